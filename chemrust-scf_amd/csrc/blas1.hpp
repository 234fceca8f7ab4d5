// Level-1 style device helpers: deterministic reductions and vector updates.
#pragma once
#include "common.hpp"

namespace chs {

constexpr int kRedBlocks = 256;
constexpr int kRedThreads = 256;

// Block-level sum of a double over blockDim.x == kRedThreads threads.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  __syncthreads();
  return s;  // valid in thread 0
}

// part[block] = sum over the block's grid-stride slice of Re(conj(a) b)
// (and Im into part[kRedBlocks + block] when want_im).
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_dot_part(const cplx<T>* __restrict__ a,
                                                          const cplx<T>* __restrict__ b, size_t n,
                                                          double* __restrict__ part) {
  __shared__ double sh[8];
  double re = 0, im = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const cplx<T> x = a[i], y = b[i];
    re += (double)x.x * y.x + (double)x.y * y.y;
    im += (double)x.x * y.y - (double)x.y * y.x;
  }
  double r = block_sum(re, sh);
  double s = block_sum(im, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r;
    part[gridDim.x + blockIdx.x] = s;
  }
}

// dst[0] = sum part[0..nb), dst[1] = sum part[nb..2nb) (fixed order).
static __global__ void k_sum_parts(const double* __restrict__ part, int nb,
                                   double* __restrict__ dst) {
  __shared__ double sh[8];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a += part[i];
    b += part[nb + i];
  }
  double r = block_sum(a, sh);
  double s = block_sum(b, sh);
  if (threadIdx.x == 0) {
    dst[0] = r;
    dst[1] = s;
  }
}

}  // namespace chs
