// Plane-wave sphere enumeration and the PW -> FFT index map.
#include <cmath>

#include "impl.hpp"

namespace chs {

// Reference src/scf.rs:387-401: negative Miller indices wrap by +n; the
// linear index is ix + ngx * (iy + ngy * iz).
__global__ void k_fft_indices(const int32_t* __restrict__ miller, int n_pw, int ngx, int ngy,
                              int ngz, int32_t* __restrict__ idx) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  int h = miller[3 * g], k = miller[3 * g + 1], l = miller[3 * g + 2];
  int ix = h >= 0 ? h : h + ngx;
  int iy = k >= 0 ? k : k + ngy;
  int iz = l >= 0 ? l : l + ngz;
  idx[g] = ix + ngx * (iy + ngy * iz);
}

// Sphere test for candidate c of the (2 hmax + 1)^3 cube in lexicographic
// (h, k, l) order. Every product and sum is separately rounded (no fused
// multiply-add), matching the CPU definition bit for bit.
__device__ __forceinline__ bool in_sphere(long c, int hmax, double b, double kf0, double kf1,
                                          double kf2, double ecut, int* h, int* k, int* l,
                                          double* e, double* gx, double* gy, double* gz) {
  // no contraction into fused multiply-adds anywhere in this function
#pragma clang fp contract(off)
  const int d = 2 * hmax + 1;
  *l = (int)(c % d) - hmax;
  *k = (int)((c / d) % d) - hmax;
  *h = (int)(c / ((long)d * d)) - hmax;
  *gx = b * ((double)*h + kf0);
  *gy = b * ((double)*k + kf1);
  *gz = b * ((double)*l + kf2);
  const double g2 = (*gx * *gx + *gy * *gy) + *gz * *gz;
  *e = 0.5 * g2;
  return *e <= ecut;
}

constexpr int kScanBlock = 1024;

__global__ void k_sphere_count(long n_cand, int hmax, double b, double kf0, double kf1, double kf2,
                               double ecut, int32_t* __restrict__ block_count) {
  __shared__ int red[kScanBlock / 64];
  const long c = (long)blockIdx.x * kScanBlock + threadIdx.x;
  int h, k, l;
  double e, gx, gy, gz;
  int f = (c < n_cand) && in_sphere(c, hmax, b, kf0, kf1, kf2, ecut, &h, &k, &l, &e, &gx, &gy, &gz);
  // wave reduction then block
  unsigned long long bal = __ballot(f);
  int cnt = __popcll(bal);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) s += red[w];
    block_count[blockIdx.x] = s;
  }
}

// Exclusive scan of block counts by one block (n_blocks small).
__global__ void k_scan_blocks(const int32_t* __restrict__ cnt, int n, int32_t* __restrict__ off,
                              int32_t* __restrict__ total) {
  if (threadIdx.x != 0) return;
  int s = 0;
  for (int i = 0; i < n; ++i) {
    off[i] = s;
    s += cnt[i];
  }
  *total = s;
}

__global__ void k_sphere_emit(long n_cand, int hmax, double b, double kf0, double kf1, double kf2,
                              double ecut, int n_grid, const int32_t* __restrict__ off,
                              int32_t* __restrict__ miller, int32_t* __restrict__ idx,
                              double* __restrict__ kin, int32_t* __restrict__ err) {
  __shared__ int wsum[kScanBlock / 64];
  const long c = (long)blockIdx.x * kScanBlock + threadIdx.x;
  int h = 0, k = 0, l = 0;
  double e = 0, gx, gy, gz;
  int f = (c < n_cand) && in_sphere(c, hmax, b, kf0, kf1, kf2, ecut, &h, &k, &l, &e, &gx, &gy, &gz);
  unsigned long long bal = __ballot(f);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int before = __popcll(bal & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(bal);
  __syncthreads();
  int base = off[blockIdx.x];
  for (int i = 0; i < w; ++i) base += wsum[i];
  if (f) {
    const int g = base + before;
    if (2 * abs(h) >= n_grid || 2 * abs(k) >= n_grid || 2 * abs(l) >= n_grid) atomicOr(err, 1);
    miller[3 * g] = h;
    miller[3 * g + 1] = k;
    miller[3 * g + 2] = l;
    const int ix = h >= 0 ? h : h + n_grid;
    const int iy = k >= 0 ? k : k + n_grid;
    const int iz = l >= 0 ? l : l + n_grid;
    idx[g] = ix + n_grid * (iy + n_grid * iz);
    kin[g] = e;
  }
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_pw_coords_to_fft_indices(chs_context* ctx, const int32_t* d_miller, int n_pw, int ngx,
                                 int ngy, int ngz, int32_t* d_idx) {
  if (!ctx || n_pw < 0 || (n_pw && (!d_miller || !d_idx)) || ngx <= 0 || ngy <= 0 || ngz <= 0)
    return CHS_ERR_INVALID_ARG;
  if (n_pw == 0) return CHS_OK;
  hipLaunchKernelGGL(k_fft_indices, dim3(ceil_div(n_pw, 256)), dim3(256), 0, ctx->stream, d_miller,
                     n_pw, ngx, ngy, ngz, d_idx);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int chs_gsphere(chs_context* ctx, double L, double ecut, const double kfrac[3], int n_grid,
                int max_pw, int* n_pw, int32_t* d_miller, int32_t* d_idx, double* d_kin) {
  if (!ctx || !n_pw || !kfrac || !(L > 0) || !(ecut > 0) || n_grid <= 0) return CHS_ERR_INVALID_ARG;
  const double b = (2.0 * M_PI) / L;
  const int hmax = (int)std::ceil(std::sqrt(2.0 * ecut) / b) + 1;
  const long d = 2 * hmax + 1;
  const long n_cand = d * d * d;
  const int nblk = (int)((n_cand + kScanBlock - 1) / kScanBlock);
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->small, sizeof(int32_t) * (2 * (size_t)nblk + 4)));
  int32_t* cnt = (int32_t*)im->small.p;
  int32_t* off = cnt + nblk;
  int32_t* tot = off + nblk;  // tot[0] total, tot[1] error flag
  hipLaunchKernelGGL(k_sphere_count, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, n_cand, hmax, b,
                     kfrac[0], kfrac[1], kfrac[2], ecut, cnt);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(64), 0, ctx->stream, cnt, nblk, off, tot);
  CHS_LAUNCH_CHECK(ctx);
  int32_t h_tot[2] = {0, 0};
  CHS_HIP(ctx, hipMemcpyAsync(h_tot, tot, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  ctx->d2h += sizeof(int32_t);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_pw = h_tot[0];
  if (!d_miller) return CHS_OK;
  if (!d_idx || !d_kin || max_pw < h_tot[0]) return CHS_ERR_INVALID_ARG;
  CHS_HIP(ctx, hipMemsetAsync(tot + 1, 0, sizeof(int32_t), ctx->stream));
  hipLaunchKernelGGL(k_sphere_emit, dim3(nblk), dim3(kScanBlock), 0, ctx->stream, n_cand, hmax, b,
                     kfrac[0], kfrac[1], kfrac[2], ecut, n_grid, off, d_miller, d_idx, d_kin,
                     tot + 1);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(h_tot + 1, tot + 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += sizeof(int32_t);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h_tot[1]) return set_error(ctx, CHS_ERR_INVALID_ARG, "plane-wave sphere exceeds the %d^3 grid", n_grid);
  return CHS_OK;
}

}  // extern "C"
