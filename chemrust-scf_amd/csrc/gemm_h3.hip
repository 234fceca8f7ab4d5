// Complex GEMM in FP32 precision on the fp16 matrix cores ("fp16x3"):
//   C = alpha op(A) B + beta C,  op(A) m x k, B k x n (K-contiguous columns).
//
// Every real operand value x is scaled by a power of two and split into
// hi = fp16(x), lo = fp16(x - hi): hi + lo carry 22 significant bits. A real
// product uses hi*hi + hi*lo + lo*hi (the dropped lo*lo and the split
// residuals are ~2^-22 of |a||b|), accumulated in FP32 by
// v_mfma_f32_16x16x32_f16. The complex product is the 3M form:
// P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi); re = P1 - P2,
// im = P3 - P1 - P2. So one complex k-step costs nine fp16 MFMAs, each
// sixteen times the FP32 matrix rate of MI355X.
//
// A is the projector block of the bound Hamiltonian (fixed per binding): it
// is split once into six fp16 planes [m][k] (ar, ai, ar + ai, each hi and lo)
// with one exponent. B (the filter iterate or the coefficient block) is split
// on the fly in the load stage with one exponent per column, so a column far
// smaller than the others keeps its own 22 bits.
//
// Used, when CHS_GEMM_FP16X3=1, by the complex64 Chebyshev filter for its two
// n_pw-long projector products (G = beta^H O and O + beta C,
// chebyshev.rs:416-500 as restated in coefficient space). It is opt-in:
// with two resident wavefronts per SIMD and 576 matrix-core cycles per K
// stage this kernel cannot hide the operand latency (the matrix cores are
// busy ~8% of its run) and is slower than the FP32 form (3.5 vs 2.3 ms for
// 864 x 512 x 80,725); chs_cgemm keeps the plain FP32 matrix-core form.
#include <algorithm>
#include <cmath>
#include <vector>

#include "impl.hpp"
#include "linalg.hpp"

namespace chs {

constexpr int kHBM = 64, kHBN = 64, kHBK = 32;
constexpr int kHP = 40;  // LDS row pitch in halfs (80 B: the 16-B operand reads of 8 lanes
                         // land on disjoint bank quads)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Exponent e with max * 2^e < 2^15 (the sum plane ar + ai stays below the
// fp16 maximum 65504), clamped to a range that keeps 2^e finite in FP32.
__host__ __device__ inline int h3_exponent(double mx) {
  if (!(mx > 0.0)) return 0;
  int x;
  (void)frexp(mx, &x);  // mx < 2^x
  const int e = 15 - x;
  return e < -100 ? -100 : (e > 100 ? 100 : e);
}

__device__ __forceinline__ void h3_split(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// Six planes of op(A) for A c128 (P x n_pw projector-major, element (p, g) at
// A[g + p * lda]). cn: op(A) = A^H as [p][g] (m = P, k = n_pw);
// otherwise op(A) = A as [g][p] (m = n_pw, k = P). Padding rows / columns
// are written as zeros.
__global__ void k_h3_split_a(const z64* __restrict__ A, int lda, int P, int n_pw, int cn, int m_pad,
                             int ld, double scale, _Float16* __restrict__ planes, size_t plane) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)m_pad * ld) return;
  const int r = (int)(i / ld), c = (int)(i % ld);
  double ar = 0.0, ai = 0.0;
  if (cn) {
    if (r < P && c < n_pw) {
      const z64 v = A[(size_t)c + (size_t)r * lda];
      ar = v.x;
      ai = -v.y;
    }
  } else if (r < n_pw && c < P) {
    const z64 v = A[(size_t)r + (size_t)c * lda];
    ar = v.x;
    ai = v.y;
  }
  const double x[3] = {ar * scale, ai * scale, (ar + ai) * scale};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const _Float16 hi = (_Float16)x[q];
    const _Float16 lo = (_Float16)(x[q] - (double)hi);
    planes[(2 * q) * plane + i] = hi;
    planes[(2 * q + 1) * plane + i] = lo;
  }
}

// Per-block maxima of |re|, |im|, |re + im|, |re - im| over A (c128, count
// n), one partial per block.
__global__ void k_h3_absmax(const z64* __restrict__ A, size_t n, double* __restrict__ part) {
  __shared__ double sh[256];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const z64 v = A[i];
    m = fmax(m, fmax(fmax(fabs(v.x), fabs(v.y)), fmax(fabs(v.x + v.y), fabs(v.x - v.y))));
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Per-column exponents of B (c64, k x n, column c at B + c * ldb): one block
// per column, max of |re|, |im|, |re + im|.
__global__ void __launch_bounds__(256) k_h3_colexp(const c32* __restrict__ B, int k, int ldb,
                                                   int* __restrict__ eb) {
  __shared__ float sh[256];
  const c32* col = B + (size_t)blockIdx.x * ldb;
  float m = 0.f;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const c32 v = col[i];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.x + v.y)));
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) eb[blockIdx.x] = h3_exponent((double)sh[0]);
}

// XCD-aware block order as gemm_block in gemm.hip: each XCD takes a
// contiguous range of logical tiles; with nfast the column tiles sharing a
// row tile of A are consecutive (A, the larger operand, then streams from
// HBM about once), otherwise the row tile is fastest.
__device__ __forceinline__ bool h3_block(int mt, int nt, int kz, int nfast, int* bm, int* bn,
                                         int* bz) {
  const int P = gridDim.x, p = blockIdx.x;
  const int L = (p & 7) * (P >> 3) + (p >> 3);
  if (L >= mt * nt * kz) return false;
  if (nfast) {
    *bn = L % nt;
    *bm = (L / nt) % mt;
  } else {
    *bm = L % mt;
    *bn = (L / mt) % nt;
  }
  *bz = L / (mt * nt);
  return true;
}

// 64 x 64 output tile per workgroup of 256 threads (four wavefronts as 2 x 2,
// each 2 x 2 MFMA tiles of 16 x 16), K stages of 32, register prefetch of the
// next stage. Operand layout of v_mfma_f32_16x16x32_f16: lane l supplies
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8; the result
// holds rows 4 (l / 16) + v, column l % 16.
__global__ void __launch_bounds__(256) k_gemm_h3(int m, int n, int k, c32 alpha,
                                                 const _Float16* __restrict__ Ah, int lda_h,
                                                 size_t plane, int ea, const c32* __restrict__ B,
                                                 int ldb, const int* __restrict__ eb, c32 beta,
                                                 c32* __restrict__ C, int ldc, int klen, int ksplit,
                                                 int nfast, c32* __restrict__ part) {
  int bm, bn, bz;
  if (!h3_block((m + kHBM - 1) / kHBM, (n + kHBN - 1) / kHBN, ksplit, nfast, &bm, &bn, &bz))
    return;
  __shared__ __attribute__((aligned(16))) _Float16 As[6][kHBM][kHP];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[6][kHBN][kHP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm0 = (w >> 1) * 32, wn0 = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = bm * kHBM, n0 = bn * kHBN;
  const int kb = bz * klen, ke = min(k, kb + klen);
  // loader coordinates: A row / B column lr, k octet lc
  const int lr = threadIdx.x >> 2, lc = (threadIdx.x & 3) * 8;
  const int bc = n0 + lr;
  const float sb = bc < n ? ldexpf(1.f, eb[bc]) : 0.f;
  const c32* bcol = B + (size_t)(bc < n ? bc : 0) * ldb;
  f4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) p1[i][j] = p2[i][j] = p3[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  uint4 ra[6];
  c32 rb[8];
  auto load = [&](int k0) {
    const _Float16* arow = Ah + (size_t)(m0 + lr) * lda_h + k0 + lc;
#pragma unroll
    for (int q = 0; q < 6; ++q) ra[q] = *reinterpret_cast<const uint4*>(arow + (size_t)q * plane);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = k0 + lc + j;
      rb[j] = (bc < n && kk < ke) ? bcol[kk] : mk(0.f, 0.f);
    }
  };
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += kHBK) {
#pragma unroll
    for (int q = 0; q < 6; ++q) *reinterpret_cast<uint4*>(&As[q][lr][lc]) = ra[q];
    {
      h8 v[6];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float re = rb[j].x * sb, im = rb[j].y * sb;
        _Float16 h, l;
        h3_split(re, h, l);
        v[0][j] = h;
        v[1][j] = l;
        h3_split(im, h, l);
        v[2][j] = h;
        v[3][j] = l;
        h3_split(re + im, h, l);
        v[4][j] = h;
        v[5][j] = l;
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) *reinterpret_cast<h8*>(&Bs[q][lr][lc]) = v[q];
    }
    __syncthreads();
    if (k0 + kHBK < ke) load(k0 + kHBK);  // next stage in flight during the MFMAs
    // one real product class: acc += hi*hi + hi*lo + lo*hi over planes (q, q + 1)
    auto product = [&](f4(&acc)[2][2], int q) {
      h8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *reinterpret_cast<const h8*>(&As[q][wm0 + 16 * i + li][8 * lk]);
        al[i] = *reinterpret_cast<const h8*>(&As[q + 1][wm0 + 16 * i + li][8 * lk]);
        bh[i] = *reinterpret_cast<const h8*>(&Bs[q][wn0 + 16 * i + li][8 * lk]);
        bl[i] = *reinterpret_cast<const h8*>(&Bs[q + 1][wn0 + 16 * i + li][8 * lk]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
    };
    product(p1, 0);  // ar br
    product(p2, 2);  // ai bi
    product(p3, 4);  // (ar + ai)(br + bi)
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = n0 + wn0 + 16 * j + li;
      if (c >= n) continue;
      const float unscale = ldexpf(1.f, -(ea + eb[c]));
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = m0 + wm0 + 16 * i + 4 * lk + v;
        if (r >= m) continue;
        const float r1 = p1[i][j][v], r2 = p2[i][j][v];
        const c32 acc = mk((r1 - r2) * unscale, (p3[i][j][v] - r1 - r2) * unscale);
        if (part) {
          part[(size_t)bz * m * n + (size_t)r + (size_t)c * m] = acc;
        } else {
          c32 val = alpha * acc;
          if (beta.x != 0.f || beta.y != 0.f) val = val + beta * C[(size_t)r + (size_t)c * ldc];
          C[(size_t)r + (size_t)c * ldc] = val;
        }
      }
    }
}

__global__ void k_h3_reduce(int m, int n, int nz, const c32* __restrict__ part, c32 alpha,
                            c32 beta, c32* __restrict__ C, int ldc) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)m * n) return;
  c32 s = part[idx];
  for (int z = 1; z < nz; ++z) s = s + part[(size_t)z * m * n + idx];
  const int r = (int)(idx % m), c = (int)(idx / m);
  c32 v = alpha * s;
  if (beta.x != 0.f || beta.y != 0.f) v = v + beta * C[(size_t)r + (size_t)c * ldc];
  C[(size_t)r + (size_t)c * ldc] = v;
}

// Split planes of op(A) for a c128 projector block (P x n_pw, lda), cn as in
// k_h3_split_a. Reads the block's maximum back once (set-up, not per step).
int h3_split_planes(chs_context* ctx, const z64* A, int lda, int P, int n_pw, bool cn,
                    H3Planes& out) {
  const int m = cn ? P : n_pw, k = cn ? n_pw : P;
  const int m_pad = (m + kHBM - 1) / kHBM * kHBM;
  const int ld = (k + kHBK - 1) / kHBK * kHBK;
  const size_t plane = (size_t)m_pad * ld;
  auto* im = ctx->impl;
  constexpr int kMaxBlocks = 1024;
  CHS_TRY(ensure(ctx, im->part, kMaxBlocks * sizeof(double)));
  const size_t cnt = (size_t)P * lda;
  hipLaunchKernelGGL(k_h3_absmax, dim3(kMaxBlocks), dim3(256), 0, ctx->stream, A, cnt,
                     (double*)im->part.p);
  CHS_LAUNCH_CHECK(ctx);
  std::vector<double> hp(kMaxBlocks);
  CHS_HIP(ctx, hipMemcpyAsync(hp.data(), im->part.p, kMaxBlocks * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += kMaxBlocks * sizeof(double);
  double mx = 0.0;
  for (double v : hp) mx = std::max(mx, v);
  const int ea = h3_exponent(mx);
  CHS_TRY(ensure(ctx, out.buf, 6 * plane * sizeof(_Float16)));
  hipLaunchKernelGGL(k_h3_split_a, dim3(ceil_div(plane, 256)), dim3(256), 0, ctx->stream, A, lda, P,
                     n_pw, cn ? 1 : 0, m_pad, ld, std::ldexp(1.0, ea), (_Float16*)out.buf.p, plane);
  CHS_LAUNCH_CHECK(ctx);
  out.m = m;
  out.k = k;
  out.m_pad = m_pad;
  out.ld = ld;
  out.plane = plane;
  out.ea = ea;
  return CHS_OK;
}

int h3_column_exponents(chs_context* ctx, const c32* B, int k, int n, int ldb, int* d_eb) {
  if (n <= 0) return CHS_OK;
  hipLaunchKernelGGL(k_h3_colexp, dim3(n), dim3(256), 0, ctx->stream, B, k, ldb, d_eb);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// C = alpha op(A) B + beta C with op(A) from its split planes (m x k), B
// c64 k x n (column stride ldb) with per-column exponents d_eb (device).
int h3_gemm(chs_context* ctx, const H3Planes& A, int n, c32 alpha, const c32* B, int ldb,
            const int* d_eb, c32 beta, c32* C, int ldc) {
  const int m = A.m, k = A.k;
  if (m <= 0 || n <= 0) return CHS_OK;
  const int tiles = ((m + kHBM - 1) / kHBM) * ((n + kHBN - 1) / kHBN);
  // split K when the tiles alone cannot fill the device: up to 2048
  // workgroups, slices of whole K stages
  int ks = 1;
  constexpr int kTarget = 2048;
  if (tiles < kTarget && k > 4 * kHBK) ks = std::max(1, std::min(kTarget / tiles, k / 1024));
  int klen = (k + ks - 1) / ks;
  klen = (klen + kHBK - 1) / kHBK * kHBK;
  ks = (k + klen - 1) / klen;
  c32* part = nullptr;
  if (ks > 1) {
    CHS_TRY(ensure(ctx, ctx->impl->gemm_ws, (size_t)ks * m * n * sizeof(c32)));
    part = (c32*)ctx->impl->gemm_ws.p;
  }
  const unsigned nblk = (unsigned)tiles * ks;
  hipLaunchKernelGGL(k_gemm_h3, dim3((nblk + 7) / 8 * 8), dim3(256), 0, ctx->stream, m, n, k, alpha,
                     (const _Float16*)A.buf.p, A.ld, A.plane, A.ea, B, ldb, d_eb, beta, C, ldc,
                     klen, ks, m > n ? 1 : 0, part);
  CHS_LAUNCH_CHECK(ctx);
  if (ks > 1) {
    hipLaunchKernelGGL(k_h3_reduce, dim3(ceil_div((size_t)m * n, 256)), dim3(256), 0, ctx->stream,
                       m, n, ks, (const c32*)part, alpha, beta, C, ldc);
    CHS_LAUNCH_CHECK(ctx);
  }
  return CHS_OK;
}

}  // namespace chs
