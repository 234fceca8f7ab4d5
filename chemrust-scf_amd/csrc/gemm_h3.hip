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
// per call into the same planes with one exponent per column, so a column
// far smaller than the others keeps its own 22 bits.
//
// Used, when CHS_GEMM_FP16X3=1, by the complex64 Chebyshev filter for its two
// n_pw-long projector products (G = beta^H O and O + beta C,
// chebyshev.rs:416-500 as restated in coefficient space); chs_cgemm keeps
// the plain FP32 matrix-core form.
#include <algorithm>
#include <cmath>
#include <vector>

#include "impl.hpp"
#include "linalg.hpp"

namespace chs {

constexpr int kHBM = 64, kHBN = 64, kHBK = 32;

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

// Six planes [n_pad][ld] of B (c64, k x n, column stride ldb) scaled by
// 2^eb[column]; padding written as zeros.
__global__ void k_h3_split_b(const c32* __restrict__ B, int k, int n, int ldb,
                             const int* __restrict__ eb, int ld, _Float16* __restrict__ planes,
                             size_t plane) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  const int c = (int)(i / ld), kk = (int)(i % ld);
  float re = 0.f, im = 0.f;
  if (c < n && kk < k) {
    const c32 v = B[(size_t)kk + (size_t)c * ldb];
    const float sc = ldexpf(1.f, eb[c]);
    re = v.x * sc;
    im = v.y * sc;
  }
  const float x[3] = {re, im, re + im};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    _Float16 hi, lo;
    h3_split(x[q], hi, lo);
    planes[(2 * q) * plane + i] = hi;
    planes[(2 * q + 1) * plane + i] = lo;
  }
}

// Register-direct form: both operands as split planes ([rows][k], row
// pitches multiples of 32 halfs, rows padded to 64, zero padding), so every
// wavefront loads its MFMA fragments straight from memory (16 B per lane per
// plane and tile) and no LDS or barrier is needed; the four wavefronts of a
// workgroup each own a 32 x 32 quarter of the 64 x 64 tile. Occupancy (three
// wavefronts per SIMD) hides the operand latency.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_gemm_h3r(int m, int n, int k, c32 alpha, const _Float16* __restrict__ Ah, int lda_h,
           size_t aplane, int ea, const _Float16* __restrict__ Bh, int ldb_h, size_t bplane,
           const int* __restrict__ eb, c32 beta, c32* __restrict__ C, int ldc, int klen,
           int ksplit, int nfast, c32* __restrict__ part) {
  int bm, bn, bz;
  if (!h3_block((m + kHBM - 1) / kHBM, (n + kHBN - 1) / kHBN, ksplit, nfast, &bm, &bn, &bz))
    return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm0 = bm * kHBM + (w >> 1) * 32, wn0 = bn * kHBN + (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int kb = bz * klen, ke = min(k, kb + klen);
  const _Float16* arow[2];
  const _Float16* brow[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    arow[i] = Ah + (size_t)(wm0 + 16 * i + li) * lda_h + 8 * lk;
    brow[i] = Bh + (size_t)(wn0 + 16 * i + li) * ldb_h + 8 * lk;
  }
  f4 acc[3][2][2];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[q][i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kHBK) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {  // P1 (planes 0/1), P2 (2/3), P3 (4/5)
      h8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *reinterpret_cast<const h8*>(arow[i] + (size_t)(2 * q) * aplane + k0);
        al[i] = *reinterpret_cast<const h8*>(arow[i] + (size_t)(2 * q + 1) * aplane + k0);
        bh[i] = *reinterpret_cast<const h8*>(brow[i] + (size_t)(2 * q) * bplane + k0);
        bl[i] = *reinterpret_cast<const h8*>(brow[i] + (size_t)(2 * q + 1) * bplane + k0);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[q][i][j], 0, 0, 0);
          acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[q][i][j], 0, 0, 0);
          acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[q][i][j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = wn0 + 16 * j + li;
      if (c >= n) continue;
      const float unscale = ldexpf(1.f, -(ea + eb[c]));
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm0 + 16 * i + 4 * lk + v;
        if (r >= m) continue;
        const float r1 = acc[0][i][j][v], r2 = acc[1][i][j][v];
        const c32 a = mk((r1 - r2) * unscale, (acc[2][i][j][v] - r1 - r2) * unscale);
        if (part) {
          part[(size_t)bz * m * n + (size_t)r + (size_t)c * m] = a;
        } else {
          c32 val = alpha * a;
          if (beta.x != 0.f || beta.y != 0.f) val = val + beta * C[(size_t)r + (size_t)c * ldc];
          C[(size_t)r + (size_t)c * ldc] = val;
        }
      }
    }
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
// c64 k x n (column stride ldb) with per-column exponents d_eb (device): B
// is split into planes in a context workspace, then the register-direct
// kernel runs (split K over up to 2048 workgroups when the tiles alone are
// fewer).
int h3_gemm(chs_context* ctx, const H3Planes& A, int n, c32 alpha, const c32* B, int ldb,
            const int* d_eb, c32 beta, c32* C, int ldc) {
  const int m = A.m, k = A.k;
  if (m <= 0 || n <= 0) return CHS_OK;
  auto* im = ctx->impl;
  const int n_pad = (n + kHBN - 1) / kHBN * kHBN;
  const size_t bplane = (size_t)n_pad * A.ld;
  CHS_TRY(ensure(ctx, im->h3_bws, 6 * bplane * sizeof(_Float16)));
  hipLaunchKernelGGL(k_h3_split_b, dim3(ceil_div(bplane, 256)), dim3(256), 0, ctx->stream, B, k, n,
                     ldb, d_eb, A.ld, (_Float16*)im->h3_bws.p, bplane);
  CHS_LAUNCH_CHECK(ctx);
  const int tiles = ((m + kHBM - 1) / kHBM) * (n_pad / kHBN);
  int ks = 1;
  constexpr int kTarget = 2048;
  if (tiles < kTarget && k > 4 * kHBK) ks = std::max(1, std::min(kTarget / tiles, k / 1024));
  int klen = (k + ks - 1) / ks;
  klen = (klen + kHBK - 1) / kHBK * kHBK;
  ks = (k + klen - 1) / klen;
  c32* part = nullptr;
  if (ks > 1) {
    CHS_TRY(ensure(ctx, im->gemm_ws, (size_t)ks * m * n * sizeof(c32)));
    part = (c32*)im->gemm_ws.p;
  }
  const unsigned nblk = (unsigned)tiles * ks;
  hipLaunchKernelGGL(k_gemm_h3r, dim3((nblk + 7) / 8 * 8), dim3(256), 0, ctx->stream, m, n, k,
                     alpha, (const _Float16*)A.buf.p, A.ld, A.plane, A.ea,
                     (const _Float16*)im->h3_bws.p, A.ld, bplane, d_eb, beta, C, ldc, klen, ks,
                     m > n ? 1 : 0, part);
  CHS_LAUNCH_CHECK(ctx);
  if (ks > 1) {
    hipLaunchKernelGGL(k_h3_reduce, dim3(ceil_div((size_t)m * n, 256)), dim3(256), 0, ctx->stream,
                       m, n, ks, (const c32*)part, alpha, beta, C, ldc);
    CHS_LAUNCH_CHECK(ctx);
  }
  return CHS_OK;
}

}  // namespace chs
