// Complex GEMM C = alpha op(A) op(B) + beta C (column-major, op in {N, T, C})
// in FP64 (ZGEMM, BlasHandle::gemm_c64, src/device/blas.rs:24-61) and FP32
// with FP32 accumulation (the projector products of the c64 filter).
//
// Block tile 64 x 64 with 256 threads (4 x 4 complex outputs per thread),
// K tiles of 16 staged in LDS. When the output tiles alone cannot fill the
// device (the tall-skinny Gram products Y^H X with K = n_pw), K is split
// across blocks; each slice writes its partial tile and k_gemm_reduce sums
// the slices in a fixed order (deterministic).
#include <algorithm>
#include <utility>

#include "impl.hpp"
#include "linalg.hpp"

namespace chs {

constexpr int GBM = 64, GBN = 64, GBK = 16;

template <typename T, int OP>
__device__ __forceinline__ cplx<T> load_op(const cplx<T>* __restrict__ A, int lda, int r, int c) {
  if (OP == 0) return A[(size_t)r + (size_t)c * lda];
  cplx<T> v = A[(size_t)c + (size_t)r * lda];
  return OP == 2 ? conj(v) : v;
}

template <typename T, int OPA, int OPB>
__global__ void __launch_bounds__(256) k_gemm(int m, int n, int k, cplx<T> alpha,
                                              const cplx<T>* __restrict__ A, int lda,
                                              const cplx<T>* __restrict__ B, int ldb, cplx<T> beta,
                                              cplx<T>* __restrict__ C, int ldc, int klen,
                                              cplx<T>* __restrict__ part) {
  __shared__ cplx<T> As[GBK][GBM + 1];
  __shared__ cplx<T> Bs[GBK][GBN + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
  const int kb = blockIdx.z * klen, ke = min(k, kb + klen);
  T ar[4][4], ai[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) ar[i][j] = ai[i][j] = (T)0;
  for (int k0 = kb; k0 < ke; k0 += GBK) {
    for (int t = threadIdx.x; t < GBM * GBK; t += 256) {
      int r, kk;
      if (OPA == 0) { r = t % GBM; kk = t / GBM; } else { kk = t % GBK; r = t / GBK; }
      const int gr = m0 + r, gk = k0 + kk;
      As[kk][r] = (gr < m && gk < ke) ? load_op<T, OPA>(A, lda, gr, gk) : mk((T)0, (T)0);
    }
    for (int t = threadIdx.x; t < GBK * GBN; t += 256) {
      int c, kk;
      if (OPB == 0) { kk = t % GBK; c = t / GBK; } else { c = t % GBN; kk = t / GBN; }
      const int gc = n0 + c, gk = k0 + kk;
      Bs[kk][c] = (gc < n && gk < ke) ? load_op<T, OPB>(B, ldb, gk, gc) : mk((T)0, (T)0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GBK; ++kk) {
      cplx<T> a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ar[i][j] = fma(a[i].x, b[j].x, ar[i][j]);
          ar[i][j] = fma(-a[i].y, b[j].y, ar[i][j]);
          ai[i][j] = fma(a[i].x, b[j].y, ai[i][j]);
          ai[i][j] = fma(a[i].y, b[j].x, ai[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = m0 + tx + 16 * i, c = n0 + ty + 16 * j;
      if (r < m && c < n) {
        const cplx<T> acc = mk(ar[i][j], ai[i][j]);
        if (part) {
          part[(size_t)blockIdx.z * m * n + (size_t)r + (size_t)c * m] = acc;
        } else {
          cplx<T> v = alpha * acc;
          if (beta.x != (T)0 || beta.y != (T)0) v = v + beta * C[(size_t)r + (size_t)c * ldc];
          C[(size_t)r + (size_t)c * ldc] = v;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Matrix-core variant: LDS-staged block tiles (below), register prefetch of
// the next K stage during the MFMAs; the complex product is three real
// 16x16x4 MFMAs per k-step (the 3M form: P1 += ar br, P2 += ai bi,
// P3 += (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2).
// Operand layout of both v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32:
// lane l supplies
// A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]. Result layout: lane l
// holds column l % 16, rows 4 (l / 16) + v (f32) or l / 16 + 4 v (f64).
// ---------------------------------------------------------------------------
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = __attribute__((ext_vector_type(4))) double;
  __device__ __forceinline__ static acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
  using acc_t = __attribute__((ext_vector_type(4))) float;
  __device__ __forceinline__ static acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Block tiles of the matrix-core kernel; 4 wavefronts as 2 x 2, each owning
// WTM x WTN MFMA tiles of 16 x 16. Small: 64 x 64 (both precisions); large:
// 128 x 128 x 16 (complex64) or 128 x 64 x 8 (complex128). The autotuner
// picks per GEMM shape.
// WAVES: wavefronts per SIMD the register allocation must leave room for
// (the small tiles keep at least two resident, so one wave's LDS waits and
// barriers do not idle the matrix core).
template <typename T, int V> struct MfmaTile;
template <typename T> struct MfmaTile<T, 1> {
  static constexpr int BM = 64, BN = 64, BK = 16, WTM = 2, WTN = 2;
  static constexpr int WAVES = sizeof(T) == 8 ? 2 : 4;
};
template <> struct MfmaTile<float, 2> {
  static constexpr int BM = 128, BN = 128, BK = 16, WTM = 4, WTN = 4;
  static constexpr int WAVES = 1;
};
// 64 x 64 with a 32-deep K stage: half the barriers per matrix-core product
// and twice the loads in flight per stage, for the long-K products (its
// 16 staged elements per thread need the registers of one wavefront per
// SIMD fewer than the 16-deep tile to stay out of scratch)
template <typename T> struct MfmaTile<T, 3> {
  static constexpr int BM = 64, BN = 64, BK = 32, WTM = 2, WTN = 2;
  static constexpr int WAVES = sizeof(T) == 8 ? 1 : 2;
};
template <> struct MfmaTile<double, 2> {
  static constexpr int BM = 128, BN = 64, BK = 8, WTM = 4, WTN = 2;
  static constexpr int WAVES = 1;
};

// XCD-aware block order: workgroups are dealt to the eight XCDs round-robin
// by launch index, so physical block p runs on XCD p % 8. Logical block
// L = (p % 8) * (P / 8) + p / 8 gives every XCD a contiguous range of
// logical blocks, ordered (k-slice, then tiles) so that the tiles sharing a
// k-slice sit in one L2. Within a slice the tiles that share the larger
// operand are consecutive (nfast: column tile fastest, so the column tiles
// of one row tile of A run together and A streams from HBM about once;
// otherwise row tile fastest, for a B wider than A is tall).
__device__ __forceinline__ bool gemm_block(int mt, int nt, int kz, int nfast, int* bm, int* bn,
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

// CL: each thread keeps one 32-bit byte offset per staged element from a
// uniform base, set once (rows and columns past the matrix edge clamped to
// the last one: they only feed output elements that are never stored); the
// bases advance by one K stage per stage, so a full stage's loads need no
// address arithmetic or bounds tests, and only a slice's last, partial stage
// tests k against its end. Usable while every offset fits in 32 bits
// (cl_offsets_fit). Otherwise each element's load sits behind its bounds
// test. Which is faster depends on the shape (the autotuner decides).
template <typename T, int OPA, int OPB, int V, bool CL>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MfmaTile<T, V>::WAVES))) k_gemm_mfma(int m, int n, int k, cplx<T> alpha,
                                                   const cplx<T>* __restrict__ A, int lda,
                                                   const cplx<T>* __restrict__ B, int ldb,
                                                   cplx<T> beta, cplx<T>* __restrict__ C, int ldc,
                                                   int klen, int ksplit, int nfast, int tri,
                                                   cplx<T>* __restrict__ part) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using TL = MfmaTile<T, V>;
  constexpr int BM = TL::BM, BN = TL::BN, BK = TL::BK;
  constexpr int WTM = TL::WTM, WTN = TL::WTN;
  int bm, bn, bz;
  if (!gemm_block((m + BM - 1) / BM, (n + BN - 1) / BN, ksplit, nfast, &bm, &bn, &bz)) return;
  if (tri && bm * BM >= (bn + 1) * BN) return;  // Hermitian product: tiles below the diagonal
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;  // staged elements per thread
  // one padding element per row: the K-major (transposing) stores of T / C
  // operands then hit distinct banks across each phase of lanes, while the
  // MFMA operand reads stay contiguous along the row
  __shared__ cplx<T> As[BK][BM + 1];
  __shared__ cplx<T> Bs[BK][BN + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm0 = (w >> 1) * (16 * WTM), wn0 = (w & 1) * (16 * WTN);
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = bm * BM, n0 = bn * BN;
  const int kb = bz * klen, ke = min(k, kb + klen);
  // three real products per complex one (Re = P1 - P2, Im = P3 - P1 - P2
  // with P1 = sum ar br, P2 = sum ai bi, P3 = sum (ar + ai)(br + bi)).
  // CL stages an op-C operand unconjugated and folds the conjugation into
  // the products: with s_a, s_b = -1 for a conjugated operand (else 1) and
  // P3 = sum (ar + s_a ai)(br + s_b bi), Re = P1 - s_a s_b P2 and
  // Im = P3 - P1 - s_a s_b P2.
  constexpr int SA = CL && OPA == 2 ? -1 : 1, SB = CL && OPB == 2 ? -1 : 1;
  constexpr int OPA_LD = CL && OPA == 2 ? 1 : OPA, OPB_LD = CL && OPB == 2 ? 1 : OPB;
  acc_t p1[WTM][WTN], p2[WTM][WTN], p3[WTM][WTN];
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j) {
      p1[i][j] = acc_t{0, 0, 0, 0};
      p2[i][j] = acc_t{0, 0, 0, 0};
      p3[i][j] = acc_t{0, 0, 0, 0};
    }
  cplx<T> ra[NA], rb[NB];
  auto a_pos = [](int e, int& r, int& kk) {
    if (OPA == 0) { r = e % BM; kk = e / BM; } else { kk = e % BK; r = e / BK; }
  };
  auto b_pos = [](int e, int& c, int& kk) {
    if (OPB == 0) { kk = e % BK; c = e / BK; } else { c = e % BN; kk = e / BN; }
  };
  // CL: per-element byte offsets from the operands' stage bases (32-bit: the
  // host takes the bounds-tested kernel when they would not fit), the bases
  // advanced by one K stage after each stage's loads
  uint32_t oa[CL ? NA : 1], ob[CL ? NB : 1];
  const char* base_a = reinterpret_cast<const char*>(A + (OPA == 0 ? (size_t)kb * lda : kb));
  const char* base_b = reinterpret_cast<const char*>(B + (OPB == 0 ? kb : (size_t)kb * ldb));
  const size_t sa = (OPA == 0 ? (size_t)BK * lda : (size_t)BK) * sizeof(cplx<T>);
  const size_t sb = (OPB == 0 ? (size_t)BK : (size_t)BK * ldb) * sizeof(cplx<T>);
  if constexpr (CL) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, kk;
      a_pos(threadIdx.x + 256 * j, r, kk);
      const size_t gr = (size_t)min(m0 + r, m - 1);
      oa[j] = (uint32_t)((OPA == 0 ? gr + (size_t)kk * lda : kk + gr * lda) * sizeof(cplx<T>));
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int c, kk;
      b_pos(threadIdx.x + 256 * j, c, kk);
      const size_t gc = (size_t)min(n0 + c, n - 1);
      ob[j] = (uint32_t)((OPB == 0 ? kk + gc * ldb : gc + (size_t)kk * ldb) * sizeof(cplx<T>));
    }
  }
  // Bounds-tested loads of stage k0 (the non-CL kernel, and the CL kernel's
  // partial last stage of a slice: k >= ke zero, rows and columns past the
  // edge clamped as the offsets are)
  auto load_tested = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, kk;
      a_pos(threadIdx.x + 256 * j, r, kk);
      const int gr = CL ? min(m0 + r, m - 1) : m0 + r, gk = k0 + kk;
      ra[j] = (gr < m && gk < ke) ? load_op<T, OPA_LD>(A, lda, gr, gk) : mk((T)0, (T)0);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int c, kk;
      b_pos(threadIdx.x + 256 * j, c, kk);
      const int gc = CL ? min(n0 + c, n - 1) : n0 + c, gk = k0 + kk;
      rb[j] = (gc < n && gk < ke) ? load_op<T, OPB_LD>(B, ldb, gk, gc) : mk((T)0, (T)0);
    }
  };
  // CL, full stages: one plain load per element at its offset, the values
  // untouched until the LDS store of the next stage, so they are waited for
  // only after this stage's matrix-core products
  auto load_plain = [&]() {
#pragma unroll
    for (int j = 0; j < NA; ++j) ra[j] = *reinterpret_cast<const cplx<T>*>(base_a + oa[j]);
#pragma unroll
    for (int j = 0; j < NB; ++j) rb[j] = *reinterpret_cast<const cplx<T>*>(base_b + ob[j]);
    base_a += sa;
    base_b += sb;
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      int r, kk;
      a_pos(threadIdx.x + 256 * j, r, kk);
      As[kk][r] = ra[j];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      int c, kk;
      b_pos(threadIdx.x + 256 * j, c, kk);
      Bs[kk][c] = rb[j];
    }
  };
  auto products = [&]() {
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
      const int kk = 4 * s + lk;
      cplx<T> a[WTM], b[WTN];
#pragma unroll
      for (int i = 0; i < WTM; ++i) a[i] = As[kk][wm0 + 16 * i + li];
#pragma unroll
      for (int j = 0; j < WTN; ++j) b[j] = Bs[kk][wn0 + 16 * j + li];
      // independent accumulators back to back: each product class over all
      // tiles before the next
#pragma unroll
      for (int i = 0; i < WTM; ++i)
#pragma unroll
        for (int j = 0; j < WTN; ++j) p1[i][j] = M::mac(a[i].x, b[j].x, p1[i][j]);
#pragma unroll
      for (int i = 0; i < WTM; ++i)
#pragma unroll
        for (int j = 0; j < WTN; ++j) p2[i][j] = M::mac(a[i].y, b[j].y, p2[i][j]);
#pragma unroll
      for (int i = 0; i < WTM; ++i)
#pragma unroll
        for (int j = 0; j < WTN; ++j)
          p3[i][j] = M::mac(a[i].x + (T)SA * a[i].y, b[j].x + (T)SB * b[j].y, p3[i][j]);
    }
  };
  if constexpr (CL) {
    // full stages with the next one's loads in flight, then the partial one
    const int kf = kb + max(0, (ke - kb) / BK) * BK;
    if (kb < kf) load_plain();
    for (int k0 = kb; k0 < kf; k0 += BK) {
      store();
      __syncthreads();
      if (k0 + BK < kf) load_plain();
      products();
      __syncthreads();
    }
    if (kf < ke) {
      load_tested(kf);
      store();
      __syncthreads();
      products();
    }
  } else {
    if (kb < ke) load_tested(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
      store();
      __syncthreads();
      if (k0 + BK < ke) load_tested(k0 + BK);  // next stage in flight during the MFMAs
      products();
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < WTM; ++i)
#pragma unroll
    for (int j = 0; j < WTN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = m0 + wm0 + 16 * i + M::row(lane, v);
        const int c = n0 + wn0 + 16 * j + li;
        if (r < m && c < n) {
          const T r1 = (T)p1[i][j][v], r2 = (T)p2[i][j][v];
          const T r2s = SA * SB < 0 ? -r2 : r2;
          const cplx<T> acc = mk(r1 - r2s, (T)p3[i][j][v] - r1 - r2s);
          if (part) {
            part[(size_t)bz * m * n + (size_t)r + (size_t)c * m] = acc;
          } else {
            cplx<T> val = alpha * acc;
            if (beta.x != (T)0 || beta.y != (T)0) val = val + beta * C[(size_t)r + (size_t)c * ldc];
            C[(size_t)r + (size_t)c * ldc] = val;
          }
        }
      }
}

// ---------------------------------------------------------------------------
// n = 1 (matrix-vector) in the memory-bound form: y = alpha op(A) x + beta y.
// Both forms cut K into slices (blockIdx.y) so that a few thousand
// wavefronts stream A, with four independent loads in flight per lane;
// with more than one slice each writes a partial vector and k_gemv_fin adds
// them in slice order (deterministic).
// op N: one thread per row, A read down its columns (coalesced across rows);
// op T / C: one wavefront per output element, a reduction along the column.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cmac(cplx<T> a, cplx<T> b, T& re, T& im) {
  re = fma(a.x, b.x, re);
  re = fma(-a.y, b.y, re);
  im = fma(a.x, b.y, im);
  im = fma(a.y, b.x, im);
}

template <typename T>
__device__ __forceinline__ void gemv_out(cplx<T> acc, cplx<T> alpha, cplx<T> beta,
                                         cplx<T>* __restrict__ y, cplx<T>* __restrict__ part,
                                         int m, int i) {
  if (part) {
    part[(size_t)blockIdx.y * m + i] = acc;
  } else {
    cplx<T> v = alpha * acc;
    if (beta.x != (T)0 || beta.y != (T)0) v = v + beta * y[i];
    y[i] = v;
  }
}

// TA: the matrix's element type (complex64 under complex128 vectors and
// accumulation is the mixed form: half the matrix bytes).
template <typename T, typename TA = T>
__device__ __forceinline__ cplx<T> widen_el(cplx<TA> a) {
  return mk((T)a.x, (T)a.y);
}

template <typename T, typename TA = T>
__global__ void __launch_bounds__(256) k_gemv_n(int m, int k, int klen, cplx<T> alpha,
                                                const cplx<TA>* __restrict__ A, int lda,
                                                const cplx<T>* __restrict__ x, cplx<T> beta,
                                                cplx<T>* __restrict__ y,
                                                cplx<T>* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int kb = blockIdx.y * klen, ke = min(k, kb + klen);
  T re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
  int kk = kb;
  for (; kk + 4 <= ke; kk += 4) {
    cplx<T> a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = widen_el<T>(A[(size_t)i + (size_t)(kk + u) * lda]);
#pragma unroll
    for (int u = 0; u < 4; ++u) cmac(a[u], x[kk + u], re[u], im[u]);
  }
  for (; kk < ke; ++kk) cmac(widen_el<T>(A[(size_t)i + (size_t)kk * lda]), x[kk], re[0], im[0]);
  gemv_out(mk((re[0] + re[1]) + (re[2] + re[3]), (im[0] + im[1]) + (im[2] + im[3])), alpha, beta,
           y, part, m, i);
}

template <typename T, int OPA, typename TA = T>
__global__ void __launch_bounds__(256) k_gemv_t(int m, int k, int klen, cplx<T> alpha,
                                                const cplx<TA>* __restrict__ A, int lda,
                                                const cplx<T>* __restrict__ x, cplx<T> beta,
                                                cplx<T>* __restrict__ y,
                                                cplx<T>* __restrict__ part) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= m) return;
  const int kb = blockIdx.y * klen, ke = min(k, kb + klen);
  const cplx<TA>* col = A + (size_t)i * lda;
  T re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
  int kk = kb + lane;
  for (; kk + 192 < ke; kk += 256) {
    cplx<T> a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = widen_el<T>(col[kk + 64 * u]);
      b[u] = x[kk + 64 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (OPA == 2) a[u].y = -a[u].y;
      cmac(a[u], b[u], re[u], im[u]);
    }
  }
  for (; kk < ke; kk += 64) {
    cplx<T> a = widen_el<T>(col[kk]);
    if (OPA == 2) a.y = -a.y;
    cmac(a, x[kk], re[0], im[0]);
  }
  T sr = (re[0] + re[1]) + (re[2] + re[3]), si = (im[0] + im[1]) + (im[2] + im[3]);
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_down(sr, off);
    si += __shfl_down(si, off);
  }
  if (lane == 0) gemv_out(mk(sr, si), alpha, beta, y, part, m, i);
}

template <typename T>
__global__ void k_gemv_fin(int m, int ns, const cplx<T>* __restrict__ part, cplx<T> alpha,
                           cplx<T> beta, cplx<T>* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  cplx<T> s = part[i];
  for (int z = 1; z < ns; ++z) s = s + part[(size_t)z * m + i];
  cplx<T> v = alpha * s;
  if (beta.x != (T)0 || beta.y != (T)0) v = v + beta * y[i];
  y[i] = v;
}

// K slices of a matrix-vector product with `waves` wavefronts per slice:
// about 4096 wavefronts in all, slices of at least `min_len` elements.
static void gemv_plan(int k, int waves, int min_len, int* ns, int* klen) {
  int s = std::max(1, std::min((4096 + waves - 1) / waves, k / min_len));
  int kl = (k + s - 1) / s;
  *klen = std::max(1, kl);
  *ns = (k + *klen - 1) / *klen;
}

template <typename T, typename TA = T>
static int gemv_launch(chs_context* ctx, int oa, int m, int k, cplx<T> alpha, const cplx<TA>* A,
                       int lda, const cplx<T>* x, cplx<T> beta, cplx<T>* y) {
  int ns, klen;
  if (oa == 0)
    gemv_plan(k, (int)ceil_div(m, 64), 16, &ns, &klen);
  else
    gemv_plan(k, m, 512, &ns, &klen);
  cplx<T>* part = nullptr;
  if (ns > 1) {
    CHS_TRY(ensure(ctx, ctx->impl->gemv_ws, (size_t)ns * m * sizeof(cplx<T>)));
    part = (cplx<T>*)ctx->impl->gemv_ws.p;
  }
  if (oa == 0)
    hipLaunchKernelGGL((k_gemv_n<T, TA>), dim3(ceil_div(m, 256), ns), dim3(256), 0, ctx->stream, m,
                       k, klen, alpha, A, lda, x, beta, y, part);
  else if (oa == 1)
    hipLaunchKernelGGL((k_gemv_t<T, 1, TA>), dim3(ceil_div(m, 4), ns), dim3(256), 0, ctx->stream, m,
                       k, klen, alpha, A, lda, x, beta, y, part);
  else
    hipLaunchKernelGGL((k_gemv_t<T, 2, TA>), dim3(ceil_div(m, 4), ns), dim3(256), 0, ctx->stream, m,
                       k, klen, alpha, A, lda, x, beta, y, part);
  CHS_LAUNCH_CHECK(ctx);
  if (ns > 1) {
    hipLaunchKernelGGL((k_gemv_fin<T>), dim3(ceil_div(m, 256)), dim3(256), 0, ctx->stream, m, ns,
                       (const cplx<T>*)part, alpha, beta, y);
    CHS_LAUNCH_CHECK(ctx);
  }
  return CHS_OK;
}

template <typename T>
__global__ void k_gemm_reduce(int m, int n, int nz, const cplx<T>* __restrict__ part,
                              cplx<T> alpha, cplx<T> beta, cplx<T>* __restrict__ C, int ldc) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)m * n) return;
  cplx<T> s = mk((T)0, (T)0);
  for (int z = 0; z < nz; ++z) s = s + part[(size_t)z * m * n + idx];
  const int r = (int)(idx % m), c = (int)(idx / m);
  cplx<T> v = alpha * s;
  if (beta.x != (T)0 || beta.y != (T)0) v = v + beta * C[(size_t)r + (size_t)c * ldc];
  C[(size_t)r + (size_t)c * ldc] = v;
}

int gemm_opcode(char op) {
  switch (op) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: return -1;
  }
}

// Split-K plan of a tile shape: when the output tiles alone cannot fill the
// device, K is cut into slices of whole K tiles until about `target`
// workgroups are in flight (more resident waves per CU hide the operand
// latency of the long-K products better; the autotuner picks the target).
// (8192 and 16384 were measured too: never faster on the 216-atom shapes)
static constexpr int kNumSplitTargets = 4;
static constexpr int kSplitTargets[kNumSplitTargets] = {512, 1024, 2048, 4096};

template <typename T>
static void tile_dims(int variant, int* bm, int* bn, int* bk) {
  *bm = GBM; *bn = GBN; *bk = GBK;
  if (variant == 1 || variant == 3) {
    *bm = MfmaTile<T, 1>::BM; *bn = MfmaTile<T, 1>::BN; *bk = MfmaTile<T, 1>::BK;
  } else if (variant == 5) {
    *bm = MfmaTile<T, 3>::BM; *bn = MfmaTile<T, 3>::BN; *bk = MfmaTile<T, 3>::BK;
  } else if (variant == 2 || variant == 4) {
    *bm = MfmaTile<T, 2>::BM; *bn = MfmaTile<T, 2>::BN; *bk = MfmaTile<T, 2>::BK;
  }
}

// tri: only the tiles that reach the diagonal or above are computed
static void split_plan(int m, int n, int k, int bm, int bn, int bk, int target, int* ksplit,
                       int* klen, bool tri = false) {
  int tiles = (int)(ceil_div(m, bm) * ceil_div(n, bn));
  if (tri) {
    tiles = 0;
    for (int i = 0; i < (int)ceil_div(m, bm); ++i)
      for (int j = 0; j < (int)ceil_div(n, bn); ++j) tiles += i * bm < (j + 1) * bn;
  }
  int ks = 1;
  // at most `target` workgroups: whole rounds of resident workgroups, no
  // nearly empty last round
  if (tiles < target && k > 4 * bk)
    ks = std::max(1, std::min(target / tiles, (k + 255) / 256));
  int kl = (k + ks - 1) / ks;
  kl = ((kl + bk - 1) / bk) * bk;
  if (kl == 0) kl = bk;
  *klen = kl;
  *ksplit = std::max(1, (k + kl - 1) / kl);
}

// Largest byte offset of a CL kernel's operand element from its stage base
// (rows and columns clamped to the matrix, k within one stage) below 2^32.
template <typename T>
static bool cl_offsets_fit(int oa, int ob, int m, int n, int lda, int ldb, int bk) {
  const double sz = (double)sizeof(cplx<T>);
  const double a = oa == 0 ? (m - 1) + (double)(bk - 1) * lda : (bk - 1) + (double)(m - 1) * lda;
  const double b = ob == 0 ? (bk - 1) + (double)(n - 1) * ldb : (n - 1) + (double)(bk - 1) * ldb;
  return a * sz < 4294967296.0 && b * sz < 4294967296.0;
}

template <typename T>
static int gemm_launch(chs_context* ctx, int choice, int oa, int ob, int m, int n, int k,
                       cplx<T> alpha, const cplx<T>* A, int lda, const cplx<T>* B, int ldb,
                       cplx<T> beta, cplx<T>* C, int ldc, DevBuf& ws, int tri = 0) {
  int variant = choice & 15;
  const int target = kSplitTargets[(choice >> 4) % kNumSplitTargets];
  // the offset-addressed kernels (3, 4, 5) only while their offsets fit 32 bits;
  // otherwise the bounds-tested kernel of the same tile
  if (variant >= 3) {
    int bm0, bn0, bk0;
    tile_dims<T>(variant, &bm0, &bn0, &bk0);
    if (!cl_offsets_fit<T>(oa, ob, m, n, lda, ldb, bk0)) variant = variant == 4 ? 2 : 1;
  }
  int bm, bn, bk;
  tile_dims<T>(variant, &bm, &bn, &bk);
  int ksplit, klen;
  split_plan(m, n, k, bm, bn, bk, target, &ksplit, &klen, tri && variant != 0);
  cplx<T>* part = nullptr;
  if (ksplit > 1) {
    CHS_TRY(ensure(ctx, ws, (size_t)ksplit * m * n * sizeof(cplx<T>)));
    part = (cplx<T>*)ws.p;
  }
  const unsigned nblk = ceil_div(m, bm) * ceil_div(n, bn) * (unsigned)ksplit;
  const int nfast = m > n ? 1 : 0;  // tiles sharing the larger operand run together
  const dim3 g1(((nblk + 7) / 8) * 8);  // multiple of 8 for the XCD-aware order
  const dim3 g3(ceil_div(m, bm), ceil_div(n, bn), ksplit);
#define CHS_GEMM_CASE(OA, OB)                                                                    \
  if (oa == OA && ob == OB) {                                                                    \
    if (variant == 1)                                                                            \
      hipLaunchKernelGGL((k_gemm_mfma<T, OA, OB, 1, false>), g1, dim3(256), 0, ctx->stream, m,   \
                         n, k, alpha, A, lda, B, ldb, beta, C, ldc, klen, ksplit, nfast, tri, part);  \
    else if (variant == 2)                                                                       \
      hipLaunchKernelGGL((k_gemm_mfma<T, OA, OB, 2, false>), g1, dim3(256), 0, ctx->stream, m,   \
                         n, k, alpha, A, lda, B, ldb, beta, C, ldc, klen, ksplit, nfast, tri, part);  \
    else if (variant == 3)                                                                       \
      hipLaunchKernelGGL((k_gemm_mfma<T, OA, OB, 1, true>), g1, dim3(256), 0, ctx->stream, m,    \
                         n, k, alpha, A, lda, B, ldb, beta, C, ldc, klen, ksplit, nfast, tri, part);  \
    else if (variant == 4)                                                                       \
      hipLaunchKernelGGL((k_gemm_mfma<T, OA, OB, 2, true>), g1, dim3(256), 0, ctx->stream, m,    \
                         n, k, alpha, A, lda, B, ldb, beta, C, ldc, klen, ksplit, nfast, tri, part);  \
    else if (variant == 5)                                                                       \
      hipLaunchKernelGGL((k_gemm_mfma<T, OA, OB, 3, true>), g1, dim3(256), 0, ctx->stream, m,    \
                         n, k, alpha, A, lda, B, ldb, beta, C, ldc, klen, ksplit, nfast, tri, part);  \
    else                                                                                         \
      hipLaunchKernelGGL((k_gemm<T, OA, OB>), g3, dim3(256), 0, ctx->stream, m, n, k, alpha, A,  \
                         lda, B, ldb, beta, C, ldc, klen, part);                                 \
  }
  CHS_GEMM_CASE(0, 0) CHS_GEMM_CASE(0, 1) CHS_GEMM_CASE(0, 2)
  CHS_GEMM_CASE(1, 0) CHS_GEMM_CASE(1, 1) CHS_GEMM_CASE(1, 2)
  CHS_GEMM_CASE(2, 0) CHS_GEMM_CASE(2, 1) CHS_GEMM_CASE(2, 2)
#undef CHS_GEMM_CASE
  CHS_LAUNCH_CHECK(ctx);
  if (ksplit > 1) {
    hipLaunchKernelGGL((k_gemm_reduce<T>), dim3(ceil_div((size_t)m * n, 256)), dim3(256), 0,
                       ctx->stream, m, n, ksplit, part, alpha, beta, C, ldc);
    CHS_LAUNCH_CHECK(ctx);
  }
  return CHS_OK;
}

// Kernel choice for a GEMM shape: variant (low 4 bits: 0 vector-FMA 64 x 64,
// 1 MFMA 64 x 64, 2 MFMA 128-wide, 3 and 4 the same with clamped loads, 5
// MFMA 64 x 64 x 32 with clamped loads) and split-K target (bits 4-5, index
// into kSplitTargets). Chosen once per (precision, ops, m, n, k) by timing every
// distinct launch into a scratch output (HIP events, context stream);
// outside a stream capture only, otherwise the MFMA 64 x 64 default.
template <typename T>
static int gemm_variant(chs_context* ctx, int oa, int ob, int m, int n, int k, const cplx<T>* A,
                        int lda, const cplx<T>* B, int ldb, int* choice, int tri = 0) {
  auto* im = ctx->impl;
  if (!im->use_mfma) {
    *choice = 0;
    return CHS_OK;
  }
  if (im->gemm_choice >= 0) {
    *choice = im->gemm_choice;
    return CHS_OK;
  }
  const GemmKey key{(int)sizeof(T), oa, ob + 16 * tri, m, n, k};
  auto it = im->gemm_tune.find(key);
  if (it != im->gemm_tune.end()) {
    *choice = it->second;
    return CHS_OK;
  }
  *choice = 1;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CHS_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone || !im->gemm_autotune) return CHS_OK;
  CHS_TRY(ensure(ctx, im->gemm_tune_c, (size_t)m * n * sizeof(cplx<T>)));
  cplx<T>* Ct = (cplx<T>*)im->gemm_tune_c.p;
  const cplx<T> one = mk((T)1, (T)0), zero = mk((T)0, (T)0);
  double best = 1e300;
  // the vector-FMA kernel never wins a large product: not timed there
  const bool big = (double)m * n * k > 1e9;
  for (int v = big ? 1 : 0; v < 6; ++v) {
    int prev_split = -1;
    for (int ti = 0; ti < kNumSplitTargets; ++ti) {
      int bm, bn, bk, ks, kl;
      tile_dims<T>(v, &bm, &bn, &bk);
      split_plan(m, n, k, bm, bn, bk, kSplitTargets[ti], &ks, &kl, tri && v != 0);
      if (ks == prev_split) continue;  // same launch as a smaller target
      prev_split = ks;
      const int c = v | (ti << 4);
      CHS_TRY(gemm_launch<T>(ctx, c, oa, ob, m, n, k, one, A, lda, B, ldb, zero, Ct, m,
                             im->gemm_tune_ws, tri));  // warm-up, workspace growth
      // best of three launches: candidates a few percent apart must not be
      // ranked by one launch's noise
      float t = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        CHS_HIP(ctx, hipEventRecord(im->ev0, ctx->stream));
        CHS_TRY(gemm_launch<T>(ctx, c, oa, ob, m, n, k, one, A, lda, B, ldb, zero, Ct, m,
                               im->gemm_tune_ws, tri));
        CHS_HIP(ctx, hipEventRecord(im->ev1, ctx->stream));
        CHS_HIP(ctx, hipEventSynchronize(im->ev1));
        float tr = 0;
        CHS_HIP(ctx, hipEventElapsedTime(&tr, im->ev0, im->ev1));
        t = std::min(t, tr);
      }
      if (t < best) {
        best = t;
        *choice = c;
      }
    }
  }
  im->gemm_tune[key] = *choice;
  tune_cache_store(ctx, "gemm " + std::to_string(key.tsize) + " " + std::to_string(key.oa) + " " +
                            std::to_string(key.ob) + " " + std::to_string(m) + " " +
                            std::to_string(n) + " " + std::to_string(k) + " " +
                            std::to_string(*choice));
  return CHS_OK;
}

template <typename T>
static int gemm_t(chs_context* ctx, char opa, char opb, int m, int n, int k, cplx<T> alpha,
                  const cplx<T>* A, int lda, const cplx<T>* B, int ldb, cplx<T> beta, cplx<T>* C,
                  int ldc) {
  const int oa = gemm_opcode(opa), ob = gemm_opcode(opb);
  if (oa < 0 || ob < 0 || m < 0 || n < 0 || k < 0) return CHS_ERR_INVALID_ARG;
  if (m == 0 || n == 0) return CHS_OK;
  if (n == 1 && ob == 0 && k > 0)  // matrix-vector
    return gemv_launch<T>(ctx, oa, m, k, alpha, A, lda, B, beta, C);
  int choice = 1;
  CHS_TRY(gemm_variant<T>(ctx, oa, ob, m, n, k, A, lda, B, ldb, &choice));
  return gemm_launch<T>(ctx, choice, oa, ob, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                        ctx->impl->gemm_ws);
}

int zgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, z64 alpha, const z64* A,
          int lda, const z64* B, int ldb, z64 beta, z64* C, int ldc) {
  return gemm_t<double>(ctx, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// C[i][j] = conj(C[j][i]) for i > j
__global__ void k_mirror_lower(z64* __restrict__ C, int n, int ldc) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i < n && j < i) C[(size_t)i + (size_t)j * ldc] = conj(C[(size_t)j + (size_t)i * ldc]);
}

int zgemm_herm(chs_context* ctx, int n, int k, const z64* A, int lda, const z64* B, int ldb, z64* C,
               int ldc) {
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  if (n < 128 || k <= 0 || !ctx->impl->use_mfma)
    return gemm_t<double>(ctx, 'C', 'N', n, n, k, one, A, lda, B, ldb, zero, C, ldc);
  int choice = 1;
  CHS_TRY(gemm_variant<double>(ctx, 2, 0, n, n, k, A, lda, B, ldb, &choice, 1));
  CHS_TRY(gemm_launch<double>(ctx, choice, 2, 0, n, n, k, one, A, lda, B, ldb, zero, C, ldc,
                              ctx->impl->gemm_ws, 1));
  hipLaunchKernelGGL(k_mirror_lower, dim3(ceil_div(n, 16), ceil_div(n, 16)), dim3(256), 0,
                     ctx->stream, C, n, ldc);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int cgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, c32 alpha, const c32* A,
          int lda, const c32* B, int ldb, c32 beta, c32* C, int ldc) {
  return gemm_t<float>(ctx, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int zgemv_c32(chs_context* ctx, char opa, int m, int k, z64 alpha, const c32* A, int lda,
              const z64* x, z64 beta, z64* y) {
  const int oa = gemm_opcode(opa);
  if (oa < 0 || m < 0 || k < 1) return CHS_ERR_INVALID_ARG;
  if (m == 0) return CHS_OK;
  return gemv_launch<double, float>(ctx, oa, m, k, alpha, A, lda, x, beta, y);
}

}  // namespace chs

using namespace chs;

extern "C" int chs_zgemm(chs_context* ctx, char opa, char opb, int m, int n, int k,
                         const double alpha[2], const double* d_a, int lda, const double* d_b,
                         int ldb, const double beta[2], double* d_c, int ldc) {
  if (!ctx || !alpha || !beta || (m > 0 && n > 0 && (!d_c || (k > 0 && (!d_a || !d_b)))))
    return CHS_ERR_INVALID_ARG;
  const int oa = gemm_opcode(opa), ob = gemm_opcode(opb);
  if (oa < 0 || ob < 0 || m < 0 || n < 0 || k < 0) return CHS_ERR_INVALID_ARG;
  if (ldc < std::max(1, m) || lda < std::max(1, oa == 0 ? m : k) ||
      ldb < std::max(1, ob == 0 ? k : n))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "zgemm: leading dimension too small");
  return zgemm(ctx, opa, opb, m, n, k, mk(alpha[0], alpha[1]), (const z64*)d_a, lda,
               (const z64*)d_b, ldb, mk(beta[0], beta[1]), (z64*)d_c, ldc);
}

// Synthetic operands for timing (values in [-0.005, 0.005)).
__global__ void k_gemm_fill(float* __restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    a[i] = ((float)(h & 0xFFFF) / 65536.0f - 0.5f) * 1e-2f;
  }
}

extern "C" int chs_cgemm(chs_context* ctx, char opa, char opb, int m, int n, int k,
                         const float alpha[2], const float* d_a, int lda, const float* d_b, int ldb,
                         const float beta[2], float* d_c, int ldc) {
  if (!ctx || !alpha || !beta || (m > 0 && n > 0 && (!d_c || (k > 0 && (!d_a || !d_b)))))
    return CHS_ERR_INVALID_ARG;
  const int oa = gemm_opcode(opa), ob = gemm_opcode(opb);
  if (oa < 0 || ob < 0 || m < 0 || n < 0 || k < 0) return CHS_ERR_INVALID_ARG;
  if (ldc < std::max(1, m) || lda < std::max(1, oa == 0 ? m : k) ||
      ldb < std::max(1, ob == 0 ? k : n))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "cgemm: leading dimension too small");
  return cgemm(ctx, opa, opb, m, n, k, mk(alpha[0], alpha[1]), (const c32*)d_a, lda,
               (const c32*)d_b, ldb, mk(beta[0], beta[1]), (c32*)d_c, ldc);
}

// Mean duration (HIP events on the context stream) of `reps` complex GEMMs
// C = A B (op(A) m x k, op(B) k x n) on synthetic operands.
extern "C" int chs_profile_gemm(chs_context* ctx, int precision, char opa, char opb, int m, int n,
                                int k, int reps, double* ms) {
  if (!ctx || !ms || m <= 0 || n <= 0 || k <= 0 || reps <= 0 ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  const int oa = gemm_opcode(opa), ob = gemm_opcode(opb);
  if (oa < 0 || ob < 0) return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  const size_t cs = precision == CHS_C128 ? 16 : 8;
  const size_t na = (size_t)m * k, nb = (size_t)k * n, nc = (size_t)m * n;
  CHS_TRY(ensure(ctx, im->ws1, na * cs));
  CHS_TRY(ensure(ctx, im->ws2, nb * cs));
  CHS_TRY(ensure(ctx, im->ws3, nc * cs));
  for (auto pr : {std::make_pair(im->ws1.p, na), std::make_pair(im->ws2.p, nb)}) {
    const size_t nf = pr.second * cs / sizeof(float);
    hipLaunchKernelGGL(k_gemm_fill, dim3(ceil_div(nf, 256)), dim3(256), 0, ctx->stream,
                       (float*)pr.first, nf);
  }
  CHS_LAUNCH_CHECK(ctx);
  const int lda = oa == 0 ? m : k, ldb = ob == 0 ? k : n;
  auto run = [&]() -> int {
    if (precision == CHS_C128)
      return zgemm(ctx, opa, opb, m, n, k, mk(1.0, 0.0), (const z64*)im->ws1.p, lda,
                   (const z64*)im->ws2.p, ldb, mk(0.0, 0.0), (z64*)im->ws3.p, m);
    return cgemm(ctx, opa, opb, m, n, k, mk(1.0f, 0.0f), (const c32*)im->ws1.p, lda,
                 (const c32*)im->ws2.p, ldb, mk(0.0f, 0.0f), (c32*)im->ws3.p, m);
  };
  CHS_TRY(run());  // warm-up (workspace growth happens here)
  CHS_HIP(ctx, hipEventRecord(im->ev0, ctx->stream));
  for (int r = 0; r < reps; ++r) CHS_TRY(run());
  CHS_HIP(ctx, hipEventRecord(im->ev1, ctx->stream));
  CHS_HIP(ctx, hipEventSynchronize(im->ev1));
  float t = 0;
  CHS_HIP(ctx, hipEventElapsedTime(&t, im->ev0, im->ev1));
  *ms = t / reps;
  return CHS_OK;
}
