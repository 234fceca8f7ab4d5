// Hermitian eigensolver for the subspace problems of Rayleigh-Ritz
// (n <= 1024, FP64): Householder reduction to real tridiagonal form,
// multisection for the eigenvalues, inverse iteration for the tridiagonal
// eigenvectors, and the blocked (compact WY) back-transformation.
//
// Reference: the generalized problem is solved by cuSOLVER ZHEGVD in
// SolverHandle::zhegvd (src/device/solver.rs:47-95) and by a dense CPU
// eigensolver in diagonalize_and_rotate (src/eigensolver/
// subspace_diag.rs:229-391); both return eigenvalues ascending with
// orthonormal eigenvectors, as this solver does.
//
// Reduction, step k (trailing size m = n - k - 1, x = A[k+1:, k]):
//   H = I - tau v v^H with H^H x = beta e_1, beta real, v_0 = 1;
//   p = tau A22 v, s = v^H A22 v, w = p - (|tau|^2 s / 2) v,
//   A22 <- H^H A22 H = A22 - v w^H - w v^H.
// Then A = Q T Q^H with Q = H_0 H_1 ... H_{n-2}, and A's eigenvectors are
// Q Z for T's eigenvectors Z. Step k is one kernel (k_hh_step: the rank-2
// update of step k - 1 fused with the partial products of step k), the n - 1
// launches replayed from a captured graph.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"

namespace chs {

constexpr int kHC = 64;   // row/column chunk of the symmetric product and update tiles
constexpr int kWY = 32;   // reflectors per compact-WY block

// Block sum visible to every thread (sh: at least 9 doubles).
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) sh[8] = s;
  __syncthreads();
  const double r = sh[8];
  __syncthreads();
  return r;
}

// Reflector of x = A[k+1:, k] (every block recomputes ||x_rest||^2: x is at
// most 1023 elements and this saves a launch). Returns tau, beta and the
// scale 1 / (alpha - beta) that maps x_j (j >= 1) to v_j.
__device__ void hh_reflector(const z64* __restrict__ x, int m, double* sh, z64* tau, double* beta,
                             z64* scale) {
  double s = 0.0;
  for (int t = threadIdx.x + 1; t < m; t += blockDim.x) s += norm2(x[t]);
  const double xn2 = block_sum_all(s, sh);
  const z64 alpha = x[0];
  if (xn2 == 0.0 && alpha.y == 0.0) {
    *tau = mk(0.0, 0.0);
    *beta = alpha.x;
    *scale = mk(0.0, 0.0);
    return;
  }
  const double b = -copysign(sqrt(norm2(alpha) + xn2), alpha.x);
  *beta = b;
  *tau = mk((b - alpha.x) / b, -alpha.y / b);
  const z64 den = mk(alpha.x - b, alpha.y);
  const double dn = norm2(den);
  *scale = mk(den.x / dn, -den.y / dn);
}

// Partial p and s of step k. Block (bi, bj) multiplies rows chunk bi of A22
// with v over columns chunk bj: ppart[bj][i] = sum_j A22[i][j] v_j,
// spart[bi][bj] = sum_i conj(v_i) ppart[bj][i]. Block (0, 0) also stores
// v in column k of Vst, tau_k, e_k = beta and d_k = Re A[k][k].
__global__ void __launch_bounds__(256) k_hh_symv(const z64* __restrict__ A, int n, int k,
                                                 z64* __restrict__ Vst, z64* __restrict__ ppart,
                                                 z64* __restrict__ spart, z64* __restrict__ tau,
                                                 double* __restrict__ dd, double* __restrict__ ee) {
  __shared__ double sh[9];
  __shared__ z64 vj[kHC];
  __shared__ z64 acc_sh[4][kHC];
  const int m = n - k - 1;
  const int nch = (m + kHC - 1) / kHC;
  const int bi = blockIdx.x, bj = blockIdx.y;
  const z64* x = A + (size_t)(k + 1) + (size_t)k * n;
  z64 t_k, scale;
  double beta;
  hh_reflector(x, m, sh, &t_k, &beta, &scale);
  if (bi == 0 && bj == 0) {
    z64* vcol = Vst + (size_t)k * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const int j = r - k - 1;
      vcol[r] = j < 0 ? mk(0.0, 0.0) : (j == 0 ? mk(1.0, 0.0) : x[j] * scale);
    }
    if (threadIdx.x == 0) {
      tau[k] = t_k;
      ee[k] = beta;
      dd[k] = A[(size_t)k + (size_t)k * n].x;
    }
  }
  for (int t = threadIdx.x; t < kHC; t += blockDim.x) {
    const int j = bj * kHC + t;
    vj[t] = j < m ? (j == 0 ? mk(1.0, 0.0) : x[j] * scale) : mk(0.0, 0.0);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = bi * kHC + lane;
  z64 acc = mk(0.0, 0.0);
  if (i < m) {
    const z64* arow = A + (size_t)(k + 1 + i);
#pragma unroll 4
    for (int jj = 0; jj < kHC / 4; ++jj) {
      const int jl = w * (kHC / 4) + jj, j = bj * kHC + jl;
      if (j < m) acc = acc + arow[(size_t)(k + 1 + j) * n] * vj[jl];
    }
  }
  acc_sh[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    z64 p = acc_sh[0][lane] + acc_sh[1][lane] + acc_sh[2][lane] + acc_sh[3][lane];
    z64 sv = mk(0.0, 0.0);
    if (i < m) {
      ppart[(size_t)bj * n + i] = p;
      const z64 vi = i == 0 ? mk(1.0, 0.0) : x[i] * scale;
      sv = cmulc(vi, p);
    }
    for (int off = 32; off > 0; off >>= 1) {
      sv.x += __shfl_down(sv.x, off);
      sv.y += __shfl_down(sv.y, off);
    }
    if (lane == 0) spart[(size_t)bi * nch + bj] = sv;
  }
}

// w of step k - 1 for every index of its trailing matrix (t in [0, mp),
// global index k + t) into LDS, with v of that step: w = tau p - c v,
// p = sum of the step's partial products, c = |tau|^2 (v^H A22 v) / 2.
// The partial products are summed before the block reduction of c so that
// their loads are in flight together (blockDim 256, mp <= 1024).
__device__ void hh_prev_vw(const z64* __restrict__ Vst, const z64* __restrict__ ppart,
                           const z64* __restrict__ spart, const z64* __restrict__ tau, int n,
                           int k, z64* __restrict__ vp, z64* __restrict__ wp, double* sh) {
  const int mp = n - k;  // trailing size of step k - 1
  const int nchp = (mp + kHC - 1) / kHC;
  const z64 t_p = tau[k - 1];
  const z64* v = Vst + (size_t)(k - 1) * n + k;
  z64 preg[4], vreg[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = threadIdx.x + 256 * q;
    z64 p = mk(0.0, 0.0), vt = mk(0.0, 0.0);
    if (t < mp) {
      for (int b = 0; b < nchp; ++b) p = p + ppart[(size_t)b * n + t];
      vt = v[t];
    }
    preg[q] = t_p * p;
    vreg[q] = vt;
  }
  double s = 0.0;
  for (int t = threadIdx.x; t < nchp * nchp; t += blockDim.x) s += spart[t].x;
  const double c = 0.5 * norm2(t_p) * block_sum_all(s, sh);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = threadIdx.x + 256 * q;
    if (t < mp) {
      vp[t] = vreg[q];
      wp[t] = mk(preg[q].x - c * vreg[q].x, preg[q].y - c * vreg[q].y);
    }
  }
  __syncthreads();
}

// Step k >= 1 of the reduction fused with the rank-2 update of step k - 1.
// The update A22 -= v w^H + w v^H of step k - 1 is applied in memory to
// [k+1, n)^2 only; column k and A[k][k] of that trailing matrix are formed
// on the fly by every block (read, never written: nothing reads them after
// this step), giving x = A[k+1:, k], the reflector of step k and d_k. Each
// block (bi, bj) then multiplies its freshly updated 64 x 64 tile with v_k:
// ppart[bj][i] = sum_j A22[i][j] v_j, spart[bi][bj] = sum_i conj(v_i) p_i.
// Block (0, 0) stores v_k (column k of Vst), tau_k, e_k = beta, d_k.
__global__ void __launch_bounds__(256) k_hh_step(z64* __restrict__ A, int n, int k,
                                                 z64* __restrict__ Vst,
                                                 const z64* __restrict__ ppart_prev,
                                                 const z64* __restrict__ spart_prev,
                                                 z64* __restrict__ ppart, z64* __restrict__ spart,
                                                 z64* __restrict__ tau, double* __restrict__ dd,
                                                 double* __restrict__ ee) {
  __shared__ double sh[9];
  __shared__ z64 vp[1024], wp[1024], xk[1024];
  __shared__ z64 vj[kHC];
  __shared__ z64 acc_sh[4][kHC];
  const int m = n - k - 1;
  const int nch = (m + kHC - 1) / kHC;
  const int bi = blockIdx.x, bj = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = bi * kHC + lane;
  // Loads that depend on nothing computed here go first, so that their
  // latency overlaps the reductions below: the thread's 16 elements of its
  // tile (addresses clamped into the trailing matrix, m >= 1; nothing in
  // this launch writes them before the update) and its entries of column k.
  z64 areg[kHC / 4];
  {
    const size_t ro = (size_t)(k + 1 + min(i, m - 1));
#pragma unroll
    for (int jj = 0; jj < kHC / 4; ++jj) {
      const int j = bj * kHC + w * (kHC / 4) + jj;
      areg[jj] = A[ro + (size_t)(k + 1 + min(j, m - 1)) * n];
    }
  }
  // column k of the step-(k-1) trailing matrix below the diagonal (rows k+1..)
  const z64* colk = A + (size_t)(k + 1) + (size_t)k * n;
  z64 creg[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = threadIdx.x + 256 * q;
    creg[q] = colk[min(r, m - 1)];
  }
  hh_prev_vw(Vst, ppart_prev, spart_prev, tau, n, k, vp, wp, sh);
  const z64 cw0 = conj(wp[0]), cv0 = conj(vp[0]);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = threadIdx.x + 256 * q;
    if (r < m) xk[r] = creg[q] - vp[r + 1] * cw0 - wp[r + 1] * cv0;
  }
  __syncthreads();
  z64 t_k, scale;
  double beta;
  hh_reflector(xk, m, sh, &t_k, &beta, &scale);
  if (bi == 0 && bj == 0) {
    z64* vcol = Vst + (size_t)k * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const int j = r - k - 1;
      vcol[r] = j < 0 ? mk(0.0, 0.0) : (j == 0 ? mk(1.0, 0.0) : xk[j] * scale);
    }
    if (threadIdx.x == 0) {
      tau[k] = t_k;
      ee[k] = beta;
      const z64 vw = vp[0] * cw0;
      dd[k] = A[(size_t)k + (size_t)k * n].x - 2.0 * vw.x;
    }
  }
  for (int t = threadIdx.x; t < kHC; t += blockDim.x) {
    const int j = bj * kHC + t;
    vj[t] = j < m ? (j == 0 ? mk(1.0, 0.0) : xk[j] * scale) : mk(0.0, 0.0);
  }
  __syncthreads();
  z64 acc = mk(0.0, 0.0);
  if (i < m) {
    z64* arow = A + (size_t)(k + 1 + i);
    const z64 vr = vp[i + 1], wr = wp[i + 1];
#pragma unroll
    for (int jj = 0; jj < kHC / 4; ++jj) {
      const int jl = w * (kHC / 4) + jj, j = bj * kHC + jl;
      if (j < m) {
        const z64 u = vr * conj(wp[j + 1]) + wr * conj(vp[j + 1]);
        const z64 upd = areg[jj] - u;
        arow[(size_t)(k + 1 + j) * n] = upd;
        acc = acc + upd * vj[jl];
      }
    }
  }
  acc_sh[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    z64 p = acc_sh[0][lane] + acc_sh[1][lane] + acc_sh[2][lane] + acc_sh[3][lane];
    z64 sv = mk(0.0, 0.0);
    if (i < m) {
      ppart[(size_t)bj * n + i] = p;
      sv = cmulc(i == 0 ? mk(1.0, 0.0) : xk[i] * scale, p);
    }
    for (int off = 32; off > 0; off >>= 1) {
      sv.x += __shfl_down(sv.x, off);
      sv.y += __shfl_down(sv.y, off);
    }
    if (lane == 0) spart[(size_t)bi * nch + bj] = sv;
  }
}

// The last steps of the reduction in one workgroup (256 threads): from step
// k0 on, the trailing matrix [k0, n)^2 (at most kHT x kHT) stays in LDS
// (column-major, As[col][row] relative to k0), so each step's product,
// reflector and rank-2 update are separated by barriers instead of kernel
// boundaries. At entry the pending update of step k0 - 1 is applied while the
// matrix is loaded; at exit d_{n-1} is written (the job of k_hh_last).
// Same formulas as k_hh_step: p = A22 v, s = v^H p, w = tau p - (|tau|^2 s / 2) v,
// A22 -= v w^H + w v^H.
constexpr int kHT = 64;

__global__ void __launch_bounds__(256) k_hh_tail(const z64* __restrict__ A, int n, int k0,
                                                 z64* __restrict__ Vst,
                                                 const z64* __restrict__ ppart_prev,
                                                 const z64* __restrict__ spart_prev,
                                                 z64* __restrict__ tau, double* __restrict__ dd,
                                                 double* __restrict__ ee) {
  __shared__ z64 As[kHT][kHT];
  __shared__ z64 vp[kHT], wp[kHT], pp[4][kHT];
  __shared__ double sh[9];
  const int mp = n - k0;  // <= kHT
  hh_prev_vw(Vst, ppart_prev, spart_prev, tau, n, k0, vp, wp, sh);
  for (int e = threadIdx.x; e < kHT * kHT; e += blockDim.x) {
    const int r = e % kHT, c = e / kHT;
    if (r < mp && c < mp)
      As[c][r] = A[(size_t)(k0 + r) + (size_t)(k0 + c) * n] - vp[r] * conj(wp[c]) -
                 wp[r] * conj(vp[c]);
  }
  __syncthreads();
  const int i4 = threadIdx.x & (kHT - 1), q4 = threadIdx.x >> 6;
  for (int k = k0; k + 1 < n; ++k) {
    const int l = k - k0, m = n - k - 1;
    const z64* x = &As[l][l + 1];  // column k below the diagonal
    z64 t_k, scale;
    double beta;
    hh_reflector(x, m, sh, &t_k, &beta, &scale);
    if (threadIdx.x == 0) {
      tau[k] = t_k;
      ee[k] = beta;
      dd[k] = As[l][l].x;
    }
    z64* vcol = Vst + (size_t)k * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const int j = r - k - 1;
      vcol[r] = j < 0 ? mk(0.0, 0.0) : (j == 0 ? mk(1.0, 0.0) : x[j] * scale);
    }
    for (int j = threadIdx.x; j < m; j += blockDim.x) vp[j] = j == 0 ? mk(1.0, 0.0) : x[j] * scale;
    __syncthreads();
    // p = A22 v: row i4, every fourth column from q4, then the four partials
    {
      z64 a = mk(0.0, 0.0);
      if (i4 < m)
        for (int j = q4; j < m; j += 4) a = a + As[l + 1 + j][l + 1 + i4] * vp[j];
      pp[q4][i4] = a;
    }
    __syncthreads();
    double sv = 0.0;
    z64 p = mk(0.0, 0.0);
    if (threadIdx.x < m) {
      p = (pp[0][threadIdx.x] + pp[1][threadIdx.x]) + (pp[2][threadIdx.x] + pp[3][threadIdx.x]);
      sv = cmulc(vp[threadIdx.x], p).x;
    }
    const double c = 0.5 * norm2(t_k) * block_sum_all(sv, sh);
    if (threadIdx.x < m) {
      const z64 tp = t_k * p;
      wp[threadIdx.x] = mk(tp.x - c * vp[threadIdx.x].x, tp.y - c * vp[threadIdx.x].y);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kHT * kHT; e += blockDim.x) {
      const int r = e % kHT, cc = e / kHT;
      if (r < m && cc < m)
        As[l + 1 + cc][l + 1 + r] =
            As[l + 1 + cc][l + 1 + r] - vp[r] * conj(wp[cc]) - wp[r] * conj(vp[cc]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) dd[n - 1] = As[mp - 1][mp - 1].x;
}

// d_{n-1}: A[n-1][n-1] after the update of step n - 2 (its last element).
__global__ void __launch_bounds__(256) k_hh_last(const z64* __restrict__ A, int n,
                                                 const z64* __restrict__ Vst,
                                                 const z64* __restrict__ ppart,
                                                 const z64* __restrict__ spart,
                                                 const z64* __restrict__ tau,
                                                 double* __restrict__ dd) {
  __shared__ double sh[9];
  __shared__ z64 vp[2], wp[2];
  if (n < 2) {
    if (threadIdx.x == 0) dd[0] = A[0].x;
    return;
  }
  hh_prev_vw(Vst, ppart, spart, tau, n, n - 1, vp, wp, sh);
  if (threadIdx.x == 0) {
    const z64 vw = vp[0] * conj(wp[0]);
    dd[n - 1] = A[(size_t)(n - 1) + (size_t)(n - 1) * n].x - 2.0 * vw.x;
  }
}

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm count).
__device__ __forceinline__ int sturm_count(const double* __restrict__ d,
                                           const double* __restrict__ e2, int n, double x,
                                           double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int j = 1; j < n; ++j) {
    q = d[j] - x - e2[j - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Eigenvalue i of the tridiagonal by 64-way multisection: one wavefront per
// eigenvalue, each lane evaluates the Sturm count at one of 64 interior
// points, and the bracket [lo, hi] with count(lo) <= i < count(hi) shrinks
// 65-fold per round. Also writes e2 = e^2 (block 0) for the caller.
__global__ void __launch_bounds__(64) k_tri_multisect(const double* __restrict__ d,
                                                      const double* __restrict__ e, int n,
                                                      double* __restrict__ e2s,
                                                      double* __restrict__ w) {
  __shared__ double e2[1024];
  const int lane = threadIdx.x, i = blockIdx.x;
  double gl = INFINITY, gu = -INFINITY, emax2 = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double el = j > 0 ? fabs(e[j - 1]) : 0.0, er = j < n - 1 ? fabs(e[j]) : 0.0;
    gl = fmin(gl, d[j] - el - er);
    gu = fmax(gu, d[j] + el + er);
    if (j < n - 1) {
      e2[j] = e[j] * e[j];
      emax2 = fmax(emax2, e2[j]);
      if (i == 0) e2s[j] = e2[j];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    gl = fmin(gl, __shfl_xor(gl, off));
    gu = fmax(gu, __shfl_xor(gu, off));
    emax2 = fmax(emax2, __shfl_xor(emax2, off));
  }
  __syncthreads();
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  double lo = gl - 2.0 * DBL_EPSILON * tnorm * n - 2.0 * pivmin;
  double hi = gu + 2.0 * DBL_EPSILON * tnorm * n + 2.0 * pivmin;
  for (int it = 0; it < 16; ++it) {
    if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const bool le = sturm_count(d, e2, n, x, pivmin) <= i;
    const unsigned long long mask = __ballot(le);
    const int below = __popcll(mask);  // lanes 0 .. below-1 have count <= i
    const double xlo = __shfl(x, below > 0 ? below - 1 : 0);
    const double xhi = __shfl(x, below < 64 ? below : 63);
    if (below > 0) lo = xlo;
    if (below < 64) hi = xhi;
  }
  if (lane == 0) w[i] = 0.5 * (lo + hi);
}

// One pass of inverse iteration over a vector stored interleaved (element j
// of column t at j * n + t), from one buffer into another. The passes work
// in blocks of kIB elements: the block's operands are loaded into registers
// first, then its results stored. (Vector stores count in the same memory
// counter as loads here, so a store between two loads would make the second
// load's wait cover the store too: one memory latency per element.)
constexpr int kIB = 16;

// Forward: the row operations of the factorisation of column i; the running
// entry stays in a register.
__device__ __forceinline__ void invit_forward(const double* __restrict__ src,
                                              double* __restrict__ dst,
                                              const double* __restrict__ mult,
                                              const double* __restrict__ swp, int n, int t, int i) {
  auto at = [n](int j, int c) { return (size_t)j * n + c; };
  double yc = src[at(0, t)];
  for (int r0 = 0; r0 < n - 1; r0 += kIB) {
    double sv[kIB], mv[kIB], wv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int r = min(r0 + u, n - 2);
      sv[u] = src[at(r + 1, t)];
      mv[u] = mult[at(r, i)];
      wv[u] = swp[at(r, i)];
    }
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int r = r0 + u;
      if (r < n - 1) {
        double yr = yc, yn = sv[u];
        if (wv[u] != 0.0) {
          const double tmp = yr;
          yr = yn;
          yn = tmp;
        }
        dst[at(r, t)] = yr;
        yc = yn - mv[u] * yr;
      }
    }
  }
  dst[at(n - 1, t)] = yc;
}

// Back substitution with U given as reciprocal pivots rinv and the two
// superdiagonals (u1 and u2 are zero where U has no entry, so every row
// reads the same operands), the two previous solution entries in registers.
__device__ __forceinline__ void invit_back(const double* __restrict__ src, double* __restrict__ dst,
                                           const double* __restrict__ rinv,
                                           const double* __restrict__ u1,
                                           const double* __restrict__ u2, int n, int t, int i) {
  auto at = [n](int j, int c) { return (size_t)j * n + c; };
  double x1 = 0.0, x2 = 0.0;
  for (int r0 = n - 1; r0 >= 0; r0 -= kIB) {
    double sv[kIB], av[kIB], bv[kIB], pv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int r = max(r0 - u, 0);
      sv[u] = src[at(r, t)];
      av[u] = u1[at(r, i)];
      bv[u] = u2[at(r, i)];
      pv[u] = rinv[at(r, i)];
    }
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int r = r0 - u;
      if (r >= 0) {
        const double x = (sv[u] - av[u] * x1 - bv[u] * x2) * pv[u];
        dst[at(r, t)] = x;
        x2 = x1;
        x1 = x;
      }
    }
  }
}

// dst_t = src_t - (zq . src_t) zq (one modified Gram-Schmidt step)
__device__ __forceinline__ void invit_mgs(const double* __restrict__ src, double* __restrict__ dst,
                                          const double* __restrict__ Z, int n, int t, int q) {
  auto at = [n](int j, int c) { return (size_t)j * n + c; };
  double dot = 0.0;
  for (int j0 = 0; j0 < n; j0 += kIB) {
    double sv[kIB], zv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int j = min(j0 + u, n - 1);
      sv[u] = src[at(j, t)];
      zv[u] = Z[at(j, q)];
    }
#pragma unroll
    for (int u = 0; u < kIB; ++u)
      if (j0 + u < n) dot += zv[u] * sv[u];
  }
  for (int j0 = 0; j0 < n; j0 += kIB) {
    double sv[kIB], zv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) {
      const int j = min(j0 + u, n - 1);
      sv[u] = src[at(j, t)];
      zv[u] = Z[at(j, q)];
    }
#pragma unroll
    for (int u = 0; u < kIB; ++u)
      if (j0 + u < n) dst[at(j0 + u, t)] = sv[u] - dot * zv[u];
  }
}

// dst_t = src_t / ||src_t||
__device__ __forceinline__ void invit_normalise(const double* __restrict__ src,
                                                double* __restrict__ dst, int n, int t) {
  auto at = [n](int j, int c) { return (size_t)j * n + c; };
  double nrm = 0.0;
  for (int j0 = 0; j0 < n; j0 += kIB) {
    double sv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) sv[u] = src[at(min(j0 + u, n - 1), t)];
#pragma unroll
    for (int u = 0; u < kIB; ++u)
      if (j0 + u < n) nrm += sv[u] * sv[u];
  }
  const double inv = 1.0 / sqrt(nrm);
  for (int j0 = 0; j0 < n; j0 += kIB) {
    double sv[kIB];
#pragma unroll
    for (int u = 0; u < kIB; ++u) sv[u] = src[at(min(j0 + u, n - 1), t)];
#pragma unroll
    for (int u = 0; u < kIB; ++u)
      if (j0 + u < n) dst[at(j0 + u, t)] = sv[u] * inv;
  }
}

// Eigenvectors of the tridiagonal by inverse iteration with partial
// pivoting LU of (T - lambda I). Eigenvalues closer than ortol form a
// cluster; one thread handles a whole cluster and orthogonalises each new
// vector against the cluster's earlier ones (modified Gram-Schmidt), so no
// thread reads another thread's vectors. Scratch and output are interleaved
// by eigenvalue index (element j of column t at j * n + t) so that
// neighbouring threads touch neighbouring addresses. Iterates move between
// the two scratch vectors xa and xb (each pass reads one and writes the
// other); the last normalisation writes Z.
__global__ void __launch_bounds__(64) k_tri_invit(const double* __restrict__ d,
                                                  const double* __restrict__ e,
                                                  const double* __restrict__ w, int n,
                                                  double rel_ortol, double* __restrict__ scr,
                                                  double* __restrict__ Z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double tnorm = 0.0;  // infinity norm of T
  for (int j = 0; j < n; ++j)
    tnorm = fmax(tnorm, fabs(d[j]) + (j > 0 ? fabs(e[j - 1]) : 0.0) + (j < n - 1 ? fabs(e[j]) : 0.0));
  const double ortol = rel_ortol * tnorm;
  if (i > 0 && w[i] - w[i - 1] <= ortol) return;  // not a cluster leader
  int end = i + 1;
  while (end < n && w[end] - w[end - 1] <= ortol) ++end;
  const size_t nn = (size_t)n * n;
  double* __restrict__ mult = scr;
  double* __restrict__ rinv = scr + nn;  // 1 / pivot (tiny pivots replaced by eps ||T||)
  double* __restrict__ u1 = scr + 2 * nn;
  double* __restrict__ u2 = scr + 3 * nn;
  double* __restrict__ swp = scr + 4 * nn;
  double* xa = scr + 5 * nn;
  double* xb = scr + 6 * nn;
  auto at = [n](int j, int t) { return (size_t)j * n + t; };
  const double eps_t = DBL_EPSILON * fmax(tnorm, DBL_MIN);
  const double pert = 10.0 * eps_t;
  auto pivot = [eps_t](double p) { return fabs(p) < eps_t ? (p < 0.0 ? -eps_t : eps_t) : p; };
  double lam_prev = -INFINITY;
  for (int t = i; t < end; ++t) {
    double lam = w[t];
    if (lam - lam_prev < pert) lam = lam_prev + pert;  // separate repeated values
    lam_prev = lam;
    // LU of T - lam I with row interchanges; U has up to two superdiagonals
    double ca = d[0] - lam, cb = n > 1 ? e[0] : 0.0, cc = 0.0;
    for (int r0 = 0; r0 < n - 1; r0 += kIB) {
      // the block's rows of T first (see kIB)
      double ev[kIB], dv[kIB], fv[kIB];
#pragma unroll
      for (int u = 0; u < kIB; ++u) {
        const int r = min(r0 + u, n - 2);
        ev[u] = e[r];
        dv[u] = d[r + 1];
        fv[u] = r + 1 < n - 1 ? e[r + 1] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kIB; ++u) {
        const int r = r0 + u;
        if (r < n - 1) {
          const double er = ev[u], an = dv[u] - lam, en = fv[u];
          if (fabs(ca) >= fabs(er)) {
            const double mu = ca != 0.0 ? er / ca : 0.0;
            mult[at(r, i)] = mu;
            swp[at(r, i)] = 0.0;
            rinv[at(r, i)] = 1.0 / pivot(ca);
            u1[at(r, i)] = cb;
            u2[at(r, i)] = cc;
            ca = an - mu * cb;
            cb = en - mu * cc;
            cc = 0.0;
          } else {
            const double mu = ca / er;
            mult[at(r, i)] = mu;
            swp[at(r, i)] = 1.0;
            rinv[at(r, i)] = 1.0 / pivot(er);
            u1[at(r, i)] = an;
            u2[at(r, i)] = en;
            const double na = cb - mu * an, nb = cc - mu * en;
            ca = na;
            cb = nb;
            cc = 0.0;
          }
        }
      }
    }
    rinv[at(n - 1, i)] = 1.0 / pivot(ca);
    // entries beyond U's last columns: zero, read against zero solution entries
    u1[at(n - 1, i)] = 0.0;
    u2[at(n - 1, i)] = 0.0;
    if (n > 1) u2[at(n - 2, i)] = 0.0;
    // start vector: deterministic pseudo-random entries in [-1, 1)
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ (unsigned long long)(t + 1) * 0xBF58476D1CE4E5B9ull;
    for (int j = 0; j < n; ++j) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      xa[at(j, t)] = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    double* cur = xa;  // the iterate
    double* oth = xb;
    for (int iter = 0; iter < 3; ++iter) {
      invit_forward(cur, oth, mult, swp, n, t, i);
      invit_back(oth, cur, rinv, u1, u2, n, t, i);
      // orthogonalise against the cluster's earlier vectors, normalise
      for (int q = i; q < t; ++q) {
        invit_mgs(cur, oth, Z, n, t, q);
        double* tmp = cur;
        cur = oth;
        oth = tmp;
      }
      if (iter == 2) {
        invit_normalise(cur, Z, n, t);
      } else {
        invit_normalise(cur, oth, n, t);
        double* tmp = cur;
        cur = oth;
        oth = tmp;
      }
    }
  }
}

// U (complex, column-major) = real Z stored interleaved (Z[j * n + t] is
// row j of eigenvector t).
__global__ void k_tri_to_complex(const double* __restrict__ Z, int n, z64* __restrict__ U) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int j = (int)(idx % n), t = (int)(idx / n);
  U[idx] = mk(Z[(size_t)j * n + t], 0.0);
}

// Triangular factor T (b x b, column-major, upper) of the compact WY form
// H_k0 ... H_{k0+b-1} = I - V T V^H (forward, columnwise):
// T_cc = tau_c, T[0:c, c] = -tau_c T[0:c, 0:c] (V[:, 0:c]^H v_c), with the
// Gram matrix Gm = V^H V (b x b, column-major) computed beforehand by a GEMM.
__global__ void __launch_bounds__(256) k_wy_tmat(const z64* __restrict__ Gm, int k0, int b,
                                                 const z64* __restrict__ tau, z64* __restrict__ T) {
  __shared__ z64 G[kWY][kWY + 1];
  __shared__ z64 Ts[kWY][kWY + 1];
  for (int t = threadIdx.x; t < kWY * kWY; t += blockDim.x) {
    const int a = t % kWY, c = t / kWY;
    G[a][c] = (a < c && c < b) ? Gm[a + (size_t)c * b] : mk(0.0, 0.0);
    Ts[a][c] = mk(0.0, 0.0);
  }
  __syncthreads();
  // row a of column c by the 8 lanes a * 8 + z (z: every eighth q), summed
  // with shuffles inside the lane group (256 threads = kWY rows x 8)
  static_assert(kWY * 8 == 256, "one lane group of 8 per row");
  const int a = threadIdx.x >> 3, z = threadIdx.x & 7;
  for (int c = 0; c < b; ++c) {
    const z64 tc = tau[k0 + c];
    z64 s = mk(0.0, 0.0);
    if (a < c)
      for (int q = a + z; q < c; q += 8) s = s + Ts[a][q] * G[q][c];
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
      s.x += __shfl_xor(s.x, off, 8);
      s.y += __shfl_xor(s.y, off, 8);
    }
    if (a < c && z == 0) Ts[a][c] = mk(-(tc.x * s.x - tc.y * s.y), -(tc.x * s.y + tc.y * s.x));
    if (threadIdx.x == 0) Ts[c][c] = tc;
    __syncthreads();
  }
  for (int t = threadIdx.x; t < b * b; t += blockDim.x) T[t] = Ts[t % b][t / b];
}

// Workspace layout (bytes aligned to 256 by construction of the sizes).
struct EighWs {
  z64 *vst, *ppart, *spart, *tau, *tmat, *w1, *w2;
  double *dd, *ee, *e2, *scr, *zt;
};
static int eigh_ws(chs_context* ctx, int n, EighWs* ws) {
  const size_t nn = (size_t)n * n;
  const int nch = (n + kHC - 1) / kHC;
  const size_t zb = (nn + 2 * (size_t)nch * n + 2 * (size_t)nch * nch + n + kWY * kWY +
                     2 * (size_t)kWY * n) * sizeof(z64);
  const size_t db = (3 * (size_t)n + 8 * nn) * sizeof(double);
  CHS_TRY(ensure(ctx, ctx->impl->eig_ws, zb + db + 1024));
  char* p = (char*)ctx->impl->eig_ws.p;
  auto take_z = [&p](size_t cnt) {
    z64* r = (z64*)p;
    p += ((cnt * sizeof(z64) + 255) / 256) * 256;
    return r;
  };
  auto take_d = [&p](size_t cnt) {
    double* r = (double*)p;
    p += ((cnt * sizeof(double) + 255) / 256) * 256;
    return r;
  };
  ws->vst = take_z(nn);
  ws->ppart = take_z(2 * (size_t)nch * n);   // two halves: steps alternate
  ws->spart = take_z(2 * (size_t)nch * nch);
  ws->tau = take_z(n);
  ws->tmat = take_z(kWY * kWY);
  ws->w1 = take_z((size_t)kWY * n);
  ws->w2 = take_z((size_t)kWY * n);
  ws->dd = take_d(n);
  ws->ee = take_d(n);
  ws->e2 = take_d(n);
  ws->scr = take_d(7 * nn);  // LU factors (5 n x n) and two iterate vectors
  ws->zt = take_d(nn);
  return CHS_OK;
}

static void tridiagonalise_launches(chs_context* ctx, z64* A, int n, const EighWs& ws) {
  // partial products alternate between two halves of ppart / spart
  const int nch0 = (n + kHC - 1) / kHC;
  z64* pp[2] = {ws.ppart, ws.ppart + (size_t)nch0 * n};
  z64* sp[2] = {ws.spart, ws.spart + (size_t)nch0 * nch0};
  // steps from k0 on (trailing matrix [k0, n)^2 of at most kHT^2) run in one
  // workgroup with the matrix in LDS
  const int k0 = n >= 3 ? std::max(1, n - kHT) : n;
  for (int k = 0; k + 1 < n; ++k) {
    if (k == k0) {
      hipLaunchKernelGGL(k_hh_tail, dim3(1), dim3(256), 0, ctx->stream, A, n, k0, ws.vst,
                         pp[(k0 - 1) & 1], sp[(k0 - 1) & 1], ws.tau, ws.dd, ws.ee);
      return;
    }
    const int m = n - k - 1;
    const unsigned nc = ceil_div(m, kHC);
    if (k == 0)
      hipLaunchKernelGGL(k_hh_symv, dim3(nc, nc), dim3(256), 0, ctx->stream, A, n, k, ws.vst,
                         pp[0], sp[0], ws.tau, ws.dd, ws.ee);
    else
      hipLaunchKernelGGL(k_hh_step, dim3(nc, nc), dim3(256), 0, ctx->stream, A, n, k, ws.vst,
                         pp[(k - 1) & 1], sp[(k - 1) & 1], pp[k & 1], sp[k & 1], ws.tau, ws.dd,
                         ws.ee);
  }
  hipLaunchKernelGGL(k_hh_last, dim3(1), dim3(256), 0, ctx->stream, A, n, ws.vst,
                     pp[n >= 2 ? (n - 2) & 1 : 0], sp[n >= 2 ? (n - 2) & 1 : 0], ws.tau, ws.dd);
}

// Hermitian eigensolve: A (n x n, destroyed) -> eigenvalues ascending in
// device w, orthonormal eigenvectors in U (n x n column-major).
int heev_tridiag(chs_context* ctx, z64* A, int n, double* w, z64* U) {
  if (n < 1 || n > 1024) return set_error(ctx, CHS_ERR_INVALID_ARG, "eigensolver limited to 1 <= n <= 1024");
  EighWs ws;
  CHS_TRY(eigh_ws(ctx, n, &ws));
  // 1. tridiagonal reduction (2 (n - 1) launches; graph-replayed)
  EighGraph& P = ctx->impl->eigh_graph;
  if (ctx->impl->use_graphs && n > 1) {
    if (!P.exec || P.n != n || P.gA != A || P.gws != ctx->impl->eig_ws.p) {
      // the key names every address the graph holds: A and the workspace
      // (all of ws lies in eig_ws, laid out by n)
      if (P.exec) (void)hipGraphExecDestroy(P.exec);
      P = EighGraph{};
      CHS_TRY(capture_graph(
          ctx,
          [&](chs_context* c) {
            tridiagonalise_launches(c, A, n, ws);
            return CHS_OK;
          },
          &P.exec));
      P.n = n;
      P.gA = A;
      P.gws = ctx->impl->eig_ws.p;
    }
    CHS_HIP(ctx, hipGraphLaunch(P.exec, ctx->stream));
  } else {
    tridiagonalise_launches(ctx, A, n, ws);
  }
  CHS_LAUNCH_CHECK(ctx);
  // 2. eigenvalues (multisection) and tridiagonal eigenvectors
  hipLaunchKernelGGL(k_tri_multisect, dim3(n), dim3(64), 0, ctx->stream, ws.dd, ws.ee, n, ws.e2, w);
  CHS_LAUNCH_CHECK(ctx);
  // clusters: gaps below 1e-6 ||T||. A vector whose neighbours lie further
  // away carries an angle error of about eps ||T|| / gap < 3e-10 without
  // reorthogonalisation, while a looser cut lets ill-conditioned generalised
  // problems (cond S ~ 1e5) form clusters of hundreds of eigenvalues, each
  // of which one thread works through at O(n c^2)
  hipLaunchKernelGGL(k_tri_invit, dim3(ceil_div(n, 64)), dim3(64), 0, ctx->stream, ws.dd, ws.ee, w,
                     n, 1e-6, ws.scr, ws.zt);
  hipLaunchKernelGGL(k_tri_to_complex, dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0,
                     ctx->stream, ws.zt, n, U);
  CHS_LAUNCH_CHECK(ctx);
  // 3. U <- Q U, blocks of reflectors from the last to the first
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0), mone = mk(-1.0, 0.0);
  const int nref = n - 1;
  for (int kb = ((nref - 1) / kWY) * kWY; nref > 0 && kb >= 0; kb -= kWY) {
    const int b = std::min(kWY, nref - kb);
    const int rows = n - kb - 1;
    const z64* V = ws.vst + (size_t)kb * n + (kb + 1);
    // V^H V into w2 (free until the T w1 product below)
    CHS_TRY(zgemm(ctx, 'C', 'N', b, b, rows, one, V, n, V, n, zero, ws.w2, b));
    hipLaunchKernelGGL(k_wy_tmat, dim3(1), dim3(256), 0, ctx->stream, ws.w2, kb, b, ws.tau,
                       ws.tmat);
    CHS_LAUNCH_CHECK(ctx);
    CHS_TRY(zgemm(ctx, 'C', 'N', b, n, rows, one, V, n, U + (kb + 1), n, zero, ws.w1, b));
    CHS_TRY(zgemm(ctx, 'N', 'N', b, n, b, one, ws.tmat, b, ws.w1, b, zero, ws.w2, b));
    CHS_TRY(zgemm(ctx, 'N', 'N', rows, n, b, mone, V, n, ws.w2, b, one, U + (kb + 1), n));
  }
  return CHS_OK;
}

}  // namespace chs
