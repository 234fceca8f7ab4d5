// FP64 dense linear algebra on the device: complex GEMM (split-K for the
// tall-skinny Gram products), Hermitisation, Cholesky, triangular inverse,
// a parallel cyclic Jacobi Hermitian eigensolver, and the Rayleigh-Ritz /
// Cholesky-QR drivers built from them.
//
// Reference: rayleigh_ritz (src/eigensolver/rayleigh_ritz.rs:48-245) calls
// cuBLAS ZGEMM for H_sub and S_sub and cuSOLVER ZHEGVD; the semantics
// matched here are diagonalize_and_rotate (src/eigensolver/
// subspace_diag.rs:229-391): S = L L^H, B = L^{-1} H L^{-H} (Hermitised),
// B U = U diag(eps) ascending, V = L^{-H} U, psi <- psi V.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <numeric>
#include <vector>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"

namespace chs {

// ---------------------------------------------------------------------------
// Small dense kernels on n x n column-major matrices
// ---------------------------------------------------------------------------

// M <- (M + M^H) / 2 (subspace_diag.rs:117-125)
__global__ void k_hermitise(z64* __restrict__ M, int n) {
  const int j = blockIdx.x, i0 = threadIdx.x;
  for (int i = i0; i <= j; i += blockDim.x) {
    z64 a = M[(size_t)i + (size_t)j * n], b = M[(size_t)j + (size_t)i * n];
    z64 v = mk(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    if (i == j) v.y = 0.0;
    M[(size_t)i + (size_t)j * n] = v;
    M[(size_t)j + (size_t)i * n] = conj(v);
  }
}

// Blocked right-looking lower Cholesky, panel width kCB. Per panel:
// k_chol_panel factors the diagonal block in LDS (every block of the launch
// does it redundantly, which is cheaper than a launch) and solves its own
// rows of the panel, X L_kk^H = A; k_chol_update subtracts L_p L_p^H from
// the trailing lower triangle by 32 x 32 tiles. info[0] = failing column + 1
// (0 on success); every kernel returns at once when it is set.
constexpr int kCB = 32;

__global__ void __launch_bounds__(256) k_chol_panel(z64* __restrict__ A, int n, int kb,
                                                    z64* __restrict__ dscr,
                                                    int* __restrict__ info) {
  __shared__ z64 D[kCB][kCB + 1];
  __shared__ int bad;
  if (info[0]) return;
  const int b = min(kCB, n - kb);
  for (int t = threadIdx.x; t < kCB * kCB; t += blockDim.x) {
    const int i = t % kCB, j = t / kCB;
    D[i][j] = (i < b && j < b && j <= i) ? A[(size_t)(kb + i) + (size_t)(kb + j) * n] : mk(0.0, 0.0);
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    const double d = D[j][j].x;
    if (!(d > 0.0)) {
      if (threadIdx.x == 0) bad = kb + j + 1;
      break;  // uniform: every thread reads the same d
    }
    const double piv = sqrt(d), inv = 1.0 / piv;
    __syncthreads();
    if (threadIdx.x == 0) D[j][j] = mk(piv, 0.0);
    for (int i = j + 1 + threadIdx.x; i < b; i += blockDim.x) D[i][j] = mk(D[i][j].x * inv, D[i][j].y * inv);
    __syncthreads();
    for (int t = threadIdx.x; t < kCB * kCB; t += blockDim.x) {
      const int i = t % kCB, c = t / kCB;
      if (c > j && c <= i && i < b) D[i][c] = D[i][c] - D[i][j] * conj(D[c][j]);
    }
    __syncthreads();
  }
  __syncthreads();
  if (bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) info[0] = bad;
    return;
  }
  // the factored diagonal block goes to dscr, not A: other blocks of this
  // launch may still be reading A's diagonal block
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < kCB * kCB; t += blockDim.x) {
      const int i = t % kCB, j = t / kCB;
      dscr[(size_t)(kb / kCB) * kCB * kCB + t] = j <= i ? D[i][j] : mk(0.0, 0.0);
    }
  // rows of the panel below the diagonal block: x L_kk^H = a, one row per thread
  const int i = kb + b + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  z64 x[kCB];
#pragma unroll
  for (int c = 0; c < kCB; ++c) x[c] = c < b ? A[(size_t)i + (size_t)(kb + c) * n] : mk(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < kCB; ++c) {
    if (c < b) {
      z64 s = x[c];
#pragma unroll
      for (int k = 0; k < c; ++k) s = s - x[k] * conj(D[c][k]);
      const double inv = 1.0 / D[c][c].x;
      x[c] = mk(s.x * inv, s.y * inv);
    }
  }
#pragma unroll
  for (int c = 0; c < kCB; ++c)
    if (c < b) A[(size_t)i + (size_t)(kb + c) * n] = x[c];
}

// A[i][j] -= sum_c L[i][c] conj(L[j][c]) over the panel columns c, for the
// lower-triangle tiles (ti >= tj) of the trailing matrix starting at k0.
__global__ void __launch_bounds__(256) k_chol_update(z64* __restrict__ A, int n, int kb, int k0,
                                                     const int* __restrict__ info) {
  __shared__ z64 Li[32][kCB + 1], Lj[32][kCB + 1];
  if (info[0]) return;
  // blockIdx.x enumerates lower tiles (ti, tj), tj <= ti
  int t = blockIdx.x, ti = 0;
  while (t > ti) {
    t -= ti + 1;
    ++ti;
  }
  const int tj = t;
  const int r0 = k0 + 32 * ti, c0 = k0 + 32 * tj;
  const int b = min(kCB, n - kb);
  for (int e = threadIdx.x; e < 32 * kCB; e += blockDim.x) {
    const int r = e % 32, c = e / 32;
    const bool okc = c < b;
    Li[r][c] = (okc && r0 + r < n) ? A[(size_t)(r0 + r) + (size_t)(kb + c) * n] : mk(0.0, 0.0);
    Lj[r][c] = (okc && c0 + r < n) ? A[(size_t)(c0 + r) + (size_t)(kb + c) * n] : mk(0.0, 0.0);
  }
  __syncthreads();
  // 256 threads x 4 outputs: thread (r, cq) computes rows r, columns cq*4 .. +3
  const int r = threadIdx.x % 32, cq = threadIdx.x / 32;
  z64 acc[4] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
  for (int c = 0; c < b; ++c) {
    const z64 a = Li[r][c];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = acc[q] + a * conj(Lj[cq * 4 + q][c]);
  }
  const int i = r0 + r;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = c0 + cq * 4 + q;
    if (j <= i && j < n) {
      z64* p = A + (size_t)i + (size_t)j * n;
      *p = *p - acc[q];
    }
  }
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// Final Cholesky factor: diagonal blocks from dscr, strict upper triangle 0.
__global__ void k_chol_finish(z64* __restrict__ A, int n, const z64* __restrict__ dscr,
                              const int* __restrict__ info) {
  if (info[0]) return;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * n) return;
  const int i = (int)(t % n), j = (int)(t / n);
  if (i < j)
    A[t] = mk(0.0, 0.0);
  else if (i / kCB == j / kCB)
    A[t] = dscr[(size_t)(i / kCB) * kCB * kCB + (i % kCB) + (size_t)(j % kCB) * kCB];
}

// X = L^{-1} for n <= 64 R: one wavefront per CPW columns j0 .. j0+CPW-1,
// column-oriented forward substitution (x_k /= L_kk; x_i -= L_ik x_k for
// i > k) with lane l holding rows l + 64 r in registers, so L is read one
// coalesced column at a time and no workgroup barrier is needed.
template <int R, int CPW>
__global__ void __launch_bounds__(64) k_tri_inv_wave(const z64* __restrict__ L, z64* __restrict__ X,
                                                     int n) {
  const int lane = threadIdx.x;
  const int j0 = blockIdx.x * CPW;
  z64 x[CPW][R];
#pragma unroll
  for (int c = 0; c < CPW; ++c)
#pragma unroll
    for (int r = 0; r < R; ++r) x[c][r] = mk(lane + 64 * r == j0 + c ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int rb = 0; rb < R; ++rb) {
    const int kbeg = max(j0, 64 * rb), kend = min(n, 64 * rb + 64);
    for (int k = kbeg; k < kend; ++k) {
      const int lk = k & 63;
      const z64* col = L + (size_t)k * n;
      const double inv = 1.0 / col[k].x;
      z64 xk[CPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        xk[c] = mk(__shfl(x[c][rb].x, lk) * inv, __shfl(x[c][rb].y, lk) * inv);
        if (lane == lk) x[c][rb] = xk[c];
      }
#pragma unroll
      for (int r = rb; r < R; ++r) {
        const int row = lane + 64 * r;
        if (row > k && row < n) {
          const z64 l = col[row];
#pragma unroll
          for (int c = 0; c < CPW; ++c) x[c][r] = x[c][r] - l * xk[c];
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    if (j0 + c >= n) break;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = lane + 64 * r;
      if (row < n) X[(size_t)row + (size_t)(j0 + c) * n] = x[c][r];
    }
  }
}

// X = L^{-1} (lower) for n > 1024: one block, rows in sequence.

__global__ void __launch_bounds__(1024) k_tri_inv(const z64* __restrict__ L, z64* __restrict__ X, int n) {
  for (long t = threadIdx.x; t < (long)n * n; t += blockDim.x) X[t] = mk(0.0, 0.0);
  __syncthreads();
  for (int i = 0; i < n; ++i) {
    const z64 lii = L[(size_t)i + (size_t)i * n];
    const double inv = 1.0 / lii.x;  // real positive diagonal
    for (int c = threadIdx.x; c <= i; c += blockDim.x) {
      z64 s = (c == i) ? mk(1.0, 0.0) : mk(0.0, 0.0);
      for (int k = c; k < i; ++k) s = s - L[(size_t)i + (size_t)k * n] * X[(size_t)k + (size_t)c * n];
      X[(size_t)i + (size_t)c * n] = mk(s.x * inv, s.y * inv);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Parallel cyclic Jacobi (round-robin ordering). Per round, k_jacobi_cols
// computes each pair's rotation U = [[c, s e], [-s e*, c]] from its own
// 2 x 2 block and applies A <- A U and V <- V U on columns p, q;
// k_jacobi_rows then applies A <- U^H A on rows p, q.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_jacobi_cols(z64* __restrict__ A, z64* __restrict__ V, int n,
                                                     const int2* __restrict__ pairs,
                                                     double4* __restrict__ rot) {
  const int2 pq = pairs[blockIdx.x];
  const int p = pq.x, q = pq.y;
  __shared__ double4 r;
  if (threadIdx.x == 0) {
    double c = 1.0, s = 0.0, er = 1.0, ei = 0.0;
    if (q < n && p < n) {
      const z64 apq = A[(size_t)p + (size_t)q * n];
      const double mag = sqrt(apq.x * apq.x + apq.y * apq.y);
      const double app = A[(size_t)p + (size_t)p * n].x, aqq = A[(size_t)q + (size_t)q * n].x;
      if (mag > 1e-300 && mag > 1e-18 * sqrt(fabs(app * aqq))) {
        er = apq.x / mag;
        ei = apq.y / mag;
        const double tau = (aqq - app) / (2.0 * mag);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        c = 1.0 / sqrt(1.0 + t * t);
        s = t * c;
      }
    }
    r = make_double4(c, s, er, ei);
    rot[blockIdx.x] = r;
  }
  __syncthreads();
  if (p >= n || q >= n || r.y == 0.0) return;
  const double c = r.x, s = r.y;
  const z64 e = mk(r.z, r.w), ec = mk(r.z, -r.w);
  z64* cp = A + (size_t)p * n;
  z64* cq = A + (size_t)q * n;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const z64 akp = cp[k], akq = cq[k];
    cp[k] = c * akp - s * (ec * akq);
    cq[k] = s * (e * akp) + c * akq;
  }
  z64* vp = V + (size_t)p * n;
  z64* vq = V + (size_t)q * n;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const z64 vkp = vp[k], vkq = vq[k];
    vp[k] = c * vkp - s * (ec * vkq);
    vq[k] = s * (e * vkp) + c * vkq;
  }
}

__global__ void __launch_bounds__(256) k_jacobi_rows(z64* __restrict__ A, int n,
                                                     const int2* __restrict__ pairs,
                                                     const double4* __restrict__ rot) {
  const int2 pq = pairs[blockIdx.x];
  const int p = pq.x, q = pq.y;
  const double4 r = rot[blockIdx.x];
  if (p >= n || q >= n || r.y == 0.0) return;
  const double c = r.x, s = r.y;
  const z64 e = mk(r.z, r.w), ec = mk(r.z, -r.w);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const z64 apk = A[(size_t)p + (size_t)k * n], aqk = A[(size_t)q + (size_t)k * n];
    A[(size_t)p + (size_t)k * n] = c * apk - s * (e * aqk);
    A[(size_t)q + (size_t)k * n] = s * (ec * apk) + c * aqk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A[(size_t)p + (size_t)q * n] = mk(0.0, 0.0);
    A[(size_t)q + (size_t)p * n] = mk(0.0, 0.0);
    A[(size_t)p + (size_t)p * n].y = 0.0;
    A[(size_t)q + (size_t)q * n].y = 0.0;
  }
}

// out[0] = sum_{i != j} |A_ij|^2, out[1] = sum |A_ij|^2 (one block).
__global__ void k_offnorm(const z64* __restrict__ A, int n, double* __restrict__ out) {
  __shared__ double sh[8];
  double off = 0, all = 0;
  for (long t = threadIdx.x; t < (long)n * n; t += blockDim.x) {
    const int i = (int)(t % n), j = (int)(t / n);
    const double v = norm2(A[t]);
    all += v;
    if (i != j) off += v;
  }
  double a = block_sum(off, sh);
  double b = block_sum(all, sh);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

__global__ void k_set_identity(z64* __restrict__ V, int n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (size_t)n * n) V[t] = mk((t % n) == (t / n) ? 1.0 : 0.0, 0.0);
}

// Ascending order of the diagonal (ties by index); eigenvalues to w, columns
// of V permuted into U.
__global__ void __launch_bounds__(1024) k_sort_eig(const z64* __restrict__ A, const z64* __restrict__ V,
                                                   int n, double* __restrict__ w,
                                                   z64* __restrict__ U) {
  __shared__ int rank_of[1024];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double di = A[(size_t)i + (size_t)i * n].x;
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[(size_t)j + (size_t)j * n].x;
      r += (dj < di) || (dj == di && j < i);
    }
    rank_of[i] = r;
    w[r] = di;
  }
  __syncthreads();
  for (long t = threadIdx.x; t < (long)n * n; t += blockDim.x) {
    const int k = (int)(t % n), col = (int)(t / n);
    U[(size_t)k + (size_t)rank_of[col] * n] = V[t];
  }
}

static int jacobi_plan(chs_context* ctx, int n, JacobiPlan& P) {
  if (P.n == n && P.sched.p) return CHS_OK;
  if (P.exec) (void)hipGraphExecDestroy(P.exec);  // the captured sweep bakes in n and buffers
  P.exec = nullptr;
  P.gA = P.gV = nullptr;
  const int m = (n % 2) ? n + 1 : n;  // add a dummy index for odd n
  P.n = n;
  P.rounds = m - 1;
  P.pairs = m / 2;
  std::vector<int> ring(m);
  std::iota(ring.begin(), ring.end(), 0);
  std::vector<int2> s((size_t)P.rounds * P.pairs);
  for (int r = 0; r < P.rounds; ++r) {
    for (int k = 0; k < P.pairs; ++k) {
      int a = ring[k], b = ring[m - 1 - k];
      if (a > b) std::swap(a, b);
      s[(size_t)r * P.pairs + k] = make_int2(a, b);  // b == n marks the dummy
    }
    // circle method: keep ring[0], rotate the rest
    const int last = ring[m - 1];
    for (int k = m - 1; k > 1; --k) ring[k] = ring[k - 1];
    ring[1] = last;
  }
  CHS_TRY(ensure(ctx, P.sched, s.size() * sizeof(int2)));
  CHS_HIP(ctx, hipMemcpyAsync(P.sched.p, s.data(), s.size() * sizeof(int2), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += s.size() * sizeof(int2);
  CHS_TRY(ensure(ctx, P.rot, (size_t)P.pairs * sizeof(double4)));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

// Hermitian eigensolver: A (n x n) is destroyed; eigenvalues ascending in
// device w, eigenvectors in U. Work: V (n x n).
int heev_jacobi(chs_context* ctx, z64* A, int n, double* w, z64* U, z64* V, int max_sweeps,
                int* sweeps_out) {
  JacobiPlan& P = ctx->impl->jacobi;
  CHS_TRY(jacobi_plan(ctx, n, P));
  hipLaunchKernelGGL(k_set_identity, dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0, ctx->stream,
                     V, n);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(ensure(ctx, ctx->impl->small, 64 * sizeof(double)));
  double* dn = (double*)ctx->impl->small.p + 48;
  int sweeps = 0;
  for (; sweeps < max_sweeps; ++sweeps) {
    hipLaunchKernelGGL(k_offnorm, dim3(1), dim3(256), 0, ctx->stream, A, n, dn);
    CHS_LAUNCH_CHECK(ctx);
    double h[2];
    CHS_HIP(ctx, hipMemcpyAsync(h, dn, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->d2h += 2 * sizeof(double);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h[0] <= 1e-30 * h[1] || h[0] == 0.0) break;
    // one sweep = 2 (n-1) small launches: captured once into a graph and
    // replayed (no memset nodes inside)
    if (!ctx->impl->use_graphs) {
      for (int r = 0; r < P.rounds; ++r) {
        const int2* pr = (const int2*)P.sched.p + (size_t)r * P.pairs;
        hipLaunchKernelGGL(k_jacobi_cols, dim3(P.pairs), dim3(256), 0, ctx->stream, A, V, n, pr,
                           (double4*)P.rot.p);
        hipLaunchKernelGGL(k_jacobi_rows, dim3(P.pairs), dim3(256), 0, ctx->stream, A, n, pr,
                           (const double4*)P.rot.p);
      }
      CHS_LAUNCH_CHECK(ctx);
      continue;
    }
    if (!P.exec || P.gA != A || P.gV != V) {
      // the schedule and rotation buffers are fixed for this n (jacobi_plan
      // drops the graph when n changes); A and V are the key
      if (P.exec) (void)hipGraphExecDestroy(P.exec);
      P.exec = nullptr;
      P.gA = P.gV = nullptr;
      CHS_TRY(capture_graph(
          ctx,
          [&](chs_context* c) {
            for (int r = 0; r < P.rounds; ++r) {
              const int2* pr = (const int2*)P.sched.p + (size_t)r * P.pairs;
              hipLaunchKernelGGL(k_jacobi_cols, dim3(P.pairs), dim3(256), 0, c->stream, A, V, n,
                                 pr, (double4*)P.rot.p);
              hipLaunchKernelGGL(k_jacobi_rows, dim3(P.pairs), dim3(256), 0, c->stream, A, n, pr,
                                 (const double4*)P.rot.p);
            }
            return CHS_OK;
          },
          &P.exec));
      P.gA = A;
      P.gV = V;
    }
    CHS_HIP(ctx, hipGraphLaunch(P.exec, ctx->stream));
  }
  if (sweeps_out) *sweeps_out = sweeps;
  if (n > 1024) return set_error(ctx, CHS_ERR_SOLVER, "eigensolver limited to n <= 1024");
  hipLaunchKernelGGL(k_sort_eig, dim3(1), dim3(1024), 0, ctx->stream, A, V, n, w, U);
  CHS_LAUNCH_CHECK(ctx);
  if (sweeps == max_sweeps)
    return set_error(ctx, CHS_ERR_SOLVER, "Jacobi eigensolver did not converge in %d sweeps",
                     max_sweeps);
  return CHS_OK;
}

int cholesky(chs_context* ctx, z64* A, int n, int* info_host) {
  CHS_TRY(ensure(ctx, ctx->impl->small, 64 * sizeof(double)));
  int* dinfo = (int*)((double*)ctx->impl->small.p + 56);
  CHS_TRY(ensure(ctx, ctx->impl->chol_ws, (size_t)ceil_div(n, kCB) * kCB * kCB * sizeof(z64)));
  z64* dscr = (z64*)ctx->impl->chol_ws.p;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, dinfo, 1, 0);
  for (int kb = 0; kb < n; kb += kCB) {
    const int b = std::min(kCB, n - kb);
    const int rows = n - kb - b;
    hipLaunchKernelGGL(k_chol_panel, dim3(std::max(1u, ceil_div(rows, 256))), dim3(256), 0,
                       ctx->stream, A, n, kb, dscr, dinfo);
    if (rows > 0) {
      const int t = (int)ceil_div(rows, 32);
      hipLaunchKernelGGL(k_chol_update, dim3(t * (t + 1) / 2), dim3(256), 0, ctx->stream, A, n, kb,
                         kb + b, dinfo);
    }
  }
  hipLaunchKernelGGL(k_chol_finish, dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0, ctx->stream,
                     A, n, dscr, dinfo);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(info_host, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ctx->d2h += sizeof(int);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

int tri_inv(chs_context* ctx, const z64* L, z64* X, int n) {
  if (n <= 64)
    hipLaunchKernelGGL((k_tri_inv_wave<1, 4>), dim3(ceil_div(n, 4)), dim3(64), 0, ctx->stream, L, X, n);
  else if (n <= 128)
    hipLaunchKernelGGL((k_tri_inv_wave<2, 4>), dim3(ceil_div(n, 4)), dim3(64), 0, ctx->stream, L, X, n);
  else if (n <= 256)
    hipLaunchKernelGGL((k_tri_inv_wave<4, 4>), dim3(ceil_div(n, 4)), dim3(64), 0, ctx->stream, L, X, n);
  else if (n <= 512)
    hipLaunchKernelGGL((k_tri_inv_wave<8, 2>), dim3(ceil_div(n, 2)), dim3(64), 0, ctx->stream, L, X, n);
  else if (n <= 1024)
    hipLaunchKernelGGL((k_tri_inv_wave<16, 1>), dim3(n), dim3(64), 0, ctx->stream, L, X, n);
  else
    hipLaunchKernelGGL(k_tri_inv, dim3(1), dim3(1024), 0, ctx->stream, L, X, n);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int hermitise(chs_context* ctx, z64* M, int n) {
  hipLaunchKernelGGL(k_hermitise, dim3(n), dim3(256), 0, ctx->stream, M, n);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// Generalized (B != nullptr) or standard Hermitian eigensolve of device
// matrices (n x n column-major); eigenvectors overwrite A, eigenvalues in
// device w. B is overwritten by its Cholesky factor.
int zhegv(chs_context* ctx, z64* A, z64* B, int n, double* w) {
  auto* im = ctx->impl;
  const size_t nn = (size_t)n * n;
  CHS_TRY(ensure(ctx, im->lin_ws, 4 * nn * sizeof(z64)));
  z64* Li = (z64*)im->lin_ws.p;
  z64* T = Li + nn;
  z64* U = T + nn;
  z64* V = U + nn;
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  if (B) {
    int info = 0;
    CHS_TRY(cholesky(ctx, B, n, &info));
    if (info) {
      ctx->diag.solver_info = info;
      return set_error(ctx, CHS_ERR_RAYLEIGH_RITZ_FAILED,
                       "Cholesky of S_sub failed at column %d (not positive definite)", info);
    }
    CHS_TRY(tri_inv(ctx, B, Li, n));
    // T = Li A ; A <- T Li^H
    CHS_TRY(zgemm(ctx, 'N', 'N', n, n, n, one, Li, n, A, n, zero, T, n));
    CHS_TRY(zgemm(ctx, 'N', 'C', n, n, n, one, T, n, Li, n, zero, A, n));
    CHS_TRY(hermitise(ctx, A, n));
  }
  if (std::getenv("CHS_EIGH_JACOBI")) {
    int sweeps = 0;
    CHS_TRY(heev_jacobi(ctx, A, n, w, U, V, 60, &sweeps));
  } else {
    CHS_TRY(heev_tridiag(ctx, A, n, w, U));
  }
  if (B) {
    // A <- Li^H U
    CHS_TRY(zgemm(ctx, 'C', 'N', n, n, n, one, Li, n, U, n, zero, A, n));
  } else {
    CHS_HIP(ctx, hipMemcpyAsync(A, U, nn * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  }
  return CHS_OK;
}

// Rayleigh-Ritz on a c128 block Y (n_b x n_pw band-major): see file header.
__global__ void k_widen(const c32* __restrict__ a, z64* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mk((double)a[i].x, (double)a[i].y);
}

int rayleigh_ritz_impl(chs_context* ctx, z64* Y, int n_b, double* eps_host, z64* HY_out,
                       double* eps_dev_out, z64* w_out, const c32* Y32) {
  CHS_TRY(ensure_precision_ready(ctx, CHS_C128));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const int n_pw = H.n_pw;
  const size_t ne = (size_t)n_b * n_pw;
  CHS_TRY(ensure(ctx, im->ws4, ne * sizeof(z64)));
  CHS_TRY(ensure(ctx, im->ws5, ne * sizeof(z64)));
  z64* HY = (z64*)im->ws4.p;
  z64* SY = (z64*)im->ws5.p;
  const bool gen = H.has_s;
  // with projectors and no H|Y> wanted: the local H|Y> only, the projector
  // parts of H_sub and S_sub in coefficient space (subspace_nl)
  const bool nl = H.n_proj > 0 && !HY_out;
  const bool lo = nl && Y32;
  if (lo) {
    CHS_TRY(ensure_precision_ready(ctx, CHS_C64));
    CHS_TRY(ensure(ctx, im->rr32, ne * sizeof(c32)));
    CHS_TRY(vloc_pipeline(ctx, CHS_C64, Y32, n_b, im->rr32.p, EPI_HPSI, nullptr, nullptr, 0.0, 0.0,
                          nullptr, nullptr));
    hipLaunchKernelGGL(k_widen, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                       (const c32*)im->rr32.p, HY, ne);
    CHS_LAUNCH_CHECK(ctx);
  } else if (nl)
    CHS_TRY(vloc_pipeline(ctx, CHS_C128, Y, n_b, HY, EPI_HPSI, nullptr, nullptr, 0.0, 0.0, nullptr,
                          nullptr));
  else
    CHS_TRY(apply_h_impl(ctx, CHS_C128, Y, n_b, HY, gen ? SY : nullptr));
  const size_t nn = (size_t)n_b * n_b;
  CHS_TRY(ensure(ctx, im->scf_ws, (2 * nn + n_b) * sizeof(z64) + 64));
  z64* Hs = (z64*)im->scf_ws.p;
  z64* Ss = Hs + nn;
  double* w = (double*)(Ss + nn);
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  // both Gram products are Hermitian: upper tiles only, lower triangle mirrored
  CHS_TRY(zgemm_herm(ctx, n_b, n_pw, Y, n_pw, HY, n_pw, Hs, n_b));
  CHS_TRY(zgemm_herm(ctx, n_b, n_pw, Y, n_pw, gen && !nl ? SY : Y, n_pw, Ss, n_b));
  if (nl) CHS_TRY(subspace_nl(ctx, Y, n_b, Hs, gen ? Ss : nullptr, lo ? Y32 : nullptr));
  CHS_TRY(hermitise(ctx, Hs, n_b));
  CHS_TRY(hermitise(ctx, Ss, n_b));
  CHS_TRY(zhegv(ctx, Hs, Ss, n_b, w));
  // Y <- Y V (V in Hs); HY <- HY V
  CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, n_b, one, Y, n_pw, Hs, n_b, zero, SY, n_pw));
  CHS_HIP(ctx, hipMemcpyAsync(Y, SY, ne * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  if (HY_out) CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, n_b, one, HY, n_pw, Hs, n_b, zero, HY_out, n_pw));
  // beta^H (Y V) = (beta^H Y) V: subspace_nl left beta^H Y at the head of proj_ws
  if (nl && w_out)
    CHS_TRY(zgemm(ctx, 'N', 'N', H.n_proj, n_b, n_b, one, (const z64*)im->proj_ws.p, H.n_proj, Hs,
                  n_b, zero, w_out, H.n_proj));
  if (eps_dev_out)
    CHS_HIP(ctx, hipMemcpyAsync(eps_dev_out, w, n_b * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream));
  if (eps_host) {
    CHS_HIP(ctx, hipMemcpyAsync(eps_host, w, n_b * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->d2h += n_b * sizeof(double);
  }
  return CHS_OK;
}

// Y <- Y L^{-H} with S_sub = Y^H S Y = L L^H.
int cholqr_impl(chs_context* ctx, z64* Y, int n_b) {
  CHS_TRY(ensure_precision_ready(ctx, CHS_C128));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const int n_pw = H.n_pw;
  const size_t ne = (size_t)n_b * n_pw, nn = (size_t)n_b * n_b;
  CHS_TRY(ensure(ctx, im->ws5, ne * sizeof(z64)));
  z64* SY = (z64*)im->ws5.p;
  const bool gen = H.has_s;
  if (gen) {
    CHS_TRY(ensure(ctx, im->ws4, ne * sizeof(z64)));
    CHS_TRY(apply_h_impl(ctx, CHS_C128, Y, n_b, im->ws4.p, SY));
  }
  CHS_TRY(ensure(ctx, im->scf_ws, 2 * nn * sizeof(z64)));
  z64* Ss = (z64*)im->scf_ws.p;
  z64* Li = Ss + nn;
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  CHS_TRY(zgemm_herm(ctx, n_b, n_pw, Y, n_pw, gen ? SY : Y, n_pw, Ss, n_b));
  CHS_TRY(hermitise(ctx, Ss, n_b));
  int info = 0;
  CHS_TRY(cholesky(ctx, Ss, n_b, &info));
  if (info) {
    ctx->diag.solver_info = info;
    return set_error(ctx, CHS_ERR_RAYLEIGH_RITZ_FAILED, "Cholesky-QR failed at column %d", info);
  }
  CHS_TRY(tri_inv(ctx, Ss, Li, n_b));
  CHS_TRY(zgemm(ctx, 'N', 'C', n_pw, n_b, n_b, one, Y, n_pw, Li, n_b, zero, SY, n_pw));
  CHS_HIP(ctx, hipMemcpyAsync(Y, SY, ne * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_zhegvd(chs_context* ctx, int n, double* d_a, double* d_b, double* w) {
  if (!ctx || n <= 0 || !d_a || !w) return CHS_ERR_INVALID_ARG;
  if (n > 1024) return set_error(ctx, CHS_ERR_INVALID_ARG, "zhegvd limited to n <= 1024");
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->ws6, (size_t)n * sizeof(double) + 64));
  double* dw = (double*)im->ws6.p;
  CHS_TRY(zhegv(ctx, (z64*)d_a, (z64*)d_b, n, dw));
  CHS_HIP(ctx, hipMemcpyAsync(w, dw, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += n * sizeof(double);
  return CHS_OK;
}

int chs_rayleigh_ritz(chs_context* ctx, double* d_y, int n_b, double* eps, double* d_hy) {
  if (!ctx || !d_y || n_b <= 0 || !eps) return CHS_ERR_INVALID_ARG;
  if (n_b > 1024) return set_error(ctx, CHS_ERR_INVALID_ARG, "n_bands limited to 1024");
  return rayleigh_ritz_impl(ctx, (z64*)d_y, n_b, eps, (z64*)d_hy, nullptr);
}

int chs_cholqr(chs_context* ctx, double* d_y, int n_b) {
  if (!ctx || !d_y || n_b <= 0) return CHS_ERR_INVALID_ARG;
  return cholqr_impl(ctx, (z64*)d_y, n_b);
}

}  // extern "C"
