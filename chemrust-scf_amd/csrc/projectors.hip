// Nonlocal (USPP) projector terms as complex GEMMs.
//
// Reference: apply_v_nl_hamiltonian (src/eigensolver/chebyshev.rs:416-500)
// runs, per ion, three skinny ZGEMMs (beta^H psi, D C, beta C) with two
// allocations each; the gate tests' operator form adds S = 1 + beta Q
// beta^H (tests/phase0_gate1_consistency.rs:320-385). Here all ions'
// projectors form one (n_pw x P) block and each term is one GEMM pair over
// all P projectors and all bands, with the diagonal D (or Q) applied as a
// row scale between them:
//   W = beta^H psi        (P x n_b)
//   H psi += beta (D W),  S psi = psi + beta (Q W)
// and S^{-1} x = x - beta M beta^H x with M = (Q^{-1} + beta^H beta)^{-1}
// (Woodbury; the same identity as uspp_preconditioner.rs:80-131).
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"
#include "projectors.hpp"

namespace chs {

template <typename T>
__global__ void k_row_scale2(const cplx<T>* __restrict__ w, int P, int n_b,
                             const double* __restrict__ d, const double* __restrict__ q,
                             cplx<T>* __restrict__ dw, cplx<T>* __restrict__ qw) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)P * n_b) return;
  const int p = (int)(i % P);
  const cplx<T> v = w[i];
  const T dp = (T)d[p];
  dw[i] = mk(dp * v.x, dp * v.y);
  if (qw) {
    const T qp = (T)q[p];
    qw[i] = mk(qp * v.x, qp * v.y);
  }
}

__global__ void k_diag_add_inv(z64* __restrict__ a, const double* __restrict__ q, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) a[(size_t)i + (size_t)i * P].x += 1.0 / q[i];
}

template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mk((decltype(b[i].x))a[i].x, (decltype(b[i].x))a[i].y);
}

static int ensure_c64_projectors(chs_context* ctx) {
  Hamiltonian& H = ctx->impl->ham;
  const size_t nb = (size_t)H.n_proj * H.n_pw;
  if (!H.beta_f.p || H.beta_f.bytes < nb * sizeof(c32)) {
    CHS_TRY(ensure(ctx, H.beta_f, nb * sizeof(c32)));
    hipLaunchKernelGGL((k_convert<z64, c32>), dim3(ceil_div(nb, 256)), dim3(256), 0, ctx->stream,
                       (const z64*)H.beta, (c32*)H.beta_f.p, nb);
    CHS_LAUNCH_CHECK(ctx);
    const size_t pp = (size_t)H.n_proj * H.n_proj;
    CHS_TRY(ensure(ctx, H.bmat_f, pp * sizeof(c32)));
    hipLaunchKernelGGL((k_convert<z64, c32>), dim3(ceil_div(pp, 256)), dim3(256), 0, ctx->stream,
                       (const z64*)H.bmat.p, (c32*)H.bmat_f.p, pp);
    if (H.has_s) {
      CHS_TRY(ensure(ctx, H.sinv_m_f, pp * sizeof(c32)));
      hipLaunchKernelGGL((k_convert<z64, c32>), dim3(ceil_div(pp, 256)), dim3(256), 0,
                         ctx->stream, (const z64*)H.sinv_m.p, (c32*)H.sinv_m_f.p, pp);
    }
    CHS_LAUNCH_CHECK(ctx);
  }
  if (ctx->impl->use_h3 && !H.h3_cn.buf.p) {
    CHS_TRY(h3_split_planes(ctx, (const z64*)H.beta, H.n_pw, H.n_proj, H.n_pw, true, H.h3_cn));
    CHS_TRY(h3_split_planes(ctx, (const z64*)H.beta, H.n_pw, H.n_proj, H.n_pw, false, H.h3_nn));
  }
  return CHS_OK;
}

template <typename T>
static int gemm_any(chs_context* ctx, char oa, char ob, int m, int n, int k, double al,
                    const cplx<T>* A, int lda, const cplx<T>* B, int ldb, double be, cplx<T>* C,
                    int ldc);
template <>
int gemm_any<double>(chs_context* ctx, char oa, char ob, int m, int n, int k, double al,
                     const z64* A, int lda, const z64* B, int ldb, double be, z64* C, int ldc) {
  return zgemm(ctx, oa, ob, m, n, k, mk(al, 0.0), A, lda, B, ldb, mk(be, 0.0), C, ldc);
}
template <>
int gemm_any<float>(chs_context* ctx, char oa, char ob, int m, int n, int k, double al,
                    const c32* A, int lda, const c32* B, int ldb, double be, c32* C, int ldc) {
  return cgemm(ctx, oa, ob, m, n, k, mk((float)al, 0.f), A, lda, B, ldb, mk((float)be, 0.f), C,
               ldc);
}

template <typename T>
static int nonlocal_t(chs_context* ctx, const cplx<T>* psi, int n_b, cplx<T>* hpsi,
                      cplx<T>* spsi) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  const cplx<T>* beta =
      sizeof(T) == 8 ? (const cplx<T>*)H.beta : (const cplx<T>*)H.beta_f.p;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(cplx<T>)));
  cplx<T>* W = (cplx<T>*)ctx->impl->proj_ws.p;
  cplx<T>* DW = W + pb;
  cplx<T>* QW = DW + pb;
  CHS_TRY(gemm_any<T>(ctx, 'C', 'N', P, n_b, n_pw, 1.0, beta, n_pw, psi, n_pw, 0.0, W, P));
  hipLaunchKernelGGL((k_row_scale2<T>), dim3(ceil_div(pb, 256)), dim3(256), 0, ctx->stream, W, P,
                     n_b, H.dion, H.qion, DW, spsi ? QW : nullptr);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(gemm_any<T>(ctx, 'N', 'N', n_pw, n_b, P, 1.0, beta, n_pw, DW, P, 1.0, hpsi, n_pw));
  if (spsi) {
    if (spsi != psi)
      CHS_HIP(ctx, hipMemcpyAsync(spsi, psi, (size_t)n_b * n_pw * sizeof(cplx<T>),
                                  hipMemcpyDeviceToDevice, ctx->stream));
    if (H.has_s)
      CHS_TRY(gemm_any<T>(ctx, 'N', 'N', n_pw, n_b, P, 1.0, beta, n_pw, QW, P, 1.0, spsi, n_pw));
  }
  return CHS_OK;
}

// ---------------------------------------------------------------------------
// Projector part of one Chebyshev step in coefficient space. With
// W_k = beta^H X_k, the step X_{k+1} = 2 (S^{-1} H X_k - c X_k) / e - X_{k-1}
// splits as X_{k+1} = O + (2/e) beta C, where O is the local-potential step
// (the fused H|psi> pipeline's Chebyshev epilogue) and
//   C = D W_k - M (Z + B D W_k),   Z = beta^H H_loc X_k,
// B = beta^H beta, M = (Q^{-1} + B)^{-1} (C = D W_k when S = 1). Z follows
// from G = beta^H O: Z = (e/2)(G + W_{k-1}) + c W_k (first step:
// Z = e G + c W_0), and W_{k+1} = G + (2/e) B C. Two n_pw-long GEMMs per
// step (G and the beta C update) instead of four, and no H|psi> block.
// ---------------------------------------------------------------------------
// coef = {c, inv_e} on the device (graph-replayable): Z = a1 G + a2 W_{k-1}
// + c W with (a1, a2) = (e, 0) on the first step and (e/2, e/2) after;
// DW = D W. Without S the scaled coefficients f D W go straight to Cs.
// Residual-based filter: O carries the extra term s_b R_b (R the residual
// block, s_b = ress[b]), so G = beta^H O holds s_b beta^H R_b too; BR =
// beta^H R removes it before Z is formed.
template <typename T>
__global__ void k_cheb_coef(const cplx<T>* __restrict__ G, const cplx<T>* __restrict__ Wp,
                            const cplx<T>* __restrict__ W, const double* __restrict__ d,
                            const double* __restrict__ coef, int first, int P, size_t n,
                            cplx<T>* __restrict__ Z, cplx<T>* __restrict__ DW,
                            const cplx<T>* __restrict__ BR, const double* __restrict__ ress) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T c = (T)coef[0], e = (T)(1.0 / coef[1]);
  const cplx<T> w = W[i];
  const T dp = (T)d[(int)(i % P)];
  if (Z) {
    const T a1 = first ? e : (T)0.5 * e, a2 = first ? (T)0 : (T)0.5 * e;
    // W_{k-1} does not exist on the first step (not read)
    cplx<T> g = G[i];
    const cplx<T> wp = first ? mk((T)0, (T)0) : Wp[i];
    if (BR) {
      const T sb = (T)ress[i / P];
      const cplx<T> br = BR[i];
      g = mk(g.x - sb * br.x, g.y - sb * br.y);
    }
    Z[i] = mk(a1 * g.x + a2 * wp.x + c * w.x, a1 * g.y + a2 * wp.y + c * w.y);
    DW[i] = mk(dp * w.x, dp * w.y);
  } else {
    const T f = (T)(first ? coef[1] : 2.0 * coef[1]);
    DW[i] = mk(f * dp * w.x, f * dp * w.y);
  }
}

// Cs *= f with f = inv_e (first step) or 2 inv_e
template <typename T>
__global__ void k_cheb_scale(cplx<T>* __restrict__ x, const double* __restrict__ coef, int first,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T f = (T)(first ? coef[1] : 2.0 * coef[1]);
  x[i] = mk(f * x[i].x, f * x[i].y);
}

template <typename T>
static cplx<T>* nl_slot(Hamiltonian& H, int n_b, int slot) {
  return (cplx<T>*)H.nl_ws.p + (size_t)slot * H.n_proj * n_b;
}

template <typename T>
static int cheb_nl_begin_t(chs_context* ctx, const cplx<T>* X0, int n_b) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj;
  CHS_TRY(ensure(ctx, H.nl_ws, 5 * (size_t)P * n_b * sizeof(cplx<T>)));
  H.nl_wp = 0;
  H.nl_w = 1;
  H.nl_g = 2;
  const cplx<T>* beta = sizeof(T) == 8 ? (const cplx<T>*)H.beta : (const cplx<T>*)H.beta_f.p;
  if constexpr (sizeof(T) == 4) {
    if (ctx->impl->use_h3) {
      CHS_TRY(ensure(ctx, H.h3_exp, 2 * (size_t)n_b * sizeof(int)));
      int* eb = (int*)H.h3_exp.p;
      CHS_TRY(h3_column_exponents(ctx, X0, H.n_pw, n_b, H.n_pw, eb));
      return h3_gemm(ctx, H.h3_cn, n_b, mk(1.f, 0.f), X0, H.n_pw, eb, mk(0.f, 0.f),
                     nl_slot<T>(H, n_b, H.nl_w), P);
    }
  }
  return gemm_any<T>(ctx, 'C', 'N', P, n_b, H.n_pw, 1.0, beta, H.n_pw, X0, H.n_pw, 0.0,
                     nl_slot<T>(H, n_b, H.nl_w), P);
}

template <typename T>
static int cheb_nl_step_t(chs_context* ctx, cplx<T>* O, int n_b, bool first, const double* coef,
                          const cplx<T>* BR, const double* ress) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  const bool f32 = sizeof(T) == 4;
  const cplx<T>* beta = f32 ? (const cplx<T>*)H.beta_f.p : (const cplx<T>*)H.beta;
  const cplx<T>* B = f32 ? (const cplx<T>*)H.bmat_f.p : (const cplx<T>*)H.bmat.p;
  const cplx<T>* M = f32 ? (const cplx<T>*)H.sinv_m_f.p : (const cplx<T>*)H.sinv_m.p;
  cplx<T>* Wp = nl_slot<T>(H, n_b, H.nl_wp);
  cplx<T>* W = nl_slot<T>(H, n_b, H.nl_w);
  cplx<T>* G = nl_slot<T>(H, n_b, H.nl_g);
  cplx<T>* Z = nl_slot<T>(H, n_b, 3);
  cplx<T>* Cs = nl_slot<T>(H, n_b, 4);
  const size_t pb = (size_t)P * n_b;
  const unsigned gp = ceil_div(pb, 256);
  // complex64 with fp16x3: both n_pw-long products on the fp16 matrix cores
  bool h3 = false;
  int* eb = nullptr;
  if constexpr (sizeof(T) == 4) {
    h3 = ctx->impl->use_h3;
    if (h3) {
      CHS_TRY(ensure(ctx, H.h3_exp, 2 * (size_t)n_b * sizeof(int)));
      eb = (int*)H.h3_exp.p;
    }
  }
  // G = beta^H O
  if constexpr (sizeof(T) == 4) {
    if (h3) {
      CHS_TRY(h3_column_exponents(ctx, O, n_pw, n_b, n_pw, eb));
      CHS_TRY(h3_gemm(ctx, H.h3_cn, n_b, mk(1.f, 0.f), O, n_pw, eb, mk(0.f, 0.f), G, P));
    }
  }
  if (!h3) CHS_TRY(gemm_any<T>(ctx, 'C', 'N', P, n_b, n_pw, 1.0, beta, n_pw, O, n_pw, 0.0, G, P));
  hipLaunchKernelGGL((k_cheb_coef<T>), dim3(gp), dim3(256), 0, ctx->stream, G, Wp, W, H.dion,
                     coef, (int)first, P, pb, H.has_s ? Z : nullptr, Cs, BR, ress);
  CHS_LAUNCH_CHECK(ctx);
  if (H.has_s) {
    // Z += B (D W); C = D W - M Z; Cs = f C
    CHS_TRY(gemm_any<T>(ctx, 'N', 'N', P, n_b, P, 1.0, B, P, Cs, P, 1.0, Z, P));
    CHS_TRY(gemm_any<T>(ctx, 'N', 'N', P, n_b, P, -1.0, M, P, Z, P, 1.0, Cs, P));
    hipLaunchKernelGGL((k_cheb_scale<T>), dim3(gp), dim3(256), 0, ctx->stream, Cs, coef,
                       (int)first, pb);
    CHS_LAUNCH_CHECK(ctx);
  }
  // X_{k+1} = O + beta Cs ; W_{k+1} = G + B Cs (into G's slot)
  if constexpr (sizeof(T) == 4) {
    if (h3) {
      CHS_TRY(h3_column_exponents(ctx, Cs, P, n_b, P, eb + n_b));
      CHS_TRY(h3_gemm(ctx, H.h3_nn, n_b, mk(1.f, 0.f), Cs, P, eb + n_b, mk(1.f, 0.f), O, n_pw));
    }
  }
  if (!h3) CHS_TRY(gemm_any<T>(ctx, 'N', 'N', n_pw, n_b, P, 1.0, beta, n_pw, Cs, P, 1.0, O, n_pw));
  CHS_TRY(gemm_any<T>(ctx, 'N', 'N', P, n_b, P, 1.0, B, P, Cs, P, 1.0, G, P));
  const int old_wp = H.nl_wp;
  H.nl_wp = H.nl_w;
  H.nl_w = H.nl_g;
  H.nl_g = old_wp;
  return CHS_OK;
}

int cheb_nl_begin(chs_context* ctx, int precision, const void* X0, int n_b) {
  if (precision == CHS_C128) return cheb_nl_begin_t<double>(ctx, (const z64*)X0, n_b);
  CHS_TRY(ensure_c64_projectors(ctx));
  return cheb_nl_begin_t<float>(ctx, (const c32*)X0, n_b);
}

int cheb_nl_step(chs_context* ctx, int precision, void* O, int n_b, bool first,
                 const double* d_coef, const void* br, const double* ress) {
  if (precision == CHS_C128)
    return cheb_nl_step_t<double>(ctx, (z64*)O, n_b, first, d_coef, (const z64*)br, ress);
  return cheb_nl_step_t<float>(ctx, (c32*)O, n_b, first, d_coef, (const c32*)br, ress);
}

// Coefficients of the residual filter's first two iterates: W_0 = beta^H R_0
// = 0 and W_1 = beta^H R_1 = inv_e BR (BR = beta^H R, R the residual block).
template <typename T>
__global__ void k_res_coef_start(const cplx<T>* __restrict__ BR, const double* __restrict__ coef,
                                 size_t n, cplx<T>* __restrict__ W0, cplx<T>* __restrict__ W1) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T f = (T)coef[1];
  W0[i] = mk((T)0, (T)0);
  W1[i] = mk(f * BR[i].x, f * BR[i].y);
}

template <typename T>
static int cheb_nl_begin_res_t(chs_context* ctx, const cplx<T>* R, int n_b, const double* coef,
                               cplx<T>* BR) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj;
  CHS_TRY(ensure(ctx, H.nl_ws, 5 * (size_t)P * n_b * sizeof(cplx<T>)));
  H.nl_wp = 0;
  H.nl_w = 1;
  H.nl_g = 2;
  const cplx<T>* beta = sizeof(T) == 8 ? (const cplx<T>*)H.beta : (const cplx<T>*)H.beta_f.p;
  CHS_TRY(gemm_any<T>(ctx, 'C', 'N', P, n_b, H.n_pw, 1.0, beta, H.n_pw, R, H.n_pw, 0.0, BR, P));
  const size_t pb = (size_t)P * n_b;
  hipLaunchKernelGGL((k_res_coef_start<T>), dim3(ceil_div(pb, 256)), dim3(256), 0, ctx->stream,
                     (const cplx<T>*)BR, coef, pb, nl_slot<T>(H, n_b, H.nl_wp),
                     nl_slot<T>(H, n_b, H.nl_w));
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int cheb_nl_begin_res(chs_context* ctx, int precision, const void* R, int n_b,
                      const double* d_coef, void* br) {
  if (precision == CHS_C128)
    return cheb_nl_begin_res_t<double>(ctx, (const z64*)R, n_b, d_coef, (z64*)br);
  CHS_TRY(ensure_c64_projectors(ctx));
  return cheb_nl_begin_res_t<float>(ctx, (const c32*)R, n_b, d_coef, (c32*)br);
}

// out += beta c with c = D w - M (g + B D w) (S present) or c = D w, where
// w = beta^H X (wx, or formed here when null) and g = beta^H out, out holding
// H_loc X on entry: out becomes S^{-1} H X (FP64). Three (two with wx)
// n_pw-long products.
int sinv_h_nonlocal(chs_context* ctx, const z64* X, int n_b, const z64* wx, z64* out) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  if (P == 0) return CHS_OK;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 4 * pb * sizeof(z64)));
  z64* w = (z64*)ctx->impl->proj_ws.p;
  z64* g = w + pb;
  z64* dw = g + pb;
  z64* qw = dw + pb;
  const z64* beta = (const z64*)H.beta;
  if (wx)
    CHS_HIP(ctx, hipMemcpyAsync(w, wx, pb * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  else
    CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0, 0.0), beta, n_pw, X, n_pw, mk(0.0, 0.0), w,
                  P));
  hipLaunchKernelGGL((k_row_scale2<double>), dim3(ceil_div(pb, 256)), dim3(256), 0, ctx->stream,
                     w, P, n_b, H.dion, H.qion, dw, (z64*)nullptr);
  CHS_LAUNCH_CHECK(ctx);
  if (H.has_s) {
    // g = beta^H H_loc X + B D w;  c = D w - M g  (into dw)
    CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0, 0.0), beta, n_pw, out, n_pw, mk(0.0, 0.0), g,
                  P));
    CHS_TRY(zgemm(ctx, 'N', 'N', P, n_b, P, mk(1.0, 0.0), (const z64*)H.bmat.p, P, dw, P,
                  mk(1.0, 0.0), g, P));
    CHS_TRY(zgemm(ctx, 'N', 'N', P, n_b, P, mk(-1.0, 0.0), (const z64*)H.sinv_m.p, P, g, P,
                  mk(1.0, 0.0), dw, P));
  }
  (void)qw;
  return zgemm(ctx, 'N', 'N', n_pw, n_b, P, mk(1.0, 0.0), beta, n_pw, dw, P, mk(1.0, 0.0), out,
               n_pw);
}

// Projector parts of the Rayleigh-Ritz subspace matrices of a c128 block Y:
// with W = beta^H Y (P x n_b), Hs += W^H D W and, when Ss is given,
// Ss += W^H Q W. One n_pw-long GEMM instead of the two (beta D W, beta Q W)
// a full H|Y>, S|Y> would need.
int subspace_nl(chs_context* ctx, const z64* Y, int n_b, z64* Hs, z64* Ss, const c32* Y32) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(z64)));
  z64* W = (z64*)ctx->impl->proj_ws.p;
  z64* DW = W + pb;
  z64* QW = DW + pb;
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  if (Y32) {
    CHS_TRY(ensure_c64_projectors(ctx));
    CHS_TRY(ensure(ctx, ctx->impl->rr32, pb * sizeof(c32)));
    c32* W32 = (c32*)ctx->impl->rr32.p;
    CHS_TRY(cgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0f, 0.0f), (const c32*)H.beta_f.p, n_pw, Y32,
                  n_pw, mk(0.0f, 0.0f), W32, P));
    hipLaunchKernelGGL((k_convert<c32, z64>), dim3(ceil_div(pb, 256)), dim3(256), 0, ctx->stream,
                       W32, W, pb);
    CHS_LAUNCH_CHECK(ctx);
  } else {
    CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, one, (const z64*)H.beta, n_pw, Y, n_pw, zero, W, P));
  }
  hipLaunchKernelGGL((k_row_scale2<double>), dim3(ceil_div(pb, 256)), dim3(256), 0, ctx->stream, W,
                     P, n_b, H.dion, H.qion, DW, Ss ? QW : nullptr);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(zgemm(ctx, 'C', 'N', n_b, n_b, P, one, W, P, DW, P, one, Hs, n_b));
  if (Ss) CHS_TRY(zgemm(ctx, 'C', 'N', n_b, n_b, P, one, W, P, QW, P, one, Ss, n_b));
  return CHS_OK;
}

int nonlocal_apply(chs_context* ctx, int precision, const void* psi, int n_b, void* hpsi,
                   void* spsi) {
  Hamiltonian& H = ctx->impl->ham;
  if (H.n_proj == 0) return CHS_OK;
  if (precision == CHS_C128)
    return nonlocal_t<double>(ctx, (const z64*)psi, n_b, (z64*)hpsi, (z64*)spsi);
  CHS_TRY(ensure_c64_projectors(ctx));
  return nonlocal_t<float>(ctx, (const c32*)psi, n_b, (c32*)hpsi, (c32*)spsi);
}

template <typename T>
static int sinv_t(chs_context* ctx, cplx<T>* x, int n_b) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  const cplx<T>* beta =
      sizeof(T) == 8 ? (const cplx<T>*)H.beta : (const cplx<T>*)H.beta_f.p;
  const cplx<T>* M =
      sizeof(T) == 8 ? (const cplx<T>*)H.sinv_m.p : (const cplx<T>*)H.sinv_m_f.p;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(cplx<T>)));
  cplx<T>* W = (cplx<T>*)ctx->impl->proj_ws.p;
  cplx<T>* MW = W + pb;
  CHS_TRY(gemm_any<T>(ctx, 'C', 'N', P, n_b, n_pw, 1.0, beta, n_pw, x, n_pw, 0.0, W, P));
  CHS_TRY(gemm_any<T>(ctx, 'N', 'N', P, n_b, P, 1.0, M, P, W, P, 0.0, MW, P));
  CHS_TRY(gemm_any<T>(ctx, 'N', 'N', n_pw, n_b, P, -1.0, beta, n_pw, MW, P, 1.0, x, n_pw));
  return CHS_OK;
}

int sinv_apply(chs_context* ctx, int precision, void* x, int n_b) {
  Hamiltonian& H = ctx->impl->ham;
  if (H.n_proj == 0 || !H.has_s) return CHS_OK;
  if (precision == CHS_C128) return sinv_t<double>(ctx, (z64*)x, n_b);
  CHS_TRY(ensure_c64_projectors(ctx));
  return sinv_t<float>(ctx, (c32*)x, n_b);
}

// ---------------------------------------------------------------------------
// Synthetic projectors, augmentation and screening (SCF setup / glue)
// ---------------------------------------------------------------------------

// beta_{a,j}(k+G) = Omega^{-1/2} f_j(k+G) e^{-i(k+G).R_a}, j = (l=0), (l=1; x, y, z):
// f_0 = Y00 pi^{3/2} s^3 exp(-G^2 s^2/4); f_x = sqrt(3/4pi) (-i s^2 G_x/2) pi^{3/2} s^3 exp(..)
__global__ void k_build_beta(const int32_t* __restrict__ miller, int n_pw, double b, double kf0,
                             double kf1, double kf2, const double* __restrict__ pos, int n_atoms,
                             int np, double s, double omega, z64* __restrict__ beta) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n_pw * n_atoms) return;
  const int g = (int)(i % n_pw), a = (int)(i / n_pw);
  const double gx = b * (miller[3 * g] + kf0), gy = b * (miller[3 * g + 1] + kf1),
               gz = b * (miller[3 * g + 2] + kf2);
  const double g2 = gx * gx + gy * gy + gz * gz;
  const double pref = pow(M_PI, 1.5) * s * s * s / sqrt(omega);
  const double rad = pref * exp(-g2 * s * s / 4.0);
  const double t = -(gx * pos[3 * a] + gy * pos[3 * a + 1] + gz * pos[3 * a + 2]);
  double sn, cs;
  sincos(t, &sn, &cs);
  const double y00 = 1.0 / sqrt(4.0 * M_PI), y1 = sqrt(3.0 / (4.0 * M_PI));
  const size_t base = (size_t)a * np * n_pw + g;
  beta[base] = mk(y00 * rad * cs, y00 * rad * sn);
  if (np > 1) {
    // y1 rad g_a (-i s^2/2) (cs + i sn) = y1 rad g_a (s^2/2) (sn - i cs)
    const double h = 0.5 * s * s * y1 * rad;
    const double ga[3] = {gx, gy, gz};
    for (int j = 0; j < 3; ++j)
      beta[base + (size_t)(1 + j) * n_pw] = mk(h * ga[j] * sn, -h * ga[j] * cs);
  }
}

// rho_p = sum_b occ_b |W[p, b]|^2 (W = beta^H psi, P x n_b)
__global__ void k_proj_occ(const z64* __restrict__ w, int P, int n_b, const double* __restrict__ occ,
                           double* __restrict__ rho_p) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double s = 0;
  for (int b = 0; b < n_b; ++b) s += occ[b] * norm2(w[(size_t)p + (size_t)b * P]);
  rho_p[p] = s;
}

// aug(G) = Omega^-1 exp(-G^2 rq^2/4) sum_a ra[a] e^{-iG.R_a} (Nyquist zero)
__global__ void k_aug_g(z64* __restrict__ g, int n, int n_atoms, const z64* __restrict__ ph,
                        const double* __restrict__ ra, double b, double rq, double omega) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n3 = (size_t)n * n * n;
  if (id >= n3) return;
  const int ix = (int)(id % n), iy = (int)((id / n) % n), iz = (int)(id / ((size_t)n * n));
  if ((n % 2 == 0) && (ix == n / 2 || iy == n / 2 || iz == n / 2)) {
    g[id] = mk(0.0, 0.0);
    return;
  }
  const int h = ix <= n / 2 ? ix : ix - n, k = iy <= n / 2 ? iy : iy - n, l = iz <= n / 2 ? iz : iz - n;
  const double g2 = b * b * (double)(h * h + k * k + l * l);
  z64 s = mk(0.0, 0.0);
  for (int a = 0; a < n_atoms; ++a) {
    const z64* pa = ph + (size_t)a * 3 * n;
    const z64 e = pa[ix] * pa[n + iy] * pa[2 * n + iz];
    s = s + ra[a] * e;
  }
  const double f = exp(-g2 * rq * rq / 4.0) / omega;
  g[id] = mk(f * s.x, f * s.y);
}

__global__ void k_real_to_complex_proj(const double* __restrict__ r, z64* __restrict__ g, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = mk(r[i], 0.0);
}

__global__ void k_add_real(const z64* __restrict__ g, double* __restrict__ r, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] += g[i].x;
}

// part[a * nchunk + c] = sum over chunk c of Re(V(G) e^{+iG.R_a}) exp(-G^2 rq^2/4)
// with the Nyquist planes of an even grid left out. Both factors separate
// over the axes (G^2 = b^2 (h^2 + k^2 + l^2), e^{iG.R} = px py pz), so a
// block tables conj(px) gx, py and pz for kScreenAtoms atoms in LDS (gx with
// its Nyquist entry zero drops those planes) and reads each V(G) once for
// all of them. A chunk is the (iz, iy) rows c, c + nchunk, ... of n points.
constexpr int kScreenAtoms = 8;

__global__ void __launch_bounds__(256) k_screen(const z64* __restrict__ vg, int n, int n_atoms,
                                                const z64* __restrict__ ph, double b, double rq,
                                                double* __restrict__ part) {
  extern __shared__ z64 tab[];  // cx, py, pz: [kScreenAtoms][n] each, then gauss[n]
  __shared__ double sh[8];
  z64* cx = tab;
  z64* py = tab + kScreenAtoms * n;
  z64* pz = tab + 2 * kScreenAtoms * n;
  double* gs = (double*)(tab + 3 * kScreenAtoms * n);
  const int c = blockIdx.x, nchunk = gridDim.x, a0 = blockIdx.y * kScreenAtoms;
  const int na = min(kScreenAtoms, n_atoms - a0);
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int h = t <= n / 2 ? t : t - n;
    gs[t] = (n % 2 == 0 && t == n / 2) ? 0.0 : exp(-b * b * (double)(h * h) * rq * rq / 4.0);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kScreenAtoms * n; t += blockDim.x) {
    const int q = t / n, i = t % n;
    z64 x = mk(0.0, 0.0), y = mk(0.0, 0.0), z = mk(0.0, 0.0);
    if (q < na) {
      const z64* pa = ph + (size_t)(a0 + q) * 3 * n;
      x = conj(pa[i]);
      x = mk(x.x * gs[i], x.y * gs[i]);
      y = pa[n + i];
      z = pa[2 * n + i];
    }
    cx[t] = x;
    py[t] = y;
    pz[t] = z;
  }
  __syncthreads();
  const int rpp = blockDim.x / n;  // rows per pass (n <= blockDim.x)
  const int ix = threadIdx.x % n, roff = threadIdx.x / n;
  double s[kScreenAtoms];
#pragma unroll
  for (int q = 0; q < kScreenAtoms; ++q) s[q] = 0.0;
  if (roff < rpp) {
    for (int r = c * rpp + roff; r < n * n; r += nchunk * rpp) {
      const int iz = r / n, iy = r % n;
      const double gyz = gs[iy] * gs[iz];
      if (gyz == 0.0) continue;
      const z64 v = vg[(size_t)r * n + ix];
#pragma unroll
      for (int q = 0; q < kScreenAtoms; ++q) {
        // Re(v conj(px py pz)) gx gy gz
        const z64 yz = conj(py[q * n + iy] * pz[q * n + iz]);
        const z64 t = cx[q * n + ix] * yz;
        s[q] += gyz * (v.x * t.x - v.y * t.y);
      }
    }
  }
  for (int q = 0; q < na; ++q) {
    const double r = block_sum(s[q], sh);
    if (threadIdx.x == 0) part[(size_t)(a0 + q) * nchunk + c] = r;
  }
}

__global__ void k_screen_finish(const double* __restrict__ part, int nchunk, int n_atoms, int np,
                                const double* __restrict__ d0, double q0, double inv_n3,
                                double* __restrict__ d) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  double s = 0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)a * nchunk + c];
  for (int j = 0; j < np; ++j) d[a * np + j] = d0[a * np + j] + q0 * s * inv_n3;
}

int build_beta(chs_context* ctx, const int32_t* miller, int n_pw, double L, const double* kf,
               const double* d_pos, int n_atoms, int np, double s, z64* beta) {
  const double b = 2.0 * M_PI / L;
  hipLaunchKernelGGL(k_build_beta, dim3(ceil_div((size_t)n_pw * n_atoms, 256)), dim3(256), 0,
                     ctx->stream, miller, n_pw, b, kf[0], kf[1], kf[2], d_pos, n_atoms, np, s,
                     L * L * L, beta);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int augment_density(chs_context* ctx, const z64* psi, int n_b, const double* d_occ, int n_atoms,
                    int np, double q0, double rq, const z64* d_ph, int n, double L, z64* cgrid,
                    double* rho, const z64* w) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(z64) + (P + n_atoms) * sizeof(double)));
  z64* W = (z64*)ctx->impl->proj_ws.p;
  double* rho_p = (double*)(W + 3 * pb);
  double* ra = rho_p + P;
  if (!w) {
    CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0, 0.0), (const z64*)H.beta, n_pw, psi, n_pw,
                  mk(0.0, 0.0), W, P));
    w = W;
  }
  hipLaunchKernelGGL(k_proj_occ, dim3(ceil_div(P, 256)), dim3(256), 0, ctx->stream, w, P, n_b,
                     d_occ, rho_p);
  CHS_LAUNCH_CHECK(ctx);
  // ra[a] = q0 sum_j rho_{a,j}
  std::vector<double> hp(P);
  CHS_HIP(ctx, hipMemcpyAsync(hp.data(), rho_p, P * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += P * sizeof(double);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<double> hra(n_atoms, 0.0);
  for (int a = 0; a < n_atoms; ++a)
    for (int j = 0; j < np; ++j) hra[a] += q0 * hp[a * np + j];
  CHS_HIP(ctx, hipMemcpyAsync(ra, hra.data(), n_atoms * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += n_atoms * sizeof(double);
  const size_t n3 = (size_t)n * n * n;
  hipLaunchKernelGGL(k_aug_g, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, cgrid, n,
                     n_atoms, d_ph, ra, 2.0 * M_PI / L, rq, L * L * L);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid, n, 1, +1));
  hipLaunchKernelGGL(k_add_real, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, cgrid, rho,
                     n3);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int screen_dion(chs_context* ctx, const double* veff, int n, double L, int n_atoms, int np,
                double q0, double rq, const z64* d_ph, const double* d_d0, z64* cgrid,
                double* d_out) {
  const size_t n3 = (size_t)n * n * n;
  const int nchunk = 64;
  if (n > 256) return set_error(ctx, CHS_ERR_INVALID_ARG, "augmentation screening limited to n <= 256");
  CHS_TRY(ensure(ctx, ctx->impl->part, sizeof(double) * ((size_t)n_atoms * nchunk + 2 * kRedBlocks)));
  hipLaunchKernelGGL(k_real_to_complex_proj, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                     veff, cgrid, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid, n, 1, -1));
  const size_t tab = (3 * (size_t)kScreenAtoms * n) * sizeof(z64) + n * sizeof(double);
  hipLaunchKernelGGL(k_screen, dim3(nchunk, ceil_div(n_atoms, kScreenAtoms)), dim3(256), tab,
                     ctx->stream, cgrid, n, n_atoms, d_ph, 2.0 * M_PI / L, rq,
                     (double*)ctx->impl->part.p);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_screen_finish, dim3(ceil_div(n_atoms, 256)), dim3(256), 0, ctx->stream,
                     (const double*)ctx->impl->part.p, nchunk, n_atoms, np, d_d0, q0,
                     1.0 / (double)n3, d_out);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

// Synthetic operand values in [-0.005, 0.005) for timing.
__global__ void k_proj_fill(float* __restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    a[i] = ((float)(h & 0xFFFF) / 65536.0f - 0.5f) * 1e-2f;
  }
}

// One filter step's two n_pw-long projector products, G = beta^H O and
// O + beta C, in the form the filter uses at `precision` (complex64: FP32
// matrix cores, or fp16x3 with its exponent kernels when enabled).
template <typename T>
static int projector_products(chs_context* ctx, cplx<T>* O, cplx<T>* Cs, cplx<T>* G, int n_b,
                              int which) {
  Hamiltonian& H = ctx->impl->ham;
  const int P = H.n_proj, n_pw = H.n_pw;
  if constexpr (sizeof(T) == 4) {
    if (ctx->impl->use_h3) {
      CHS_TRY(ensure(ctx, H.h3_exp, 2 * (size_t)n_b * sizeof(int)));
      int* eb = (int*)H.h3_exp.p;
      if (which == 0) {
        CHS_TRY(h3_column_exponents(ctx, O, n_pw, n_b, n_pw, eb));
        return h3_gemm(ctx, H.h3_cn, n_b, mk(1.f, 0.f), O, n_pw, eb, mk(0.f, 0.f), G, P);
      }
      CHS_TRY(h3_column_exponents(ctx, Cs, P, n_b, P, eb + n_b));
      return h3_gemm(ctx, H.h3_nn, n_b, mk(1.f, 0.f), Cs, P, eb + n_b, mk(1.f, 0.f), O, n_pw);
    }
  }
  const cplx<T>* beta = sizeof(T) == 4 ? (const cplx<T>*)H.beta_f.p : (const cplx<T>*)H.beta;
  if (which == 0)
    return gemm_any<T>(ctx, 'C', 'N', P, n_b, n_pw, 1.0, beta, n_pw, O, n_pw, 0.0, G, P);
  return gemm_any<T>(ctx, 'N', 'N', n_pw, n_b, P, 1.0, beta, n_pw, Cs, P, 1.0, O, n_pw);
}

extern "C" {

int chs_profile_projector_products(chs_context* ctx, int precision, int n_b, int reps,
                                   double* ms2) {
  if (!ctx || n_b <= 0 || reps <= 0 || !ms2 || (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  Hamiltonian& H = im->ham;
  if (!H.bound || H.n_proj == 0) return set_error(ctx, CHS_ERR_INVALID_ARG, "no projectors bound");
  if (precision == CHS_C64) CHS_TRY(ensure_c64_projectors(ctx));
  const size_t cs = precision == CHS_C128 ? 16 : 8;
  const size_t no = (size_t)H.n_pw * n_b, nc = (size_t)H.n_proj * n_b;
  CHS_TRY(ensure(ctx, im->ws1, no * cs));
  CHS_TRY(ensure(ctx, im->ws2, nc * cs));
  CHS_TRY(ensure(ctx, im->ws3, nc * cs));
  for (auto pr : {std::make_pair(im->ws1.p, no), std::make_pair(im->ws2.p, nc)}) {
    const size_t nf = pr.second * cs / sizeof(float);
    hipLaunchKernelGGL(k_proj_fill, dim3(ceil_div(nf, 256)), dim3(256), 0, ctx->stream,
                       (float*)pr.first, nf);
  }
  CHS_LAUNCH_CHECK(ctx);
  for (int which = 0; which < 2; ++which) {
    auto run = [&]() -> int {
      if (precision == CHS_C128)
        return projector_products<double>(ctx, (z64*)im->ws1.p, (z64*)im->ws2.p, (z64*)im->ws3.p,
                                          n_b, which);
      return projector_products<float>(ctx, (c32*)im->ws1.p, (c32*)im->ws2.p, (c32*)im->ws3.p,
                                       n_b, which);
    };
    CHS_TRY(run());  // warm-up: tuning and workspace growth
    CHS_HIP(ctx, hipEventRecord(im->ev0, ctx->stream));
    for (int r = 0; r < reps; ++r) CHS_TRY(run());
    CHS_HIP(ctx, hipEventRecord(im->ev1, ctx->stream));
    CHS_HIP(ctx, hipEventSynchronize(im->ev1));
    float t = 0;
    CHS_HIP(ctx, hipEventElapsedTime(&t, im->ev0, im->ev1));
    ms2[which] = t / reps;
  }
  return CHS_OK;
}

int chs_bind_projectors(chs_context* ctx, int n_proj, const double* d_beta, const double* d_d,
                        const double* d_q) {
  if (!ctx || n_proj < 0 || (n_proj > 0 && (!d_beta || !d_d || !d_q))) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "bind the Hamiltonian first");
  H.n_proj = n_proj;
  H.beta = d_beta;
  H.dion = d_d;
  H.qion = d_q;
  H.has_s = false;
  release(H.beta_f);
  release(H.sinv_m_f);
  release(H.bmat_f);
  release(H.h3_cn.buf);
  release(H.h3_nn.buf);
  if (n_proj == 0) return CHS_OK;
  std::vector<double> q(n_proj);
  CHS_HIP(ctx, hipMemcpyAsync(q.data(), d_q, n_proj * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += n_proj * sizeof(double);
  bool any = false, all = true;
  for (double v : q) {
    any = any || v != 0.0;
    all = all && v != 0.0;
  }
  if (any && !all)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "Q must be zero for all or no projectors");
  // B = beta^H beta (the filter's coefficient recurrences use it) and, with
  // Q != 0, M = (Q^{-1} + B)^{-1} through Cholesky
  const int P = n_proj;
  const size_t pp = (size_t)P * P;
  CHS_TRY(ensure(ctx, H.bmat, pp * sizeof(z64)));
  CHS_TRY(zgemm(ctx, 'C', 'N', P, P, H.n_pw, mk(1.0, 0.0), (const z64*)d_beta, H.n_pw,
                (const z64*)d_beta, H.n_pw, mk(0.0, 0.0), (z64*)H.bmat.p, P));
  CHS_TRY(hermitise(ctx, (z64*)H.bmat.p, P));
  if (!any) {
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CHS_OK;
  }
  CHS_TRY(ensure(ctx, H.sinv_m, pp * sizeof(z64)));
  CHS_TRY(ensure(ctx, ctx->impl->lin_ws, 2 * pp * sizeof(z64)));
  z64* A = (z64*)ctx->impl->lin_ws.p;
  z64* Li = A + pp;
  CHS_HIP(ctx, hipMemcpyAsync(A, H.bmat.p, pp * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  hipLaunchKernelGGL(k_diag_add_inv, dim3(ceil_div(P, 256)), dim3(256), 0, ctx->stream, A, d_q, P);
  CHS_LAUNCH_CHECK(ctx);
  int info = 0;
  CHS_TRY(cholesky(ctx, A, P, &info));
  if (info)
    return set_error(ctx, CHS_ERR_SOLVER, "Q^-1 + beta^H beta is not positive definite (column %d)",
                     info);
  CHS_TRY(tri_inv(ctx, A, Li, P));
  CHS_TRY(zgemm(ctx, 'C', 'N', P, P, P, mk(1.0, 0.0), Li, P, Li, P, mk(0.0, 0.0),
                (z64*)H.sinv_m.p, P));
  H.has_s = true;
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

int chs_set_dion(chs_context* ctx, const double* d_d) {
  if (!ctx || !d_d) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (H.n_proj == 0) return set_error(ctx, CHS_ERR_INVALID_ARG, "no projectors bound");
  H.dion = d_d;
  return CHS_OK;
}

}  // extern "C"
