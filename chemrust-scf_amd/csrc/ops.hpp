// Internal operator entry points shared across translation units.
#pragma once
#include "impl.hpp"

namespace chs {

int ensure_precision_ready(chs_context* ctx, int precision);
int apply_h_impl(chs_context* ctx, int precision, const void* psi, int n_b, void* hpsi, void* spsi);
int chebyshev_filter_impl(chs_context* ctx, int precision, const void* psi, int n_b, int degree,
                          const double* b4, void* out);
int filter_steps(chs_context* ctx, int precision, void* a, void* b, int n_b, int degree,
                 const double* d_coef, double* flag, void** res);
int filter_check(chs_context* ctx, const double* flag);
int lanczos_impl(chs_context* ctx, int k, uint64_t seed, double* upper, double* ritz_min,
                 bool beta_c64 = false);
// Nonlocal projector terms (projectors.hip): hpsi += beta D beta^H psi and,
// when spsi != nullptr, spsi = psi + beta Q beta^H psi.
int nonlocal_apply(chs_context* ctx, int precision, const void* psi, int n_b, void* hpsi,
                   void* spsi);
// Hs += W^H D W, Ss += W^H Q W (Ss may be null) with W = beta^H Y (c128);
// with Y32 (Y held exactly in complex64) W is formed in complex64 and widened.
int subspace_nl(chs_context* ctx, const z64* Y, int n_b, z64* Hs, z64* Ss,
                const c32* Y32 = nullptr);
// x <- S^{-1} x = x - beta M beta^H x (Woodbury).
int sinv_apply(chs_context* ctx, int precision, void* x, int n_b);
// Chebyshev filter with projectors, in projector-coefficient space:
// cheb_nl_begin sets W_0 = beta^H X_0; cheb_nl_step turns the local-potential
// step O (in place) into X_{k+1} and advances W.
int cheb_nl_begin(chs_context* ctx, int precision, const void* X0, int n_b);
// d_coef: device {c, inv_e} of the filter window. With br (beta^H of the
// residual block, filter precision) and ress (device per-band scales) the
// step is the residual filter's: O carries + ress[b] R_b.
int cheb_nl_step(chs_context* ctx, int precision, void* O, int n_b, bool first,
                 const double* d_coef, const void* br = nullptr, const double* ress = nullptr);
// Residual filter start: br = beta^H R, W_0 = 0, W_1 = inv_e br.
int cheb_nl_begin_res(chs_context* ctx, int precision, const void* R, int n_b,
                      const double* d_coef, void* br);
// out (H_loc X on entry, c128) <- S^{-1} H X; wx = beta^H X or null.
int sinv_h_nonlocal(chs_context* ctx, const z64* X, int n_b, const z64* wx, z64* out);
// Residual-based Chebyshev filter (R-ChFSI): Y (c128) from the Rayleigh-Ritz
// block X (c128) and its eigenvalues eps (host); the recurrence in
// `precision`, residual and assembly in FP64. wx = beta^H X or null.
int chebyshev_filter_res_impl(chs_context* ctx, int precision, const z64* X, int n_b,
                              const double* eps, int degree, const double* b4, const z64* wx,
                              z64* Y);

}  // namespace chs
