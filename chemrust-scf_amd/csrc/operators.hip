// Operator entry points: H|psi>, the Chebyshev filter, spectral bounds and
// the Lanczos upper bound; level-1 primitives.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"

namespace chs {

int ensure_precision_ready(chs_context* ctx, int precision);

static size_t csize(int precision) { return precision == CHS_C128 ? 16 : 8; }

// H|psi> = T psi + V_loc psi (+ nonlocal), S|psi> when spsi != nullptr.
int apply_h_impl(chs_context* ctx, int precision, const void* psi, int n_b, void* hpsi,
                 void* spsi) {
  CHS_TRY(ensure_precision_ready(ctx, precision));
  Hamiltonian& H = ctx->impl->ham;
  CHS_TRY(vloc_pipeline(ctx, precision, psi, n_b, hpsi, EPI_HPSI, nullptr, nullptr, 0, 0, nullptr,
                        nullptr));
  if (H.n_proj > 0) {
    CHS_TRY(nonlocal_apply(ctx, precision, psi, n_b, hpsi, spsi));
  } else if (spsi && spsi != psi) {
    CHS_HIP(ctx, hipMemcpyAsync(spsi, psi, (size_t)n_b * H.n_pw * csize(precision),
                                hipMemcpyDeviceToDevice, ctx->stream));
  }
  return CHS_OK;
}

__global__ void k_norm_guard2(const double* __restrict__ part, int n_part, int k,
                              double* __restrict__ flag) {
  __shared__ double sh[8];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  double r = block_sum(a, sh);
  double s = block_sum(b, sh);
  if (threadIdx.x == 0) {
    const double nc = sqrt(r), np = sqrt(s);
    const double growth = np > 0.0 ? nc / np : 1.0;
    if (flag[0] == 0.0 && growth > 10.0) {
      flag[0] = k;
      flag[1] = np;
      flag[2] = nc;
      flag[3] = growth;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_sqnorm_part(const cplx<T>* __restrict__ x, size_t n,
                                                             double* __restrict__ part) {
  __shared__ double sh[8];
  double a = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const cplx<T> v = x[i];
    a += (double)v.x * v.x + (double)v.y * v.y;
  }
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Growth guard on consecutive iterates: flag[4] keeps |Y_{k-1}|^2 of the
// previous step; from the second step on, growth > 10 sets flag[0..3] as
// k_norm_guard2 does.
__global__ void k_norm_guard_roll(const double* __restrict__ part, int n_part, int k, int first,
                                  double* __restrict__ flag) {
  __shared__ double sh[8];
  double a = 0;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) a += part[i];
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) {
    if (!first) {
      const double nc = sqrt(r), np = sqrt(flag[4]);
      const double growth = np > 0.0 ? nc / np : 1.0;
      if (flag[0] == 0.0 && growth > 10.0) {
        flag[0] = k;
        flag[1] = np;
        flag[2] = nc;
        flag[3] = growth;
      }
    }
    flag[4] = r;
  }
}

// One Chebyshev step Y_out = [2] (A Y_in - c Y_in)/e [- Y_prev] with the
// growth guard; A = H (fused epilogue) or S^{-1} H (projector path). With a
// residual block (resw, per-band scales ress, its projector coefficients br)
// the step is the residual filter's, Y_out += ress[b] resw_b; guard_first
// then marks the first step the growth guard sees.
static int cheb_step(chs_context* ctx, int precision, const void* yin, int n_b, void* yout,
                     const void* yprev, bool first, const double* d_coef, int k, double* flag,
                     const void* resw = nullptr, const double* ress = nullptr,
                     const void* br = nullptr, bool guard_first = false) {
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const size_t ne = (size_t)n_b * H.n_pw;
  const int mode = resw ? EPI_CHEBK_RES : first ? EPI_CHEB1 : EPI_CHEBK;
  if (resw) first = false;
  if (H.n_proj == 0) {
    const int gz = ceil_div(H.st.n_st, 16);  // partials of the smallest stick chunk
    CHS_TRY(ensure(ctx, im->part, sizeof(double) * 2 * (size_t)gz * n_b + 64));
    int n_part = 0;
    CHS_TRY(vloc_pipeline(ctx, precision, yin, n_b, yout, mode, yprev, d_coef, 0.0, 0.0,
                          (double*)im->part.p, &n_part, resw, ress));
    if (!first) {
      hipLaunchKernelGGL(k_norm_guard2, dim3(1), dim3(256), 0, ctx->stream,
                         (const double*)im->part.p, n_part, k, flag);
      CHS_LAUNCH_CHECK(ctx);
    }
    return CHS_OK;
  }
  // projector path: the local step with the fused epilogue, then the
  // projector terms in coefficient space (cheb_nl_step), then the norm
  CHS_TRY(vloc_pipeline(ctx, precision, yin, n_b, yout, mode, yprev, d_coef, 0.0, 0.0, nullptr,
                        nullptr, resw, ress));
  CHS_TRY(cheb_nl_step(ctx, precision, yout, n_b, first, d_coef, resw ? br : nullptr, ress));
  // |Y_k|^2 over the whole block: enough workgroups to stream it at HBM rate
  constexpr int kNormBlocks = 2048;
  CHS_TRY(ensure(ctx, im->part, sizeof(double) * kNormBlocks + 64));
  if (precision == CHS_C128)
    hipLaunchKernelGGL((k_sqnorm_part<double>), dim3(kNormBlocks), dim3(kRedThreads), 0,
                       ctx->stream, (const z64*)yout, ne, (double*)im->part.p);
  else
    hipLaunchKernelGGL((k_sqnorm_part<float>), dim3(kNormBlocks), dim3(kRedThreads), 0,
                       ctx->stream, (const c32*)yout, ne, (double*)im->part.p);
  hipLaunchKernelGGL(k_norm_guard_roll, dim3(1), dim3(256), 0, ctx->stream,
                     (const double*)im->part.p, kNormBlocks, k, (resw ? guard_first : first) ? 1 : 0,
                     flag);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// The degree-m recurrence of the bound Hamiltonian from Y0 in `a` (`b`: the
// second buffer), window {c, inv_e} in device memory. Stream operations
// only, so it can be captured once every workspace exists. *res receives
// the buffer holding Y_m.
int filter_steps(chs_context* ctx, int precision, void* a, void* b, int n_b, int degree,
                 const double* d_coef, double* flag, void** res) {
  Hamiltonian& H = ctx->impl->ham;
  if (H.n_proj > 0) CHS_TRY(cheb_nl_begin(ctx, precision, a, n_b));
  CHS_TRY(cheb_step(ctx, precision, a, n_b, b, nullptr, true, d_coef, 1, flag));
  for (int k = 2; k <= degree; ++k) {
    CHS_TRY(cheb_step(ctx, precision, b, n_b, a, a, false, d_coef, k, flag));
    std::swap(a, b);
  }
  *res = b;
  return CHS_OK;
}

// Divergence report of a filter's flag block (flag[0..3]); synchronises.
int filter_check(chs_context* ctx, const double* flag) {
  double h_flag[4];
  CHS_HIP(ctx, hipMemcpyAsync(h_flag, flag, 4 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += 4 * sizeof(double);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h_flag[0] != 0.0) {
    ctx->diag.iteration = (int)h_flag[0];
    ctx->diag.norm_previous = h_flag[1];
    ctx->diag.norm_current = h_flag[2];
    ctx->diag.growth_factor = h_flag[3];
    return set_error(ctx, CHS_ERR_CHEBYSHEV_DIVERGED,
                     "Chebyshev filter diverged at iteration %d: norm %.6e -> %.6e (x%.3f)",
                     (int)h_flag[0], h_flag[1], h_flag[2], h_flag[3]);
  }
  return CHS_OK;
}

// Degree-m filter; reference src/eigensolver/chebyshev.rs:626-816. Buffers
// rotate without copies: a = Y_{k-2}, b = Y_{k-1}; Y_k overwrites a.
int chebyshev_filter_impl(chs_context* ctx, int precision, const void* psi, int n_b, int degree,
                          const double* b4, void* out) {
  CHS_TRY(ensure_precision_ready(ctx, precision));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const size_t bytes = (size_t)n_b * H.n_pw * csize(precision);
  if (degree == 0) {
    if (out != psi)
      CHS_HIP(ctx, hipMemcpyAsync(out, psi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return CHS_OK;
  }
  CHS_TRY(ensure(ctx, im->ws1, bytes));
  CHS_TRY(ensure(ctx, im->ws2, bytes));
  CHS_TRY(ensure(ctx, im->small, 64 * sizeof(double)));
  double* flag = (double*)im->small.p;
  double* d_coef = flag + 8;
  const double coef[2] = {b4[2], 1.0 / b4[3]};
  CHS_HIP(ctx, hipMemsetAsync(flag, 0, 8 * sizeof(double), ctx->stream));
  CHS_HIP(ctx, hipMemcpyAsync(d_coef, coef, sizeof(coef), hipMemcpyHostToDevice, ctx->stream));
  ctx->h2d += sizeof(coef);
  void* a = im->ws1.p;
  void* b = im->ws2.p;
  CHS_HIP(ctx, hipMemcpyAsync(a, psi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  void* res = nullptr;
  CHS_TRY(filter_steps(ctx, precision, a, b, n_b, degree, d_coef, flag, &res));
  CHS_HIP(ctx, hipMemcpyAsync(out, res, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return filter_check(ctx, flag);
}

// ---- residual-based filter (R-ChFSI) ------------------------------------
// 2025 R-ChFSI paper (reference_paper/2025-rchfsi-inexact-mv-paper,
// main.tex:539-610, Algorithm 3 with the approximate inverse D^{-1} taken
// as the exact Woodbury S^{-1}). With X the Rayleigh-Ritz block and Lambda
// its eigenvalues, the unscaled recurrence Y_{k+1} = 2 (A - c) Y_k / e -
// Y_{k-1} (A = S^{-1} H, chebyshev.rs:737-762) is carried for the residuals
// R_k = Y_k - X T_k((Lambda - c)/e):
//   R_0 = 0,  R_1 = W / e,  W = A X - X Lambda (formed in FP64),
//   R_{k+1} = 2 (A - c) R_k / e - R_{k-1} + (2/e) T_k(x_b) W_b,
//   Y_m = R_m + X T_m(x_b),  x_b = (eps_b - c) / e.
// In exact arithmetic Y_m is the plain filter's; run in complex64 the
// products' rounding enters in proportion to |W|, which shrinks as the
// subspace converges, so the filtered block keeps FP64 accuracy where the
// plain complex64 filter stalls near FP32 rounding.

// W = SH - eps_b X (SH may alias W)
__global__ void k_res_start(const z64* __restrict__ X, const z64* SH, const double* __restrict__ eps,
                            size_t n_pw, size_t ne, z64* W) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const double e = eps[i / n_pw];
  const z64 x = X[i], h = SH[i];
  W[i] = mk(h.x - e * x.x, h.y - e * x.y);
}

// R_0 = 0, R_1 = inv_e W (W in the filter precision)
template <typename T>
__global__ void k_res_init(const cplx<T>* __restrict__ W, const double* __restrict__ coef, size_t n,
                           cplx<T>* __restrict__ R0, cplx<T>* __restrict__ R1) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T f = (T)coef[1];
  R0[i] = mk((T)0, (T)0);
  R1[i] = mk(f * W[i].x, f * W[i].y);
}

// Y = R + lam_b X (FP64; Y may alias X)
template <typename T>
__global__ void k_res_assemble(const cplx<T>* __restrict__ R, const z64* X,
                               const double* __restrict__ lam, size_t n_pw, size_t ne, z64* Y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const double l = lam[i / n_pw];
  const z64 x = X[i];
  Y[i] = mk((double)R[i].x + l * x.x, (double)R[i].y + l * x.y);
}

template <typename A, typename B>
__global__ void k_res_convert(const A* __restrict__ a, B* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mk((decltype(b[i].x))a[i].x, (decltype(b[i].x))a[i].y);
}

int chebyshev_filter_res_impl(chs_context* ctx, int precision, const z64* X, int n_b,
                              const double* eps, int degree, const double* b4, const z64* wx,
                              z64* Y) {
  CHS_TRY(ensure_precision_ready(ctx, CHS_C128));
  if (precision == CHS_C64) CHS_TRY(ensure_precision_ready(ctx, CHS_C64));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const int n_pw = H.n_pw, P = H.n_proj, m = degree;
  const size_t ne = (size_t)n_b * n_pw, cs = csize(precision);
  if (m == 0) {
    if (Y != X)
      CHS_HIP(ctx, hipMemcpyAsync(Y, X, ne * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
    return CHS_OK;
  }
  // per-band tables: eps | s_1 .. s_{m-1} | T_m, with s_k = (2/e) T_k(x_b)
  const double c = b4[2], inv_e = 1.0 / b4[3];
  std::vector<double> tab((size_t)(m + 1) * n_b);
  for (int b = 0; b < n_b; ++b) {
    const double x = (eps[b] - c) * inv_e;
    double t0 = 1.0, t1 = x;  // T_{k-1}, T_k at k = 1
    tab[b] = eps[b];
    for (int k = 1; k < m; ++k) {
      tab[(size_t)k * n_b + b] = 2.0 * inv_e * t1;
      const double t2 = 2.0 * x * t1 - t0;
      t0 = t1;
      t1 = t2;
    }
    tab[(size_t)m * n_b + b] = t1;  // T_m
  }
  CHS_TRY(ensure(ctx, im->res_tab, tab.size() * sizeof(double) + 16 * sizeof(double)));
  double* d_tab = (double*)im->res_tab.p;
  double* flag = d_tab + tab.size();      // 8 doubles
  double* d_coef = flag + 8;              // {c, inv_e}
  const double coef[2] = {c, inv_e};
  CHS_HIP(ctx, hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  ctx->h2d += tab.size() * sizeof(double);
  CHS_HIP(ctx, hipMemcpyAsync(d_coef, coef, sizeof(coef), hipMemcpyHostToDevice, ctx->stream));
  ctx->h2d += sizeof(coef);
  CHS_HIP(ctx, hipMemsetAsync(flag, 0, 8 * sizeof(double), ctx->stream));
  // FP64 residual W = S^{-1} H X - X Lambda
  CHS_TRY(ensure(ctx, im->res64, ne * sizeof(z64)));
  z64* W64 = (z64*)im->res64.p;
  CHS_TRY(vloc_pipeline(ctx, CHS_C128, X, n_b, W64, EPI_HPSI, nullptr, nullptr, 0.0, 0.0, nullptr,
                        nullptr));
  if (P > 0) CHS_TRY(sinv_h_nonlocal(ctx, X, n_b, wx, W64));
  hipLaunchKernelGGL(k_res_start, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream, X,
                     (const z64*)W64, (const double*)d_tab, (size_t)n_pw, ne, W64);
  CHS_LAUNCH_CHECK(ctx);
  void* Wf = W64;
  if (precision == CHS_C64) {
    CHS_TRY(ensure(ctx, im->res32, ne * sizeof(c32)));
    hipLaunchKernelGGL((k_res_convert<z64, c32>), dim3(ceil_div(ne, 256)), dim3(256), 0,
                       ctx->stream, (const z64*)W64, (c32*)im->res32.p, ne);
    CHS_LAUNCH_CHECK(ctx);
    Wf = im->res32.p;
  }
  // R_0, R_1 and their projector coefficients
  CHS_TRY(ensure(ctx, im->ws1, ne * cs));
  CHS_TRY(ensure(ctx, im->ws2, ne * cs));
  void* a = im->ws1.p;
  void* bb = im->ws2.p;
  if (precision == CHS_C128)
    hipLaunchKernelGGL((k_res_init<double>), dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                       (const z64*)Wf, (const double*)d_coef, ne, (z64*)a, (z64*)bb);
  else
    hipLaunchKernelGGL((k_res_init<float>), dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                       (const c32*)Wf, (const double*)d_coef, ne, (c32*)a, (c32*)bb);
  CHS_LAUNCH_CHECK(ctx);
  void* br = nullptr;
  if (P > 0) {
    CHS_TRY(ensure(ctx, im->res_br, (size_t)P * n_b * cs));
    br = im->res_br.p;
    CHS_TRY(cheb_nl_begin_res(ctx, precision, Wf, n_b, d_coef, br));
  }
  for (int k = 1; k < m; ++k) {
    CHS_TRY(cheb_step(ctx, precision, bb, n_b, a, a, false, d_coef, k + 1, flag, Wf,
                      d_tab + (size_t)k * n_b, br, k == 1));
    std::swap(a, bb);
  }
  // Y = R_m + X T_m
  if (precision == CHS_C128)
    hipLaunchKernelGGL((k_res_assemble<double>), dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                       (const z64*)bb, X, (const double*)(d_tab + (size_t)m * n_b), (size_t)n_pw, ne,
                       Y);
  else
    hipLaunchKernelGGL((k_res_assemble<float>), dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                       (const c32*)bb, X, (const double*)(d_tab + (size_t)m * n_b), (size_t)n_pw, ne,
                       Y);
  CHS_LAUNCH_CHECK(ctx);
  return filter_check(ctx, flag);
}

// ---- Lanczos -----------------------------------------------------------
// The bound runs on S^{-1} H in the S inner product with the projector
// coefficients w_x = beta^H x of every Lanczos vector carried along, so a
// step needs one local H|v>, one beta^H product and one beta product (the
// Chebyshev step's coefficient-space form):
//   g = beta^H Hloc v,  y = beta^H H v = g + B D w_v,
//   c = D w_v - M y  (c = D w_v without S),  S^{-1} H v = Hloc v + beta c,
//   beta^H S^{-1} H v = g + B c,
//   alpha = Re<v|Hloc v> + sum_p D_p |w_v,p|^2,
//   f = S^{-1} H v - alpha v - beta_{j-1} v_prev (same recurrence on w),
//   beta_j^2 = <f|f> + sum_p Q_p |w_f,p|^2.
// B = beta^H beta and M = (Q^{-1} + B)^{-1} are Hermitian, so their rows are
// read as conjugated columns.

// out[p] = a[p] + s sum_q Mat[p][q] x[q], x[q] = (dx ? dx[q] : 1) xin[q],
// a[p] = (da ? da[p] : 1) ain[p] (0 without ain); no sum without Mat.
// One wavefront per row.
__global__ void __launch_bounds__(256) k_herm_mv(int P, const z64* __restrict__ Mat,
                                                 const z64* __restrict__ xin,
                                                 const double* __restrict__ dx,
                                                 const z64* __restrict__ ain,
                                                 const double* __restrict__ da, double s,
                                                 z64* __restrict__ out) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (p >= P) return;
  double re = 0, im = 0;
  if (Mat) {
    const z64* col = Mat + (size_t)p * P;
    for (int q = lane; q < P; q += 64) {
      const z64 m = col[q];
      z64 x = xin[q];
      if (dx) x = mk(dx[q] * x.x, dx[q] * x.y);
      re += m.x * x.x + m.y * x.y;  // conj(m) x
      im += m.x * x.y - m.y * x.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_down(re, off);
      im += __shfl_down(im, off);
    }
  }
  if (lane == 0) {
    z64 a = mk(0.0, 0.0);
    if (ain) {
      a = ain[p];
      if (da) a = mk(da[p] * a.x, da[p] * a.y);
    }
    out[p] = mk(a.x + s * re, a.y + s * im);
  }
}

// dst = red[0] + sum_p d_p Re(conj(x_p) y_p) (d null: no projector part),
// then ab[2 j] = dst (store_sqrt 0) or ab[2 j + 1] = sqrt(max(dst, 0)).
__global__ void k_lanczos_scalar(int P, const double* __restrict__ d, const z64* __restrict__ x,
                                 const z64* __restrict__ y, const double* __restrict__ red,
                                 double* __restrict__ ab, int j, int store_sqrt) {
  __shared__ double sh[8];
  double a = 0;
  if (d)
    for (int p = threadIdx.x; p < P; p += blockDim.x) a += d[p] * (x[p].x * y[p].x + x[p].y * y[p].y);
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) {
    const double t = red[0] + r;
    if (store_sqrt) ab[2 * j + 1] = t > 0 ? sqrt(t) : 0.0;
    else ab[2 * j] = t;
  }
}

// f = u - alpha_j v - beta_{j-1} vp over the n plane waves and the P
// coefficients (indices n .. n + P - 1 address the coefficient arrays).
__global__ void k_lanczos_f(const z64* __restrict__ u, const z64* __restrict__ v,
                            const z64* __restrict__ vp, const z64* __restrict__ wu,
                            const z64* __restrict__ wv, const z64* __restrict__ wvp,
                            const double* __restrict__ ab, int j, size_t n, int P,
                            z64* __restrict__ f, z64* __restrict__ wf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n + (size_t)P) return;
  const double al = ab[2 * j];
  const bool in_pw = i < n;
  const size_t q = in_pw ? i : i - n;
  const z64 ui = in_pw ? u[q] : wu[q], vi = in_pw ? v[q] : wv[q];
  z64 fi = mk(ui.x - al * vi.x, ui.y - al * vi.y);
  if (j > 0) {
    const double be = ab[2 * (j - 1) + 1];
    const z64 pi = in_pw ? vp[q] : wvp[q];
    fi = mk(fi.x - be * pi.x, fi.y - be * pi.y);
  }
  if (in_pw) f[q] = fi;
  else wf[q] = fi;
}

// x /= ab[2 j + 1] over n plane waves and P coefficients (x untouched when
// the divisor is 0, an invariant subspace).
__global__ void k_lanczos_scale(z64* __restrict__ x, z64* __restrict__ wx,
                                const double* __restrict__ ab, int j, size_t n, int P) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n + (size_t)P) return;
  const double be = ab[2 * j + 1];
  if (!(be > 0)) return;
  const double s = 1.0 / be;
  z64* e = i < n ? x + i : wx + (i - n);
  *e = mk(e->x * s, e->y * s);
}

// Extreme eigenvalues of the k x k Lanczos tridiagonal T (alpha_i = ab[2i],
// beta_i = ab[2i+1]) by bisection on the Sturm count, one thread each;
// out = {lambda_max(T) + beta_{k-1}, lambda_min(T)}.
__global__ void k_lanczos_finish(const double* __restrict__ ab, int k, double* __restrict__ out) {
  const int which = threadIdx.x;  // 0: largest, 1: smallest
  if (which > 1 || blockIdx.x != 0) return;
  double gl = INFINITY, gu = -INFINITY, bmax2 = 0.0;
  for (int i = 0; i < k; ++i) {
    const double bl = i > 0 ? fabs(ab[2 * (i - 1) + 1]) : 0.0, br = i + 1 < k ? fabs(ab[2 * i + 1]) : 0.0;
    gl = fmin(gl, ab[2 * i] - bl - br);
    gu = fmax(gu, ab[2 * i] + bl + br);
    if (i + 1 < k) bmax2 = fmax(bmax2, ab[2 * i + 1] * ab[2 * i + 1]);
  }
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, bmax2);
  const double tn = fmax(fabs(gl), fabs(gu));
  double lo = gl - 4.0 * 2.220446049250313e-16 * tn * k - 2.0 * pivmin;
  double hi = gu + 4.0 * 2.220446049250313e-16 * tn * k + 2.0 * pivmin;
  const int idx = which == 0 ? k - 1 : 0;  // eigenvalue index (ascending)
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    int cnt = 0;  // eigenvalues below mid
    double q = ab[0] - mid;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < k; ++i) {
      const double b = ab[2 * (i - 1) + 1];
      q = ab[2 * i] - mid - b * b / q;
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
    if (cnt <= idx) lo = mid;
    else hi = mid;
  }
  const double lam = 0.5 * (lo + hi);
  if (which == 0) out[0] = lam + ab[2 * (k - 1) + 1];
  else out[1] = lam;
}

static int dot_re_into(chs_context* ctx, const z64* a, const z64* b, size_t n, double* dst) {
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->part, sizeof(double) * 2 * kRedBlocks + 64));
  hipLaunchKernelGGL((k_dot_part<double>), dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, a,
                     b, n, (double*)im->part.p);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                     (const double*)im->part.p, kRedBlocks, dst);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// Lanczos bound in the S inner product (S = I without projectors); start
// vector complex N(0,1) per plane wave from the LCG with `seed`. beta_c64:
// the projector products read the complex64 copy of beta (vectors and sums
// stay complex128), for the complex64 filter phase, whose own products use
// that copy: half the bytes of the two matrix-vector passes per step.
int lanczos_impl(chs_context* ctx, int k, uint64_t seed, double* upper, double* ritz_min,
                 bool beta_c64) {
  CHS_TRY(ensure_precision_ready(ctx, CHS_C128));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const int n = H.n_pw;
  const int P = H.n_proj;
  if (beta_c64 && P > 0) CHS_TRY(ensure_precision_ready(ctx, CHS_C64));
  const bool gen = H.has_s;
  if (k < 1 || k > 16) return set_error(ctx, CHS_ERR_INVALID_ARG, "Lanczos steps must be 1..16");
  const size_t vb = (size_t)n * sizeof(z64);
  // start vector from the LCG (n_pw complex), generated on the host once per
  // (seed, n_pw) and kept on the device: the SCF asks for it every iteration
  if (!im->lanczos_v0.p || im->lanczos_seed != seed || im->lanczos_n != n) {
    std::vector<double> h0(2 * (size_t)n);
    uint64_t s = seed;
    auto uni = [&s]() {
      s = (s * 25214903917ull + 11ull) & 0xFFFFFFFFFFFFull;
      return (double)s / 281474976710656.0;
    };
    auto nrm = [&]() {
      double u1 = 1.0 - uni();
      double u2 = uni();
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    };
    for (int g = 0; g < n; ++g) {
      h0[2 * g] = nrm();
      h0[2 * g + 1] = nrm();
    }
    im->lanczos_n = 0;  // invalid until the upload has landed
    CHS_TRY(ensure(ctx, im->lanczos_v0, vb));
    // stream-ordered after any earlier reader of the buffer; h0 must outlive it
    CHS_HIP(ctx, hipMemcpyAsync(im->lanczos_v0.p, h0.data(), vb, hipMemcpyHostToDevice,
                                ctx->stream));
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->h2d += vb;
    im->lanczos_seed = seed;
    im->lanczos_n = n;
  }
  CHS_TRY(ensure(ctx, im->lin_ws, 4 * vb + 7 * (size_t)std::max(P, 1) * sizeof(z64) +
                                      64 * sizeof(double)));
  z64* base = (z64*)im->lin_ws.p;
  // three rotating Lanczos vectors (current, previous, next) and H|v>
  z64* vec[3] = {base, base + n, base + 2 * (size_t)n};
  z64* u = base + 3 * (size_t)n;
  z64* pb = base + 4 * (size_t)n;  // projector coefficients
  const size_t P1 = (size_t)std::max(P, 1);
  z64* wvec[3] = {pb, pb + P1, pb + 2 * P1};
  z64 *g = pb + 3 * P1, *y = pb + 4 * P1, *c = pb + 5 * P1, *wu = pb + 6 * P1;
  double* ab = (double*)(pb + 7 * P1);  // alpha/beta pairs + scratch
  double* red = ab + 40;
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const z64* beta = (const z64*)H.beta;
  const c32* beta32 = beta_c64 ? (const c32*)H.beta_f.p : nullptr;
  // op(beta) x: op C gives the coefficients beta^H x, op N adds beta c to u
  auto bprod = [&](char op, int m, int kk, const z64* x, z64 bt, z64* yv) -> int {
    if (beta32) return zgemv_c32(ctx, op, m, kk, one, beta32, n, x, bt, yv);
    return zgemm(ctx, op, 'N', m, 1, kk, one, beta, n, x, kk, bt, yv, m);
  };
  const z64* B = (const z64*)H.bmat.p;
  const z64* M = gen ? (const z64*)H.sinv_m.p : nullptr;
  const double* Q = gen ? H.qion : nullptr;
  const unsigned nblk = (unsigned)ceil_div((size_t)n + P, 256);
  int cur = 0, prv = 1, nxt = 2;
  CHS_HIP(ctx, hipMemcpyAsync(vec[cur], im->lanczos_v0.p, vb, hipMemcpyDeviceToDevice,
                              ctx->stream));
  CHS_HIP(ctx, hipMemsetAsync(ab, 0, 40 * sizeof(double), ctx->stream));
  // normalise: v /= sqrt(<v|v> + sum Q |w_v|^2)
  if (P > 0) CHS_TRY(bprod('C', P, n, vec[cur], zero, wvec[cur]));
  CHS_TRY(dot_re_into(ctx, vec[cur], vec[cur], n, red));
  hipLaunchKernelGGL(k_lanczos_scalar, dim3(1), dim3(256), 0, ctx->stream, P, Q, wvec[cur],
                     wvec[cur], red, ab, 16, 1);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_lanczos_scale, dim3(nblk), dim3(256), 0, ctx->stream, vec[cur], wvec[cur],
                     ab, 16, (size_t)n, P);
  CHS_LAUNCH_CHECK(ctx);
  for (int j = 0; j < k; ++j) {
    CHS_TRY(vloc_pipeline(ctx, CHS_C128, vec[cur], 1, u, EPI_HPSI, nullptr, nullptr, 0, 0, nullptr,
                          nullptr));
    CHS_TRY(dot_re_into(ctx, vec[cur], u, n, red));
    hipLaunchKernelGGL(k_lanczos_scalar, dim3(1), dim3(256), 0, ctx->stream, P, H.dion,
                       wvec[cur], wvec[cur], red, ab, j, 0);
    CHS_LAUNCH_CHECK(ctx);
    if (P > 0) {
      const unsigned pblk = (unsigned)ceil_div(P, 4);
      CHS_TRY(bprod('C', P, n, u, zero, g));
      hipLaunchKernelGGL(k_herm_mv, dim3(pblk), dim3(256), 0, ctx->stream, P, B, wvec[cur],
                         H.dion, g, nullptr, 1.0, y);
      CHS_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_herm_mv, dim3(pblk), dim3(256), 0, ctx->stream, P, M, y, nullptr,
                         wvec[cur], H.dion, -1.0, c);
      CHS_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_herm_mv, dim3(pblk), dim3(256), 0, ctx->stream, P, B, c, nullptr, g,
                         nullptr, 1.0, wu);
      CHS_LAUNCH_CHECK(ctx);
      CHS_TRY(bprod('N', n, P, c, one, u));
    }
    hipLaunchKernelGGL(k_lanczos_f, dim3(nblk), dim3(256), 0, ctx->stream, u, vec[cur], vec[prv],
                       wu, wvec[cur], wvec[prv], ab, j, (size_t)n, P, vec[nxt], wvec[nxt]);
    CHS_LAUNCH_CHECK(ctx);
    CHS_TRY(dot_re_into(ctx, vec[nxt], vec[nxt], n, red));
    hipLaunchKernelGGL(k_lanczos_scalar, dim3(1), dim3(256), 0, ctx->stream, P, Q, wvec[nxt],
                       wvec[nxt], red, ab, j, 1);
    CHS_LAUNCH_CHECK(ctx);
    if (j + 1 < k) {
      hipLaunchKernelGGL(k_lanczos_scale, dim3(nblk), dim3(256), 0, ctx->stream, vec[nxt],
                         wvec[nxt], ab, j, (size_t)n, P);
      CHS_LAUNCH_CHECK(ctx);
      const int t = prv;
      prv = cur;
      cur = nxt;
      nxt = t;
    }
  }
  hipLaunchKernelGGL(k_lanczos_finish, dim3(1), dim3(64), 0, ctx->stream, ab, k, red + 4);
  CHS_LAUNCH_CHECK(ctx);
  double hr[2];
  CHS_HIP(ctx, hipMemcpyAsync(hr, red + 4, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += 2 * sizeof(double);
  *upper = hr[0];
  if (ritz_min) *ritz_min = hr[1];
  return CHS_OK;
}

// ---- level-1 primitives ------------------------------------------------
__global__ void k_zaxpy(size_t n, double ar, double ai, const z64* __restrict__ x, z64* y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const z64 a = mk(ar, ai);
    y[i] = y[i] + a * x[i];
  }
}
__global__ void k_zscal(size_t n, double ar, double ai, z64* x) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = mk(ar, ai) * x[i];
}

__global__ void k_minmax(const double* __restrict__ v, size_t n, double* __restrict__ out) {
  __shared__ double smin[256], smax[256];
  double mn = INFINITY, mx = -INFINITY;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
    mn = fmin(mn, v[i]);
    mx = fmax(mx, v[i]);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smin[0];
    out[1] = smax[0];
  }
}

int vloc_timed(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int reps,
               double* ms);

__global__ void k_pattern(double* __restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    a[i] = ((double)(h & 0xFFFF) / 65536.0 - 0.5) * 1e-2;
  }
}

// ---- allocation-free dot product ---------------------------------------
// Scratch of chs_zdotc_noalloc: kRedBlocks real then kRedBlocks imaginary
// partial sums, then the two totals (514 doubles, 4,112 bytes). Every GPU the
// library runs on holds its own copy until the process ends. Host code never
// passes the array's name to a kernel or copy: the name is a host-side
// stand-in, not a device address; it uses the per-device address that
// hipGetSymbolAddress returned for that device (noalloc_scratch).
static __device__ __align__(256) double scratch[2 * kRedBlocks + 2];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

// The context device's scratch address, looked up once per device with that
// device current (chs_context_create calls this at set-up).
int noalloc_scratch(chs_context* ctx, double** out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "device %d beyond the scratch table", ctx->device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    int cur = -1;
    CHS_HIP(ctx, hipGetDevice(&cur));
    if (cur != ctx->device) CHS_HIP(ctx, hipSetDevice(ctx->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != ctx->device) (void)hipSetDevice(cur);
    CHS_HIP(ctx, e);
    if (!p) return set_error(ctx, CHS_ERR_DEVICE, "no device address for the dot-product scratch");
    g_scratch_ptr[ctx->device] = (double*)p;
  }
  *out = g_scratch_ptr[ctx->device];
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_zdotc_noalloc(chs_context* ctx, int n, const double* d_x, const double* d_y,
                      double out[2]) {
  if (!ctx || n < 0 || !out || (n && (!d_x || !d_y))) return CHS_ERR_INVALID_ARG;
  double* s = nullptr;
  CHS_TRY(noalloc_scratch(ctx, &s));
  // every launch writes all 2 kRedBlocks partials and both totals before they
  // are read, so nothing carries over from an earlier call
  hipLaunchKernelGGL((k_dot_part<double>), dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                     (const z64*)d_x, (const z64*)d_y, (size_t)n, s);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double*)s,
                     kRedBlocks, s + 2 * kRedBlocks);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(out, s + 2 * kRedBlocks, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += 2 * sizeof(double);
  return CHS_OK;
}


int chs_profile_hpsi(chs_context* ctx, int precision, int n_b, int reps, double* ms3,
                     double* bytes3, int32_t* dims4) {
  if (!ctx || n_b <= 0 || reps <= 0 || !ms3 || (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  CHS_TRY(ensure_precision_ready(ctx, precision));
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  const size_t s = csize(precision), ne = (size_t)n_b * H.n_pw;
  CHS_TRY(ensure(ctx, im->ws1, ne * s));
  CHS_TRY(ensure(ctx, im->ws2, ne * s));
  const size_t nd = ne * s / sizeof(double);
  hipLaunchKernelGGL(k_pattern, dim3(ceil_div(nd, 256)), dim3(256), 0, ctx->stream,
                     (double*)im->ws1.p, nd);
  hipLaunchKernelGGL(k_pattern, dim3(ceil_div(nd, 256)), dim3(256), 0, ctx->stream,
                     (double*)im->ws2.p, nd);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(vloc_timed(ctx, precision, im->ws1.p, n_b, im->ws2.p, reps, ms3));
  const double n = H.n_grid, n3 = n * n * n, nst = H.st.n_st, npw = H.n_pw, v = s / 2.0;
  const double stick = s * n * nst * n_b;
  if (bytes3) {
    bytes3[0] = s * npw * n_b + stick;                            // psi in, sticks out
    bytes3[1] = 2.0 * stick + v * n3;                              // sticks in/out, V once
    bytes3[2] = stick + 2.0 * s * npw * n_b + 8.0 * npw + s * npw * n_b;  // sticks, 2 psi, T, out
  }
  if (dims4) {
    dims4[0] = H.n_grid;
    dims4[1] = H.n_pw;
    dims4[2] = H.st.n_st;
    dims4[3] = H.st.n_act;
  }
  return CHS_OK;
}

int chs_apply_h(chs_context* ctx, int precision, const void* d_psi, int n_b, void* d_hpsi,
                void* d_spsi) {
  if (!ctx || !d_psi || !d_hpsi || n_b <= 0 || (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  return apply_h_impl(ctx, precision, d_psi, n_b, d_hpsi, d_spsi);
}

int chs_chebyshev_filter(chs_context* ctx, int precision, const void* d_psi, int n_b, int degree,
                         const double* bounds4, void* d_out) {
  if (!ctx || !d_psi || !d_out || !bounds4 || n_b <= 0 || degree < 0 ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  if (!(bounds4[3] > 0.0))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "filter half width must be positive");
  return chebyshev_filter_impl(ctx, precision, d_psi, n_b, degree, bounds4, d_out);
}

int chs_chebyshev_filter_residual(chs_context* ctx, int precision, const double* d_x, int n_b,
                                  const double* eps, int degree, const double* bounds4,
                                  double* d_out) {
  if (!ctx || !d_x || !d_out || !eps || !bounds4 || n_b <= 0 || degree < 0 ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  if (!(bounds4[3] > 0.0))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "filter half width must be positive");
  return chebyshev_filter_res_impl(ctx, precision, (const z64*)d_x, n_b, eps, degree, bounds4,
                                   nullptr, (z64*)d_out);
}

int chs_lanczos_bound(chs_context* ctx, int k, uint64_t seed, double* upper, double* ritz_min) {
  if (!ctx || !upper) return CHS_ERR_INVALID_ARG;
  return lanczos_impl(ctx, k, seed, upper, ritz_min, false);
}

int chs_grid_lambda_max(chs_context* ctx, double L, double* out) {
  if (!ctx || !out || !(L > 0)) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  CHS_TRY(ensure(ctx, ctx->impl->small, 64 * sizeof(double)));
  double* d = (double*)ctx->impl->small.p + 16;
  const size_t n3 = (size_t)H.n_grid * H.n_grid * H.n_grid;
  hipLaunchKernelGGL(k_minmax, dim3(1), dim3(256), 0, ctx->stream, H.veff, n3, d);
  CHS_LAUNCH_CHECK(ctx);
  double mm[2];
  CHS_HIP(ctx, hipMemcpyAsync(mm, d, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += 2 * sizeof(double);
  // gmax: the grid corner, as GVectorGrid::gmax in the reference (D6)
  const double b = 2.0 * M_PI / L;
  const double gmax = b * std::sqrt(3.0) * (H.n_grid / 2);
  *out = 0.5 * gmax * gmax + (mm[1] - mm[0]);
  return CHS_OK;
}

int chs_zdotc(chs_context* ctx, int n, const double* d_x, const double* d_y, double out[2]) {
  if (!ctx || n < 0 || !out || (n && (!d_x || !d_y))) return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->small, 64 * sizeof(double)));
  double* dst = (double*)im->small.p + 32;
  CHS_TRY(ensure(ctx, im->part, sizeof(double) * 2 * kRedBlocks + 64));
  hipLaunchKernelGGL((k_dot_part<double>), dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                     (const z64*)d_x, (const z64*)d_y, (size_t)n, (double*)im->part.p);
  CHS_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                     (const double*)im->part.p, kRedBlocks, dst);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(out, dst, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += 2 * sizeof(double);
  return CHS_OK;
}

int chs_zaxpy(chs_context* ctx, int n, const double alpha[2], const double* d_x, double* d_y) {
  if (!ctx || n < 0 || !alpha || (n && (!d_x || !d_y))) return CHS_ERR_INVALID_ARG;
  if (!n) return CHS_OK;
  hipLaunchKernelGGL(k_zaxpy, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (size_t)n,
                     alpha[0], alpha[1], (const z64*)d_x, (z64*)d_y);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int chs_zscal(chs_context* ctx, int n, const double alpha[2], double* d_x) {
  if (!ctx || n < 0 || !alpha || (n && !d_x)) return CHS_ERR_INVALID_ARG;
  if (!n) return CHS_OK;
  hipLaunchKernelGGL(k_zscal, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (size_t)n,
                     alpha[0], alpha[1], (z64*)d_x);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

}  // extern "C"
