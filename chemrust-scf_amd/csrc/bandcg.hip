// Batched band-by-band CG refinement on the device.
//
// Reference: band_cg_minimize (src/eigensolver/band_cg.rs:157-437, CASTEP
// electronic_find_eigenstate) minimises one band at a time against a set
// of already converged lower bands, with the closed-form line search of
// line_search.rs:57-198, the residual / S-normalisation /
// S-orthogonalisation helpers of cg_helpers.rs:49-216 and the TPA + USPP
// preconditioner of uspp_preconditioner.rs:38-244.
//
// Here every band of a block runs that same iteration at once. Band i's
// "converged bands" are the bands j < i of the block as it was handed in
// (the Rayleigh-Ritz vectors), so each band's trajectory is exactly the
// reference's band_cg_minimize(psi_i, precond, {(psi_j, S psi_j), j < i})
// while all bands share the batched H|d>, the projector GEMMs and the
// orthogonalisation GEMM. Bands stop updating independently when they meet
// the reference's exit conditions.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "bandcg.hpp"
#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"

namespace chs {

// ---- per-band reductions ---------------------------------------------------
// out[b] = <x_b|y_b> (complex) over the band's n_pw coefficients; one block
// per band, fixed reduction order.
__global__ void __launch_bounds__(256) k_bdot(const z64* __restrict__ x, const z64* __restrict__ y,
                                              int n_pw, z64* __restrict__ out) {
  __shared__ double sh[8];
  const int b = blockIdx.x;
  const z64* xb = x + (size_t)b * n_pw;
  const z64* yb = y + (size_t)b * n_pw;
  double re = 0, im = 0;
  for (int g = threadIdx.x; g < n_pw; g += blockDim.x) {
    const z64 a = xb[g], c = yb[g];
    re += a.x * c.x + a.y * c.y;
    im += a.x * c.y - a.y * c.x;
  }
  const double r = block_sum(re, sh);
  const double i = block_sum(im, sh);
  if (threadIdx.x == 0) out[b] = mk(r, i);
}

// ---- line search on the device (line_search.rs:57-198) --------------------
__device__ void line_search_dev(double a, double hd, double cc, double d, double* step,
                                double* norm, int* status) {
  const double b = -2.0 * hd;
  double st, ev;
  int s;
  if (fabs(d) > DBL_MIN) {
    const double ad = a * d, bd = b * d;
    if (fabs(bd) < 1e-30) { *step = 0; *norm = 1; *status = -2; return; }
    const double tmp = ad * ad - 2.0 * ad * cc + cc * cc + b * bd;
    if (tmp < 0.0 || fabs(bd) < DBL_MIN) { *step = 0; *norm = 1; *status = -1; return; }
    const double det = sqrt(tmp);
    const double r1 = (-ad + cc + det) / bd, r2 = (-ad + cc - det) / bd;
    const double den1 = 1.0 + d * r1 * r1, den2 = 1.0 + d * r2 * r2;
    const bool ok1 = den1 > 0.0, ok2 = den2 > 0.0;
    const double e1 = ok1 ? (a + b * r1 + cc * r1 * r1) / den1 : 0.0;
    const double e2 = ok2 ? (a + b * r2 + cc * r2 * r2) / den2 : 0.0;
    double r, e;
    if (ok1 && ok2) {
      if (e1 < e2) { r = r1; e = e1; } else { r = r2; e = e2; }
    } else if (ok1 && e1 < a) {
      r = r1; e = e1;
    } else if (ok2 && e2 < a) {
      r = r2; e = e2;
    } else {
      *step = 0; *norm = 1; *status = -1; return;
    }
    st = -r; ev = e; s = 0;
  } else if (cc > 0.0) {
    st = -b / (2.0 * cc); ev = 0; s = 0;
  } else {
    st = 0; ev = a; s = -1;
  }
  (void)ev;
  if (st > 15.0) { st = 15.0; s = 1; } else if (st < 0.0) { s = -1; }
  const double t = 1.0 + d * st * st;
  *step = st;
  if (t > 0.0) { *norm = sqrt(t); *status = s; } else { *norm = 0.0; *status = 2; }
}

// ---- elementwise batched updates --------------------------------------------
// Per-band state, one struct per band in device memory.
struct BandState {
  double eps, prev, beta_old, step, norm, rnorm;
  int active, n_steps, converged, status, flag;
};

__global__ void k_scale_bands(z64* __restrict__ x, int n_pw, const double* __restrict__ s,
                              const BandState* __restrict__ st) {
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw || (st && !st[b].active)) return;
  const double f = s[b];
  z64* v = x + (size_t)b * n_pw + g;
  *v = mk(v->x * f, v->y * f);
}

// y_b = a_b x_b + c_b y_b for active bands
__global__ void k_axpby_bands(const z64* __restrict__ x, z64* __restrict__ y, int n_pw,
                              const z64* __restrict__ a, const z64* __restrict__ c,
                              const BandState* __restrict__ st) {
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw || !st[b].active) return;
  const size_t i = (size_t)b * n_pw + g;
  y[i] = a[b] * x[i] + c[b] * y[i];
}

// r = h - (Re<psi|h> / Re<psi|s>) s (compute_residual, cg_helpers.rs:49-82)
__global__ void k_residual(const z64* __restrict__ h, const z64* __restrict__ s, int n_pw,
                           const z64* __restrict__ dh, const z64* __restrict__ ds,
                           z64* __restrict__ r, const BandState* __restrict__ st) {
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw || !st[b].active) return;
  const double e = dh[b].x / ds[b].x;
  const size_t i = (size_t)b * n_pw + g;
  r[i] = mk(h[i].x - e * s[i].x, h[i].y - e * s[i].y);
}

// TPA preconditioner t(g) (uspp_preconditioner.rs:148-180) applied per band;
// with an MLP scale table (n_b x n_pw, f32) the table replaces t.
__global__ void k_tpa_apply(const z64* __restrict__ r, const double* __restrict__ t,
                            const float* __restrict__ mlp, int n_pw, z64* __restrict__ out,
                            const BandState* __restrict__ st) {
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw || !st[b].active) return;
  const size_t i = (size_t)b * n_pw + g;
  const double f = mlp ? (double)mlp[i] : t[g];
  out[i] = mk(r[i].x * f, r[i].y * f);
}

// out = tr + t (beta WR) for the USPP correction (elementwise part)
__global__ void k_precond_fix(z64* __restrict__ out, const z64* __restrict__ corr,
                              const double* __restrict__ t, const float* __restrict__ mlp,
                              int n_pw, const BandState* __restrict__ st) {
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw || !st[b].active) return;
  const size_t i = (size_t)b * n_pw + g;
  const double f = mlp ? (double)mlp[i] : t[g];
  out[i] = mk(out[i].x + f * corr[i].x, out[i].y + f * corr[i].y);
}

// Masked overlaps for the lower-band S-orthogonalisation:
// OV[j, i] <- Re(OV[j, i]) for j < i, else 0 (cg_helpers.rs:188 uses the real
// part); with `full` the complex OV[j, i] is kept, which removes the whole
// component along band j (the real part alone leaves i Im(OV) psi_j behind,
// so complex bands drift into lower ones over many steps)
__global__ void k_lower_real(z64* __restrict__ ov, int n_b, const BandState* __restrict__ st,
                             int full) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_b) return;
  z64& v = ov[(size_t)j + (size_t)i * n_b];
  v = (j < i && st[i].active) ? (full ? v : mk(v.x, 0.0)) : mk(0.0, 0.0);
}

// ---- per-band scalar logic ---------------------------------------------------
// After the initial H/S apply: S-normalise and set eps (band_cg.rs:178-208)
__global__ void k_cg_init(BandState* st, const z64* __restrict__ dps, double* __restrict__ scale,
                          int n_b) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b) return;
  const double so = dps[b].x;
  st[b].active = so > 0.0;
  st[b].flag = so > 0.0 ? 0 : -1;
  scale[b] = so > 0.0 ? 1.0 / sqrt(so) : 1.0;
  st[b].n_steps = 0;
  st[b].converged = 0;
  st[b].beta_old = 0.0;
}

__global__ void k_cg_set_eps(BandState* st, const z64* __restrict__ dph, int n_b) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) return;
  st[b].eps = dph[b].x;
}

// ||g|| guard (band_cg.rs:208-218) and prev eigenvalue
__global__ void k_cg_gguard(BandState* st, const z64* __restrict__ dgg, const z64* __restrict__ drr,
                            int step, int n_b) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) return;
  if (sqrt(fmax(dgg[b].x, 0.0)) < DBL_EPSILON * 1000.0) {
    st[b].active = 0;
    st[b].converged = 1;
    st[b].n_steps = step - 1;
    st[b].rnorm = sqrt(fmax(drr[b].x, 0.0));
    return;
  }
  st[b].prev = st[b].eps;
}

// S-normalisation factor of the orthogonalised gradient (cg_helpers.rs:204-215)
__global__ void k_cg_gnorm(BandState* st, const z64* __restrict__ dgs, double* __restrict__ scale,
                           int n_b) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) return;
  const double so = dgs[b].x;
  if (!(so > 0.0)) {
    st[b].active = 0;
    st[b].flag = -2;
    scale[b] = 1.0;
    return;
  }
  scale[b] = 1.0 / sqrt(so);
}

// FR coefficients: d = g (SD, steps <= 2) or gamma d_old - g (band_cg.rs:248-278)
__global__ void k_cg_dir_coef(BandState* st, const z64* __restrict__ dgr, int step, int n_b,
                              z64* __restrict__ ca, z64* __restrict__ cc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) return;
  const double beta = dgr[b].x;
  if (step <= 2) {
    ca[b] = mk(1.0, 0.0);  // d = 1 * g + 0 * d_old
    cc[b] = mk(0.0, 0.0);
  } else {
    const double gamma = fabs(st[b].beta_old) > 1e-300 ? beta / st[b].beta_old : 0.0;
    ca[b] = mk(-1.0, 0.0);  // d = -g + gamma d_old
    cc[b] = mk(gamma, 0.0);
  }
  st[b].beta_old = beta;
}

// -overlap with the current band (band_cg.rs:305-316): coefficient for
// x <- x - <psi|sd> psi, only when |<psi|sd>| > 1e-30
__global__ void k_cg_ovcoef(const BandState* st, const z64* __restrict__ dov, int n_b,
                            z64* __restrict__ ca, z64* __restrict__ cc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b) return;
  const z64 o = dov[b];
  const bool use = st[b].active && sqrt(o.x * o.x + o.y * o.y) > 1e-30;
  ca[b] = use ? mk(-o.x, -o.y) : mk(0.0, 0.0);
  cc[b] = mk(1.0, 0.0);
}

// residual guard, line search, update coefficients (band_cg.rs:326-389)
__global__ void k_cg_linesearch(BandState* st, const z64* __restrict__ drr,
                                const z64* __restrict__ dhd, const z64* __restrict__ dhdd,
                                const z64* __restrict__ dsdd, double tol, int step, int n_b,
                                z64* __restrict__ ca, z64* __restrict__ cc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) {
    if (b < n_b) { ca[b] = mk(0.0, 0.0); cc[b] = mk(1.0, 0.0); }
    return;
  }
  const double rn = sqrt(fmax(drr[b].x, 0.0));
  if (rn < tol * 0.01 && rn < 1e-12 * fmax(fabs(st[b].eps), 1.0)) {
    st[b].active = 0;
    st[b].converged = 1;
    st[b].n_steps = step - 1;
    st[b].rnorm = rn;
    ca[b] = mk(0.0, 0.0);
    cc[b] = mk(1.0, 0.0);
    return;
  }
  double s, nrm;
  int status;
  line_search_dev(st[b].eps, dhd[b].x, dhdd[b].x, dsdd[b].x, &s, &nrm, &status);
  st[b].step = s;
  st[b].norm = nrm;
  st[b].status = status;
  const double in = nrm > 0.0 ? 1.0 / nrm : 1.0;
  ca[b] = mk(s * in, 0.0);  // x <- in * x + (s * in) * dir
  cc[b] = mk(in, 0.0);
}

// convergence |eps_new - prev| < tol (band_cg.rs:408-419)
__global__ void k_cg_check(BandState* st, const z64* __restrict__ dph, const z64* __restrict__ drr,
                           double tol, int step, int max_steps, int n_b) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_b || !st[b].active) return;
  const double e = dph[b].x;
  st[b].rnorm = sqrt(fmax(drr[b].x, 0.0));
  st[b].n_steps = step;
  if (fabs(e - st[b].prev) < tol) {
    st[b].eps = e;
    st[b].converged = 1;
    st[b].active = 0;
    return;
  }
  st[b].eps = e;
  if (step == max_steps) st[b].active = 0;
}

// ---------------------------------------------------------------------------
int BandCg::ensure_ws(chs_context* c, int n_pw_, int n_b_) {
  ctx = c;
  n_pw = n_pw_;
  n_b = n_b_;
  const size_t ne = (size_t)n_pw * n_b;
  CHS_TRY(ensure(ctx, vecs, 12 * ne * sizeof(z64)));
  CHS_TRY(ensure(ctx, scal, (size_t)n_b * (12 * sizeof(z64) + sizeof(BandState) + 2 * sizeof(double)) +
                                (size_t)n_b * n_b * sizeof(z64) + 1024));
  return CHS_OK;
}

BandCg::~BandCg() {
  release(vecs);
  release(scal);
  release(tpa);
  release(rmat);
}

}  // namespace chs

namespace chs {

// S x (no H): x + beta Q beta^H x with projectors, else a copy.
static int s_only(chs_context* ctx, const z64* x, int n_b, z64* sx) {
  Hamiltonian& H = ctx->impl->ham;
  const size_t ne = (size_t)n_b * H.n_pw;
  CHS_HIP(ctx, hipMemcpyAsync(sx, x, ne * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream));
  if (H.n_proj == 0 || !H.has_s) return CHS_OK;
  const int P = H.n_proj, n_pw = H.n_pw;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(z64)));
  z64* W = (z64*)ctx->impl->proj_ws.p;
  CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0, 0.0), (const z64*)H.beta, n_pw, x, n_pw,
                mk(0.0, 0.0), W, P));
  CHS_TRY(bcg_scale_rows(ctx, W, P, n_b, H.qion));
  CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, P, mk(1.0, 0.0), (const z64*)H.beta, n_pw, W, P,
                mk(1.0, 0.0), sx, n_pw));
  return CHS_OK;
}

__global__ void k_scale_rows(z64* __restrict__ w, int P, int n_b, const double* __restrict__ q) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)P * n_b) return;
  const double f = q[i % P];
  w[i] = mk(w[i].x * f, w[i].y * f);
}

int bcg_scale_rows(chs_context* ctx, z64* w, int P, int n_b, const double* q) {
  hipLaunchKernelGGL(k_scale_rows, dim3(ceil_div((size_t)P * n_b, 256)), dim3(256), 0, ctx->stream,
                     w, P, n_b, q);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// t(g) on the device from the bound kinetic energies (mean over plane waves)
__global__ void k_tpa_table(const double* __restrict__ kin, int n, double inv_mean,
                            double* __restrict__ t) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double x = kin[g] * inv_mean;
  const double tmp = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
  t[g] = 1.0 / (1.0 + 16.0 * (x * x) * (x * x) / tmp);
}

__global__ void k_scale_beta(const z64* __restrict__ beta, const double* __restrict__ t, int n_pw,
                             size_t n, z64* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = t[i % n_pw];
  out[i] = mk(beta[i].x * f, beta[i].y * f);
}

__global__ void k_negate_add_qinv(z64* __restrict__ c, const double* __restrict__ q, int P) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)P * P) return;
  const int r = (int)(i % P), col = (int)(i / P);
  if (r == col) c[i].x += 1.0 / q[r];
}

__global__ void k_negate(z64* __restrict__ c, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = mk(-c[i].x, -c[i].y);
}

// Preconditioner setup: TPA table and, with projectors, R = (-Q^{-1} - C)^{-1},
// C = beta^H T beta (uspp_preconditioner.rs:56-73, :187-244). R = -(Q^{-1} +
// C)^{-1} is formed through the Cholesky factor of the positive definite
// Q^{-1} + C.
int BandCg::setup_precond(bool uspp) {
  Hamiltonian& H = ctx->impl->ham;
  CHS_TRY(ensure(ctx, tpa, (size_t)n_pw * sizeof(double)));
  std::vector<double> kin(n_pw);
  CHS_HIP(ctx, hipMemcpyAsync(kin.data(), H.kin, n_pw * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += n_pw * sizeof(double);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double sum = 0;
  for (double k : kin) sum += k;
  const double mean = sum / n_pw;
  hipLaunchKernelGGL(k_tpa_table, dim3(ceil_div(n_pw, 256)), dim3(256), 0, ctx->stream, H.kin,
                     n_pw, mean > 0 ? 1.0 / mean : 1.0, (double*)tpa.p);
  CHS_LAUNCH_CHECK(ctx);
  use_r = uspp && H.n_proj > 0 && H.has_s;
  if (!use_r) return CHS_OK;
  const int P = H.n_proj;
  const size_t pp = (size_t)P * P, pn = (size_t)P * n_pw;
  CHS_TRY(ensure(ctx, rmat, pp * sizeof(z64)));
  CHS_TRY(ensure(ctx, ctx->impl->lin_ws, (2 * pp + pn) * sizeof(z64)));
  z64* Cm = (z64*)ctx->impl->lin_ws.p;
  z64* Li = Cm + pp;
  z64* bt = Li + pp;
  hipLaunchKernelGGL(k_scale_beta, dim3(ceil_div(pn, 256)), dim3(256), 0, ctx->stream,
                     (const z64*)H.beta, (const double*)tpa.p, n_pw, pn, bt);
  CHS_TRY(zgemm(ctx, 'C', 'N', P, P, n_pw, mk(1.0, 0.0), (const z64*)H.beta, n_pw, bt, n_pw,
                mk(0.0, 0.0), Cm, P));
  CHS_TRY(hermitise(ctx, Cm, P));
  hipLaunchKernelGGL(k_negate_add_qinv, dim3(ceil_div(pp, 256)), dim3(256), 0, ctx->stream, Cm,
                     H.qion, P);
  int info = 0;
  CHS_TRY(cholesky(ctx, Cm, P, &info));
  if (info) return set_error(ctx, CHS_ERR_SOLVER, "Q^-1 + beta^H T beta not positive definite");
  CHS_TRY(tri_inv(ctx, Cm, Li, P));
  CHS_TRY(zgemm(ctx, 'C', 'N', P, P, P, mk(-1.0, 0.0), Li, P, Li, P, mk(0.0, 0.0), (z64*)rmat.p, P));
  return CHS_OK;
}

// G = P^{-1} R for the active bands
int BandCg::precond(const z64* R, z64* G, const float* mlp, const void* st) {
  Hamiltonian& H = ctx->impl->ham;
  const BandState* bs = (const BandState*)st;
  dim3 grid(ceil_div(n_pw, 256), n_b);
  hipLaunchKernelGGL(k_tpa_apply, grid, dim3(256), 0, ctx->stream, R, (const double*)tpa.p, mlp,
                     n_pw, G, bs);
  CHS_LAUNCH_CHECK(ctx);
  if (!use_r) return CHS_OK;
  const int P = H.n_proj;
  const size_t pb = (size_t)P * n_b;
  CHS_TRY(ensure(ctx, ctx->impl->proj_ws, 3 * pb * sizeof(z64)));
  z64* W = (z64*)ctx->impl->proj_ws.p;
  z64* WR = W + pb;
  z64* corr = vec(11);
  CHS_TRY(zgemm(ctx, 'C', 'N', P, n_b, n_pw, mk(1.0, 0.0), (const z64*)H.beta, n_pw, G, n_pw,
                mk(0.0, 0.0), W, P));
  CHS_TRY(zgemm(ctx, 'N', 'N', P, n_b, P, mk(1.0, 0.0), (const z64*)rmat.p, P, W, P, mk(0.0, 0.0),
                WR, P));
  CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, P, mk(1.0, 0.0), (const z64*)H.beta, n_pw, WR, P,
                mk(0.0, 0.0), corr, n_pw));
  hipLaunchKernelGGL(k_precond_fix, grid, dim3(256), 0, ctx->stream, G, corr, (const double*)tpa.p,
                     mlp, n_pw, bs);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int BandCg::run(z64* psi, int max_steps, double tol, const float* mlp, double* eps_out,
                int* steps_out, int* conv_out) {
  const size_t ne = (size_t)n_pw * n_b;
  z64 *HP = vec(0), *SP = vec(1), *R = vec(2), *G = vec(3), *SG = vec(4), *D = vec(5),
      *DOLD = vec(6), *HD = vec(7), *SD = vec(8), *CONV = vec(9), *SCONV = vec(10);
  unsigned char* s = (unsigned char*)scal.p;
  z64* d1 = (z64*)s;  // per-band dot scratch, 6 slots
  z64* d2 = d1 + n_b;
  z64* d3 = d2 + n_b;
  z64* d4 = d3 + n_b;
  z64* ca = d4 + n_b;
  z64* cc = ca + n_b;
  BandState* st = (BandState*)(cc + n_b);
  double* scale = (double*)(st + n_b);
  z64* OV = (z64*)(scale + n_b);
  const dim3 eg(ceil_div(n_pw, 256), n_b), sg1(ceil_div(n_b, 256));
  auto bdot = [&](const z64* x, const z64* y, z64* out) {
    hipLaunchKernelGGL(k_bdot, dim3(n_b), dim3(256), 0, ctx->stream, x, y, n_pw, out);
  };
  auto copy = [&](z64* dst, const z64* src) {
    return hipMemcpyAsync(dst, src, ne * sizeof(z64), hipMemcpyDeviceToDevice, ctx->stream);
  };
  // converged set = the block as handed in, with its S image
  CHS_HIP(ctx, copy(CONV, psi));
  CHS_HIP(ctx, hipMemsetAsync(DOLD, 0, ne * sizeof(z64), ctx->stream));
  CHS_TRY(apply_h_impl(ctx, CHS_C128, psi, n_b, HP, SP));
  CHS_HIP(ctx, copy(SCONV, SP));
  bdot(psi, SP, d1);
  hipLaunchKernelGGL(k_cg_init, sg1, dim3(256), 0, ctx->stream, st, d1, scale, n_b);
  for (z64* v : {psi, HP, SP})
    hipLaunchKernelGGL(k_scale_bands, eg, dim3(256), 0, ctx->stream, v, n_pw, scale,
                       (const BandState*)st);
  bdot(psi, HP, d1);
  bdot(psi, SP, d2);
  hipLaunchKernelGGL(k_cg_set_eps, sg1, dim3(256), 0, ctx->stream, st, d1, n_b);
  hipLaunchKernelGGL(k_residual, eg, dim3(256), 0, ctx->stream, HP, SP, n_pw, d1, d2, R,
                     (const BandState*)st);
  CHS_LAUNCH_CHECK(ctx);
  const z64 one = mk(1.0, 0.0), mone = mk(-1.0, 0.0);
  std::vector<BandState> hst(n_b);
  for (int step = 1; step <= max_steps; ++step) {
    CHS_TRY(precond(R, G, mlp, st));
    bdot(G, G, d1);
    bdot(R, R, d2);
    hipLaunchKernelGGL(k_cg_gguard, sg1, dim3(256), 0, ctx->stream, st, d1, d2, step, n_b);
    CHS_TRY(s_only(ctx, G, n_b, SG));
    // lower-band S-orthogonalisation against the handed-in block (real overlaps)
    CHS_TRY(zgemm(ctx, 'C', 'N', n_b, n_b, n_pw, one, CONV, n_pw, SG, n_pw, mk(0.0, 0.0), OV, n_b));
    hipLaunchKernelGGL(k_lower_real, dim3(ceil_div(n_b, 256), n_b), dim3(256), 0, ctx->stream, OV,
                       n_b, (const BandState*)st, full_overlap ? 1 : 0);
    CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, n_b, mone, CONV, n_pw, OV, n_b, one, G, n_pw));
    CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, n_b, mone, SCONV, n_pw, OV, n_b, one, SG, n_pw));
    bdot(G, SG, d1);
    hipLaunchKernelGGL(k_cg_gnorm, sg1, dim3(256), 0, ctx->stream, st, d1, scale, n_b);
    for (z64* v : {G, SG})
      hipLaunchKernelGGL(k_scale_bands, eg, dim3(256), 0, ctx->stream, v, n_pw, scale,
                         (const BandState*)st);
    // direction
    bdot(G, R, d1);
    hipLaunchKernelGGL(k_cg_dir_coef, sg1, dim3(256), 0, ctx->stream, st, d1, step, n_b, ca, cc);
    CHS_HIP(ctx, copy(D, DOLD));
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, G, D, n_pw, ca, cc,
                       (const BandState*)st);
    CHS_LAUNCH_CHECK(ctx);
    CHS_TRY(apply_h_impl(ctx, CHS_C128, D, n_b, HD, SD));
    // orthogonalise against the current band
    bdot(psi, SD, d1);
    hipLaunchKernelGGL(k_cg_ovcoef, sg1, dim3(256), 0, ctx->stream, st, d1, n_b, ca, cc);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, psi, D, n_pw, ca, cc,
                       (const BandState*)st);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, HP, HD, n_pw, ca, cc,
                       (const BandState*)st);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, SP, SD, n_pw, ca, cc,
                       (const BandState*)st);
    CHS_HIP(ctx, copy(DOLD, D));
    // residual guard + line search + update
    bdot(R, R, d1);
    bdot(HP, D, d2);
    bdot(HD, D, d3);
    bdot(SD, D, d4);
    hipLaunchKernelGGL(k_cg_linesearch, sg1, dim3(256), 0, ctx->stream, st, d1, d2, d3, d4, tol,
                       step, n_b, ca, cc);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, D, psi, n_pw, ca, cc,
                       (const BandState*)st);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, HD, HP, n_pw, ca, cc,
                       (const BandState*)st);
    hipLaunchKernelGGL(k_axpby_bands, eg, dim3(256), 0, ctx->stream, SD, SP, n_pw, ca, cc,
                       (const BandState*)st);
    bdot(psi, HP, d1);
    bdot(psi, SP, d2);
    hipLaunchKernelGGL(k_residual, eg, dim3(256), 0, ctx->stream, HP, SP, n_pw, d1, d2, R,
                       (const BandState*)st);
    bdot(R, R, d3);
    hipLaunchKernelGGL(k_cg_check, sg1, dim3(256), 0, ctx->stream, st, d1, d3, tol, step, max_steps,
                       n_b);
    CHS_LAUNCH_CHECK(ctx);
    CHS_HIP(ctx, hipMemcpyAsync(hst.data(), st, n_b * sizeof(BandState), hipMemcpyDeviceToHost,
                                ctx->stream));
    ctx->d2h += n_b * sizeof(BandState);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    bool any = false;
    for (const BandState& b : hst) any = any || b.active;
    if (!any) break;
  }
  CHS_HIP(ctx, hipMemcpyAsync(hst.data(), st, n_b * sizeof(BandState), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += n_b * sizeof(BandState);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < n_b; ++b) {
    if (hst[b].flag) return set_error(ctx, CHS_ERR_INVALID_ARG,
                                      "band %d: non-positive <psi|S|psi> (band_cg contract)", b);
    if (eps_out) eps_out[b] = hst[b].eps;
    if (steps_out) steps_out[b] = hst[b].n_steps;
    if (conv_out) conv_out[b] = hst[b].converged;
  }
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" int chs_band_cg(chs_context* ctx, double* d_psi, int n_b, int max_steps, double tol,
                           int precond, const float* d_mlp_scale, double* eps, int* n_steps,
                           int* converged) {
  const bool full = precond >= 0 && (precond & 256) != 0;
  if (precond >= 0) precond &= ~256;
  if (!ctx || !d_psi || n_b <= 0 || max_steps < 0 || !(tol > 0) || precond < 0 || precond > 2)
    return CHS_ERR_INVALID_ARG;
  if (precond == 2 && !d_mlp_scale) return CHS_ERR_INVALID_ARG;
  CHS_TRY(ensure_precision_ready(ctx, CHS_C128));
  Hamiltonian& H = ctx->impl->ham;
  BandCg cg;
  cg.full_overlap = full;
  CHS_TRY(cg.ensure_ws(ctx, H.n_pw, n_b));
  CHS_TRY(cg.setup_precond(precond >= 1));
  return cg.run((z64*)d_psi, max_steps, tol, precond == 2 ? d_mlp_scale : nullptr, eps, n_steps,
                converged);
}
