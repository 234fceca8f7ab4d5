// SCF driver on the device: ionic potential, Hartree + LDA potential,
// occupation-weighted density, total energy, Pulay/Kerker mixing.
//
// The reference's SCF loop (run_scf, src/scf.rs:358-376) stops at todo!()
// in build_v_eff (:157-159), construct_density (:275-279), check
// (:328-334) and DensityHistory::mix (src/mixing.rs:24-27). The glue here
// follows their specifications: PROJECT_ROOT_PLAN.md:568-611 (density),
// notes/plans/phase-2/TASKS.md:332-358 (Fermi-Dirac occupations on the
// host, integral of rho = N_e) and :386-440 (Pulay mixing, host solve of the
// small DIIS system), around the diagonalize transition (src/scf.rs:169-265)
// whose per-call setup (context, plans, uploads) is hoisted out of the loop.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "blas1.hpp"
#include "fft.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"
#include "scf_dev.hpp"

namespace chs {

// ---------------------------------------------------------------------------
// grid helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double grid_g2(int ix, int iy, int iz, int n, double b, bool* nyq) {
  const int h = ix <= n / 2 ? ix : ix - n;
  const int k = iy <= n / 2 ? iy : iy - n;
  const int l = iz <= n / 2 ? iz : iz - n;
  *nyq = (n % 2 == 0) && (ix == n / 2 || iy == n / 2 || iz == n / 2);
  return b * b * (double)(h * h + k * k + l * l);
}

// V_ion(G) = sum_a v(G) e^{-iG.R_a}; v(G) = -4 pi Z exp(-G^2 rc^2/4)/(Omega G^2),
// v(0) = pi Z rc^2 / Omega; Nyquist planes zero. ph: per-atom 1D phase
// tables [atom][axis][i] = e^{-i b f(i) x_a}.
__global__ void k_vion_g(z64* __restrict__ g, int n, int n_atoms, const z64* __restrict__ ph,
                         double b, double z, double rc, double omega) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n3 = (size_t)n * n * n;
  if (id >= n3) return;
  const int ix = (int)(id % n), iy = (int)((id / n) % n), iz = (int)(id / ((size_t)n * n));
  bool nyq;
  const double g2 = grid_g2(ix, iy, iz, n, b, &nyq);
  if (nyq) {
    g[id] = mk(0.0, 0.0);
    return;
  }
  double v;
  if (ix == 0 && iy == 0 && iz == 0)
    v = M_PI * z * rc * rc / omega;
  else
    v = -4.0 * M_PI * z * exp(-g2 * rc * rc / 4.0) / (omega * g2);
  z64 s = mk(0.0, 0.0);
  for (int a = 0; a < n_atoms; ++a) {
    const z64* pa = ph + (size_t)a * 3 * n;
    s = s + pa[ix] * pa[n + iy] * pa[2 * n + iz];
  }
  g[id] = mk(v * s.x, v * s.y);
}

__global__ void k_real_part(const z64* __restrict__ g, double* __restrict__ r, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = g[i].x;
}

__global__ void k_to_complex(const double* __restrict__ r, z64* __restrict__ g, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = mk(r[i], 0.0);
}

// rho(G) = g / N^3; V_H(G) = 4 pi rho(G)/G^2 (G != 0, not Nyquist); partial
// sums of E_H = Omega/2 sum 4 pi |rho(G)|^2 / G^2 into part[block].
__global__ void __launch_bounds__(256) k_hartree_g(z64* __restrict__ g, int n, double b,
                                                   double omega, double* __restrict__ part) {
  __shared__ double sh[8];
  const size_t n3 = (size_t)n * n * n;
  const double inv = 1.0 / (double)n3;
  double e = 0;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n3;
       id += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(id % n), iy = (int)((id / n) % n), iz = (int)(id / ((size_t)n * n));
    bool nyq;
    const double g2 = grid_g2(ix, iy, iz, n, b, &nyq);
    if (nyq || g2 <= 0.0) {
      g[id] = mk(0.0, 0.0);
      continue;
    }
    const z64 rg = mk(g[id].x * inv, g[id].y * inv);
    e += 0.5 * omega * 4.0 * M_PI * norm2(rg) / g2;
    const double f = 4.0 * M_PI / g2;
    g[id] = mk(f * rg.x, f * rg.y);
  }
  const double s = block_sum(e, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[gridDim.x + blockIdx.x] = 0.0;
  }
}

// Kerker K(G) = G^2/(G^2 + q^2) with the 1/N^3 of the round trip; G = 0 and
// Nyquist planes dropped.
__global__ void k_kerker_g(z64* __restrict__ g, int n, double b, double q2) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n3 = (size_t)n * n * n;
  if (id >= n3) return;
  const int ix = (int)(id % n), iy = (int)((id / n) % n), iz = (int)(id / ((size_t)n * n));
  bool nyq;
  const double g2 = grid_g2(ix, iy, iz, n, b, &nyq);
  if (nyq || g2 <= 0.0) {
    g[id] = mk(0.0, 0.0);
    return;
  }
  const double f = g2 / (g2 + q2) / (double)n3;
  g[id] = mk(f * g[id].x, f * g[id].y);
}

// LDA: Slater exchange + Perdew-Zunger 1981 correlation (unpolarised).
__device__ __forceinline__ void lda_pz(double rho, double* exc, double* vxc) {
  if (rho <= 1e-14) {
    *exc = 0;
    *vxc = 0;
    return;
  }
  const double cx = 0.75 * cbrt(3.0 / M_PI);
  const double ex = -cx * cbrt(rho);
  const double vx = 4.0 / 3.0 * ex;
  const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
  double ec, vc;
  if (rs >= 1.0) {
    const double gg = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sq = sqrt(rs);
    const double den = 1.0 + b1 * sq + b2 * rs;
    ec = gg / den;
    vc = ec * (1.0 + 7.0 / 6.0 * b1 * sq + 4.0 / 3.0 * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  *exc = ex + ec;
  *vxc = vx + vc;
}

// vhxc = V_H + V_xc(rho); veff = vion + vhxc (veff may be null). Partial
// sums: part[b] = sum rho e_xc, part[nb + b] = sum vin_hxc * rho (vin may
// be null).
__global__ void __launch_bounds__(256) k_vxc(const double* __restrict__ rho,
                                             const z64* __restrict__ vh_c,
                                             const double* __restrict__ vion,
                                             double* __restrict__ vhxc, double* __restrict__ veff,
                                             const double* __restrict__ vin, size_t n3,
                                             double* __restrict__ part) {
  __shared__ double sh[8];
  double se = 0, sd = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (size_t)gridDim.x * blockDim.x) {
    double e, vx;
    const double r = rho[i];
    lda_pz(r, &e, &vx);
    const double v = vh_c[i].x + vx;
    if (vhxc) vhxc[i] = v;
    if (veff) veff[i] = vion[i] + v;
    se += r * e;
    if (vin) sd += vin[i] * r;
  }
  const double a = block_sum(se, sh);
  const double c = block_sum(sd, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = c;
  }
}

__global__ void __launch_bounds__(256) k_dot_real(const double* __restrict__ a,
                                                  const double* __restrict__ b, size_t n,
                                                  double* __restrict__ part) {
  __shared__ double sh[8];
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r;
    part[gridDim.x + blockIdx.x] = 0.0;
  }
}

// out = sum_i c[i] * src_i for up to 8 source arrays (pointer table).
struct Src8 {
  const double* p[8];
  double c[8];
  int k;
};
__global__ void k_lincomb(Src8 s, double* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0;
  for (int j = 0; j < s.k; ++j) v += s.c[j] * s.p[j][i];
  out[i] = v;
}

__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ o, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = a[i] - b[i];
}

__global__ void k_axpy_re(double alpha, const z64* __restrict__ x, const double* __restrict__ y,
                          double* __restrict__ o, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = y[i] + alpha * x[i].x;
}

__global__ void k_fill(double* __restrict__ a, double v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(256) k_abs_diff(const double* __restrict__ a,
                                                  const double* __restrict__ b, size_t n,
                                                  double* __restrict__ part) {
  __shared__ double sh[8];
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += fabs(a[i] - b[i]);
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r;
    part[gridDim.x + blockIdx.x] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// reductions to host scalars
// ---------------------------------------------------------------------------
int reduce2(chs_context* ctx, double* out2) {
  auto* im = ctx->impl;
  double* part = (double*)im->part.p;
  double* dst = (double*)im->small.p + 40;
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kRedThreads), 0, ctx->stream, part, kRedBlocks,
                     dst);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(out2, dst, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->d2h += 2 * sizeof(double);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

static int prep_reduce(chs_context* ctx) {
  CHS_TRY(ensure(ctx, ctx->impl->part, sizeof(double) * 2 * kRedBlocks + 64));
  CHS_TRY(ensure(ctx, ctx->impl->small, 64 * sizeof(double)));
  return CHS_OK;
}

double dot_real(chs_context* ctx, const double* a, const double* b, size_t n, int* rc) {
  *rc = prep_reduce(ctx);
  if (*rc) return 0;
  hipLaunchKernelGGL(k_dot_real, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, a, b, n,
                     (double*)ctx->impl->part.p);
  double o[2] = {0, 0};
  *rc = reduce2(ctx, o);
  return o[0];
}

double l1_diff(chs_context* ctx, const double* a, const double* b, size_t n, int* rc) {
  *rc = prep_reduce(ctx);
  if (*rc) return 0;
  hipLaunchKernelGGL(k_abs_diff, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, a, b, n,
                     (double*)ctx->impl->part.p);
  double o[2] = {0, 0};
  *rc = reduce2(ctx, o);
  return o[0];
}

// ---------------------------------------------------------------------------
// Grid physics on the device (GridOps)
// ---------------------------------------------------------------------------
int GridOps::init(chs_context* c, int n_, double L, double omega_) {
  ctx = c;
  n = n_;
  n3 = (size_t)n * n * n;
  b = 2.0 * M_PI / L;
  omega = omega_;
  CHS_TRY(ensure(ctx, cgrid, n3 * sizeof(z64)));
  return prep_reduce(ctx);
}

// Per-atom 1D structure-phase tables (host cos/sin, one upload per geometry).
int GridOps::set_atoms(const double* pos, int n_at) {
  n_atoms = n_at;
  std::vector<double> h((size_t)n_atoms * 3 * n * 2);
  for (int a = 0; a < n_atoms; ++a)
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < n; ++i) {
        const int f = i <= n / 2 ? i : i - n;
        const double t = -b * f * pos[3 * a + c];
        h[(((size_t)a * 3 + c) * n + i) * 2] = std::cos(t);
        h[(((size_t)a * 3 + c) * n + i) * 2 + 1] = std::sin(t);
      }
  CHS_TRY(ensure(ctx, ph, h.size() * sizeof(double)));
  CHS_HIP(ctx, hipMemcpyAsync(ph.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->h2d += h.size() * sizeof(double);
  return CHS_OK;
}

// V_ion on the grid from the atom phase tables.
int GridOps::vion(double zval, double rc, double* d_vion) {
  hipLaunchKernelGGL(k_vion_g, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, (z64*)cgrid.p,
                     n, n_atoms, (const z64*)ph.p, b, zval, rc, omega);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid.p, n, 1, +1));
  hipLaunchKernelGGL(k_real_part, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                     (const z64*)cgrid.p, d_vion, n3);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// V_Hxc[rho] into vhxc (and veff = vion + vhxc); returns E_H, E_xc and
// dc = int vin * rho (when vin != null).
int GridOps::vhxc(const double* rho, const double* vion, double* vhxc_out, double* veff,
                  const double* vin, double* eh, double* exc, double* dc) {
  hipLaunchKernelGGL(k_to_complex, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, rho,
                     (z64*)cgrid.p, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid.p, n, 1, -1));
  hipLaunchKernelGGL(k_hartree_g, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                     (z64*)cgrid.p, n, b, omega, (double*)ctx->impl->part.p);
  CHS_LAUNCH_CHECK(ctx);
  double o[2];
  CHS_TRY(reduce2(ctx, o));
  *eh = o[0];
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid.p, n, 1, +1));
  hipLaunchKernelGGL(k_vxc, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, rho,
                     (const z64*)cgrid.p, vion, vhxc_out, veff, vin, n3,
                     (double*)ctx->impl->part.p);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(reduce2(ctx, o));
  *exc = o[0] * omega / (double)n3;
  if (dc) *dc = o[1] * omega / (double)n3;
  return CHS_OK;
}

// f <- K f (Kerker filter, real grid function)
int GridOps::kerker(double* f, double q) {
  hipLaunchKernelGGL(k_to_complex, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, f,
                     (z64*)cgrid.p, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid.p, n, 1, -1));
  hipLaunchKernelGGL(k_kerker_g, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                     (z64*)cgrid.p, n, b, q * q);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d(ctx, CHS_C128, cgrid.p, n, 1, +1));
  hipLaunchKernelGGL(k_real_part, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                     (const z64*)cgrid.p, f, n3);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// ---------------------------------------------------------------------------
// Pulay (DIIS) mixing with Kerker-filtered residuals, history on the device.
// ---------------------------------------------------------------------------
static bool solve_dense(std::vector<double>& a, std::vector<double>& b, int n) {
  double amax = 0;
  for (double v : a) amax = std::max(amax, std::abs(v));
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::abs(a[r * n + col]) > std::abs(a[piv * n + col])) piv = r;
    if (std::abs(a[piv * n + col]) <= 1e-14 * amax) return false;
    if (piv != col) {
      for (int j = 0; j < n; ++j) std::swap(a[col * n + j], a[piv * n + j]);
      std::swap(b[col], b[piv]);
    }
    for (int r = col + 1; r < n; ++r) {
      const double f = a[r * n + col] / a[col * n + col];
      for (int j = col; j < n; ++j) a[r * n + j] -= f * a[col * n + j];
      b[r] -= f * b[col];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int j = r + 1; j < n; ++j) s -= a[r * n + j] * b[j];
    b[r] = s / a[r * n + r];
  }
  return true;
}

int PulayDev::init(chs_context* c, GridOps* g_, int depth_, double alpha_, double q_) {
  ctx = c;
  g = g_;
  depth = std::max(1, std::min(depth_, 8));  // k_lincomb takes up to 8 sources
  alpha = alpha_;
  q = q_;
  CHS_TRY(ensure(ctx, mem, (size_t)(2 * depth + 2) * g->n3 * sizeof(double)));
  slots.clear();
  gram.assign((size_t)depth * depth, 0.0);
  return CHS_OK;
}

double* PulayDev::rin(int slot) { return (double*)mem.p + (size_t)(2 * slot) * g->n3; }
double* PulayDev::res(int slot) { return (double*)mem.p + (size_t)(2 * slot + 1) * g->n3; }

// Same algorithm as the oracle's Pulay::mix: append (rho_in, R), solve the
// bordered DIIS system (dropping the oldest pair while singular), then
// rho_next = sum c rho_in + alpha K[sum c R].
int PulayDev::mix(const double* rho_in, const double* rho_out, double* rho_next) {
  const size_t n3 = g->n3;
  // choose the slot for the new pair: a free one or the oldest
  int slot;
  if ((int)slots.size() < depth) {
    slot = (int)slots.size();
    std::vector<int> used(slots);
    std::sort(used.begin(), used.end());
    for (int s = 0; s < depth; ++s)
      if (!std::binary_search(used.begin(), used.end(), s)) {
        slot = s;
        break;
      }
  } else {
    slot = slots.front();
    slots.erase(slots.begin());
  }
  CHS_HIP(ctx, hipMemcpyAsync(rin(slot), rho_in, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
  hipLaunchKernelGGL(k_sub, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, rho_out, rho_in,
                     res(slot), n3);
  CHS_LAUNCH_CHECK(ctx);
  slots.push_back(slot);
  // Gram row of the new residual against the history
  for (int s : slots) {
    int rc = 0;
    const double v = dot_real(ctx, res(slot), res(s), n3, &rc);
    if (rc) return rc;
    gram[(size_t)slot * depth + s] = v;
    gram[(size_t)s * depth + slot] = v;
  }
  std::vector<double> c;
  std::vector<int> use(slots);
  for (;;) {
    const int k = (int)use.size();
    std::vector<double> a((size_t)(k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) a[i * (k + 1) + j] = gram[(size_t)use[i] * depth + use[j]];
    for (int i = 0; i < k; ++i) {
      a[i * (k + 1) + k] = 1.0;
      a[k * (k + 1) + i] = 1.0;
    }
    rhs[k] = 1.0;
    if (solve_dense(a, rhs, k + 1)) {
      c.assign(rhs.begin(), rhs.begin() + k);
      break;
    }
    use.erase(use.begin());
    slots.erase(slots.begin());
    if (use.empty()) {
      use.push_back(slot);
      slots.push_back(slot);
      c.assign(1, 1.0);
      break;
    }
  }
  // rho_opt -> rho_next, R_opt -> scratch slot `depth`
  double* ropt = rin(depth);
  Src8 s1{}, s2{};
  s1.k = s2.k = (int)use.size();
  for (size_t i = 0; i < use.size(); ++i) {
    s1.p[i] = rin(use[i]);
    s1.c[i] = c[i];
    s2.p[i] = res(use[i]);
    s2.c[i] = c[i];
  }
  hipLaunchKernelGGL(k_lincomb, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, s1, rho_next,
                     n3);
  hipLaunchKernelGGL(k_lincomb, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, s2, ropt, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(g->kerker(ropt, q));
  // rho_next += alpha * K R_opt
  hipLaunchKernelGGL(k_axpy_re, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, alpha,
                     (const z64*)g->cgrid.p, rho_next, rho_next, n3);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// ---------------------------------------------------------------------------
// Fermi-Dirac occupations on the host (the reference's plan does this step
// on the CPU: notes/plans/phase-2/TASKS.md:342-343).
// ---------------------------------------------------------------------------
double fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ,
                         double* ts) {
  *ts = 0;
  if (kt <= 0) {
    const int n_occ = (int)std::lround(n_e / 2.0);
    for (int i = 0; i < n; ++i) occ[i] = i < n_occ ? 2.0 : 0.0;
    return n_occ > 0 && n_occ <= n ? eps[n_occ - 1] : 0.0;
  }
  auto fd = [&](double e, double mu) {
    const double x = (e - mu) / kt;
    if (x > 200) return 0.0;
    if (x < -200) return 2.0;
    return 2.0 / (1.0 + std::exp(x));
  };
  double lo = eps[0] - 50 * kt, hi = eps[n - 1] + 50 * kt;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    double s = 0;
    for (int i = 0; i < n; ++i) s += fd(eps[i], mid);
    const double nlo = s > n_e ? lo : mid, nhi = s > n_e ? mid : hi;
    // a step that changes neither end is a fixed point: the remaining
    // iterations would repeat it, so stopping gives the same mu
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  const double mu = 0.5 * (lo + hi);
  double t = 0;
  for (int i = 0; i < n; ++i) {
    occ[i] = fd(eps[i], mu);
    const double f = occ[i] / 2.0;
    if (f > 1e-300 && f < 1.0 - 1e-16) t += f * std::log(f) + (1.0 - f) * std::log(1.0 - f);
  }
  *ts = 2.0 * kt * t;
  return mu;
}

// Occupations of states with weights w_i (the k-point weight of each
// state): sum_i w_i f_i = n_e, -TS = 2 kT sum_i w_i [f ln f + (1-f) ln(1-f)]
// (f = f_i / 2), mu by bisection (notes/plans/phase-2/TASKS.md:342-343).
// kT = 0: states filled in ascending order, two electrons of weight w_i each.
double fermi_occupations_w(const double* eps, const double* w, int n, double n_e, double kt,
                           double* occ, double* ts) {
  *ts = 0;
  double emin = eps[0], emax = eps[0];
  for (int i = 1; i < n; ++i) {
    emin = std::min(emin, eps[i]);
    emax = std::max(emax, eps[i]);
  }
  if (kt <= 0) {
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return eps[a] < eps[b]; });
    double left = n_e, mu = emin;
    for (int i : ord) {
      const double take = std::min(2.0, std::max(0.0, left / w[i]));
      occ[i] = take;
      left -= take * w[i];
      if (take > 0) mu = eps[i];
    }
    return mu;
  }
  auto fd = [&](double e, double mu) {
    const double x = (e - mu) / kt;
    if (x > 200) return 0.0;
    if (x < -200) return 2.0;
    return 2.0 / (1.0 + std::exp(x));
  };
  double lo = emin - 50 * kt, hi = emax + 50 * kt;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    double s = 0;
    for (int i = 0; i < n; ++i) s += w[i] * fd(eps[i], mid);
    const double nlo = s > n_e ? lo : mid, nhi = s > n_e ? mid : hi;
    if (nlo == lo && nhi == hi) break;  // fixed point: later steps repeat it
    lo = nlo;
    hi = nhi;
  }
  const double mu = 0.5 * (lo + hi);
  double t = 0;
  for (int i = 0; i < n; ++i) {
    occ[i] = fd(eps[i], mu);
    const double f = occ[i] / 2.0;
    if (f > 1e-300 && f < 1.0 - 1e-16) t += w[i] * (f * std::log(f) + (1.0 - f) * std::log(1.0 - f));
  }
  *ts = 2.0 * kt * t;
  return mu;
}

}  // namespace chs
