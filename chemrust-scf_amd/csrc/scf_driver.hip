// chs_scf_run: the SCF loop around the eigensolver hot path (run_scf,
// reference src/scf.rs:358-376; diagonalize, :169-265).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <thread>
#include <vector>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"
#include "projectors.hpp"
#include "scf_dev.hpp"

namespace chs {

__global__ void k_z2c(const z64* __restrict__ a, c32* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mk((float)a[i].x, (float)a[i].y);
}
__global__ void k_c2z(const c32* __restrict__ a, z64* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = mk((double)a[i].x, (double)a[i].y);
}
__global__ void k_fill_d(double* __restrict__ a, double v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// The reference's 48-bit LCG step s -> (a s + c) mod 2^48
// (src/eigensolver/subspace_diag.rs:434-447) advanced by `steps` draws at
// once: the affine map composed by repeated squaring (ring arithmetic mod 2^64
// reduces exactly mod 2^48).
static uint64_t lcg_jump(uint64_t s, uint64_t steps) {
  uint64_t A = 1, C = 0;                      // accumulated map x -> A x + C
  uint64_t a = 25214903917ull, c = 11ull;     // map of 2^j steps
  while (steps) {
    if (steps & 1) {
      A = a * A;
      C = a * C + c;
    }
    c = a * c + c;
    a = a * a;
    steps >>= 1;
  }
  return (A * s + C) & 0xFFFFFFFFFFFFull;
}

// Seeded start psi_{b,g} = (x + i y)/(1 + |k+G|^2), x, y ~ N(0,1) from the
// reference's LCG (src/eigensolver/subspace_diag.rs:434-447) with
// Box-Muller, drawn band-major then plane-wave order. Every element takes
// four draws, so host threads start their ranges at draw 4 e (lcg_jump) and
// produce the sequential stream exactly.
// psi: 2 n_b n_pw doubles, left uninitialised by the caller, so that the
// pages are first touched by the threads that fill them.
static void initial_psi_host(const std::vector<double>& kin, int n_b, uint64_t seed,
                             double* psi) {
  const size_t n_pw = kin.size();
  const size_t ne = (size_t)n_b * n_pw;
  auto fill = [&](size_t e0, size_t e1) {
    uint64_t s = lcg_jump(seed, 4 * (uint64_t)e0);
    auto uni = [&s]() {
      s = (s * 25214903917ull + 11ull) & 0xFFFFFFFFFFFFull;
      return (double)s / 281474976710656.0;
    };
    auto nrm = [&]() {
      const double u1 = 1.0 - uni();
      const double u2 = uni();
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    };
    for (size_t e = e0; e < e1; ++e) {
      const double x = nrm();
      const double y = nrm();
      const double d = 1.0 + 2.0 * kin[e % n_pw];
      psi[2 * e] = x / d;
      psi[2 * e + 1] = y / d;
    }
  };
  const size_t nt = std::max<size_t>(1, std::min<size_t>({16, std::thread::hardware_concurrency(),
                                                            ne / 65536 + 1}));
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(fill, ne * t / nt, ne * (t + 1) / nt);
  fill(0, ne / nt);
  for (auto& th : pool) th.join();
}

struct ScfState {
  DevBuf miller, idx, kin, vion, veff, vhxc, rho_in, rho_out, rho_next, psi, psi32;
  DevBuf pos, beta, d0, dscr, q, occ, wproj;  // projector part
  ~ScfState() {
    for (DevBuf* b : {&miller, &idx, &kin, &vion, &veff, &vhxc, &rho_in, &rho_out, &rho_next, &psi,
                      &psi32, &pos, &beta, &d0, &dscr, &q, &occ, &wproj})
      release(*b);
  }
};

// D0 ~ U[d_min, d_max] from the 48-bit LCG (seed_d), one value per projector
static std::vector<double> seeded_d0(const chs_scf_params* p, int P) {
  std::vector<double> d0(P);
  uint64_t s = p->seed_d;
  for (int i = 0; i < P; ++i) {
    s = (s * 25214903917ull + 11ull) & 0xFFFFFFFFFFFFull;
    d0[i] = p->d_min + (p->d_max - p->d_min) * ((double)s / 281474976710656.0);
  }
  return d0;
}

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int scf_run_impl(chs_context* ctx, const chs_scf_params* p, const double* pos, chs_scf_result* res,
                 double* energies, double* eps_out, double* rho_host) {
  const int np = p->n_proj_per_atom;
  if (np != 0 && np != 1 && np != 4)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "projectors per atom must be 0, 1 or 4");
  ScfState S;
  const int n = p->n_grid, n_b = p->n_bands;
  const double L = p->n_rep * p->a0, omega = L * L * L;
  const int n_atoms = 8 * p->n_rep * p->n_rep * p->n_rep;
  const double n_e = p->zval * n_atoms;
  const size_t n3 = (size_t)n * n * n;
  // plane waves
  int n_pw = 0;
  CHS_TRY(chs_gsphere(ctx, L, p->ecut, p->kfrac, n, 0, &n_pw, nullptr, nullptr, nullptr));
  CHS_TRY(ensure(ctx, S.miller, 3 * (size_t)n_pw * sizeof(int32_t)));
  CHS_TRY(ensure(ctx, S.idx, (size_t)n_pw * sizeof(int32_t)));
  CHS_TRY(ensure(ctx, S.kin, (size_t)n_pw * sizeof(double)));
  CHS_TRY(chs_gsphere(ctx, L, p->ecut, p->kfrac, n, n_pw, &n_pw, (int32_t*)S.miller.p,
                      (int32_t*)S.idx.p, (double*)S.kin.p));
  if (n_b > n_pw) return set_error(ctx, CHS_ERR_INVALID_ARG, "more bands than plane waves");
  // grids
  for (DevBuf* b : {&S.vion, &S.veff, &S.vhxc, &S.rho_in, &S.rho_out, &S.rho_next})
    CHS_TRY(ensure(ctx, *b, n3 * sizeof(double)));
  GridOps G;
  CHS_TRY(G.init(ctx, n, L, omega));
  CHS_TRY(G.set_atoms(pos, n_atoms));
  CHS_TRY(G.vion(p->zval, p->rc_loc, (double*)S.vion.p));
  // start wavefunctions (host generator, one upload)
  std::vector<double> kin(n_pw);
  CHS_HIP(ctx, hipMemcpy(kin.data(), S.kin.p, n_pw * sizeof(double), hipMemcpyDeviceToHost));
  ctx->d2h += n_pw * sizeof(double);
  const size_t ne = (size_t)n_b * n_pw;
  std::unique_ptr<double[]> h_psi(new double[2 * ne]);
  initial_psi_host(kin, n_b, p->seed_psi, h_psi.get());
  CHS_TRY(ensure(ctx, S.psi, ne * sizeof(z64)));
  CHS_HIP(ctx, hipMemcpy(S.psi.p, h_psi.get(), ne * sizeof(z64), hipMemcpyHostToDevice));
  ctx->h2d += ne * sizeof(z64);
  const bool mixed = p->precision == CHS_C64;
  if (mixed) CHS_TRY(ensure(ctx, S.psi32, ne * sizeof(c32)));
  CHS_TRY(chs_bind_hamiltonian(ctx, n, n_pw, (const int32_t*)S.miller.p, (const double*)S.kin.p,
                               (const double*)S.veff.p));
  // USPP projectors: beta on the device, D0 ~ U[d_min, d_max] from the LCG
  // (seed_d, atom-major), Q = q0 on every projector
  const int P = np * n_atoms;
  if (P > 0) {
    CHS_TRY(ensure(ctx, S.pos, 3 * (size_t)n_atoms * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.pos.p, pos, 3 * (size_t)n_atoms * sizeof(double),
                           hipMemcpyHostToDevice));
    CHS_TRY(ensure(ctx, S.beta, (size_t)P * n_pw * sizeof(z64)));
    CHS_TRY(build_beta(ctx, (const int32_t*)S.miller.p, n_pw, L, p->kfrac,
                       (const double*)S.pos.p, n_atoms, np, p->beta_rc, (z64*)S.beta.p));
    std::vector<double> d0 = seeded_d0(p, P), q(P, p->q0);
    for (DevBuf* b : {&S.d0, &S.dscr, &S.q}) CHS_TRY(ensure(ctx, *b, P * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.d0.p, d0.data(), P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(S.dscr.p, d0.data(), P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(S.q.p, q.data(), P * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += 3 * (size_t)n_atoms * sizeof(double) + 3 * (size_t)P * sizeof(double);
    CHS_TRY(chs_bind_projectors(ctx, P, (const double*)S.beta.p, (const double*)S.dscr.p,
                                (const double*)S.q.p));
    CHS_TRY(ensure(ctx, S.wproj, (size_t)P * n_b * sizeof(z64)));
  }
  // launch shapes of the H|psi> kernels, tuned once per (grid, precision)
  // and context (later runs reuse them)
  if (ctx->impl->hpsi_autotune)
    for (int prec : {CHS_C128, CHS_C64}) {
      if (prec == CHS_C64 && !mixed) continue;
      if (ctx->impl->vloc_tune.count(std::make_pair(n, prec))) continue;
      CHS_TRY(chs_autotune_hpsi(ctx, prec, n_b, nullptr, nullptr, nullptr, nullptr));
    }
  CHS_TRY(ensure(ctx, S.occ, (size_t)n_b * sizeof(double)));
  PulayDev mixer;
  CHS_TRY(mixer.init(ctx, &G, std::max(1, p->pulay_depth), p->mix_alpha, p->kerker_q));
  hipLaunchKernelGGL(k_fill_d, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                     (double*)S.rho_in.p, n_e / omega, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));

  std::vector<double> eps, occ(n_b, 0.0);
  double e_prev = 0, mu = 0, etot = 0;
  int small = 0, it = 0, converged = 0;
  bool fp64_phase = !mixed;
  int fp64_iters = 0, fp32_iters = 0;
  // complex64 filter until |dE|/atom < fp32_until (default 2 etol, but not
  // below 3e-7 Ha, near the energy noise the complex64 iterates leave), or
  // until |dE| stalls (> 0.9 of the previous one) within 10 fp32_until;
  // convergence is then decided on FP64 energies only
  const double fp32_until = p->fp32_until > 0 ? p->fp32_until : std::max(2.0 * p->etol, 3e-7);
  double de_prev = INFINITY;
  const double t0 = now();
  double t_filter = 0, t_rr = 0, t_den = 0, t_veff = 0, t_lan = 0;
  for (it = 1; it <= p->max_iter; ++it) {
    double ta = now();
    double eh_in, exc_in;
    CHS_TRY(G.vhxc((const double*)S.rho_in.p, (const double*)S.vion.p, (double*)S.vhxc.p,
                   (double*)S.veff.p, nullptr, &eh_in, &exc_in, nullptr));
    CHS_TRY(chs_set_veff(ctx, (const double*)S.veff.p));
    if (P > 0 && p->q0 != 0.0)
      CHS_TRY(screen_dion(ctx, (const double*)S.veff.p, n, L, n_atoms, np, p->q0, p->q_rc,
                          (const z64*)G.ph.p, (const double*)S.d0.p, (z64*)G.cgrid.p,
                          (double*)S.dscr.p));
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double tb = now();
    t_veff += tb - ta;
    double lmax;
    if (p->lanczos_k > 0) {
      // complex64 filter phase: the bound's projector products on the
      // complex64 beta the filter itself uses
      CHS_TRY(lanczos_impl(ctx, p->lanczos_k, p->seed_lanczos, &lmax, nullptr,
                           mixed && !fp64_phase));
    } else {
      CHS_TRY(chs_grid_lambda_max(ctx, L, &lmax));
    }
    ta = now();
    t_lan += ta - tb;
    const int passes = it == 1 ? std::max(1, p->first_passes) : 1;
    const bool use32 = mixed && !fp64_phase;
    if (use32) ++fp32_iters;
    for (int ps = 0; ps < passes; ++ps) {
      double b4[4];
      chs_spectral_bounds(eps.data(), (int)eps.size(), lmax, b4);
      double tf0 = now();
      if (use32) {
        hipLaunchKernelGGL(k_z2c, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                           (const z64*)S.psi.p, (c32*)S.psi32.p, ne);
        CHS_LAUNCH_CHECK(ctx);
        CHS_TRY(chebyshev_filter_impl(ctx, CHS_C64, S.psi32.p, n_b, p->cheb_degree, b4, S.psi32.p));
        hipLaunchKernelGGL(k_c2z, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                           (const c32*)S.psi32.p, (z64*)S.psi.p, ne);
        CHS_LAUNCH_CHECK(ctx);
      } else {
        CHS_TRY(chebyshev_filter_impl(ctx, CHS_C128, S.psi.p, n_b, p->cheb_degree, b4, S.psi.p));
      }
      double tf1 = now();
      t_filter += tf1 - tf0;
      eps.resize(n_b);
      // the last pass also hands over beta^H psi of the rotated block for
      // the augmentation charge
      // in the complex64 phase H|Y> and beta^H Y come from the filter's
      // complex64 output (the Gram matrices and the eigensolve stay FP64)
      CHS_TRY(rayleigh_ritz_impl(ctx, (z64*)S.psi.p, n_b, eps.data(), nullptr, nullptr,
                                 P > 0 && ps + 1 == passes ? (z64*)S.wproj.p : nullptr,
                                 use32 ? (const c32*)S.psi32.p : nullptr));
      t_rr += now() - tf1;
    }
    tb = now();
    double ts = 0;
    mu = fermi_occupations(eps.data(), n_b, n_e, p->smear_kt, occ.data(), &ts);
    if (use32 && S.psi32.p && ctx->impl->ham.st.rs_win > 0) {
      // complex64 phase: the density from the rotated block rounded to
      // complex64 (FP64 accumulation), like the filter that produced it
      hipLaunchKernelGGL(k_z2c, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                         (const z64*)S.psi.p, (c32*)S.psi32.p, ne);
      CHS_LAUNCH_CHECK(ctx);
      CHS_TRY(density_c64(ctx, (const c32*)S.psi32.p, n_b, occ.data(), omega,
                          (double*)S.rho_out.p));
    } else {
      CHS_TRY(density_impl(ctx, (const z64*)S.psi.p, n_b, occ.data(), omega,
                           (double*)S.rho_out.p));
    }
    if (P > 0 && p->q0 != 0.0) {
      CHS_HIP(ctx, hipMemcpyAsync(S.occ.p, occ.data(), n_b * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream));
      CHS_TRY(augment_density(ctx, (const z64*)S.psi.p, n_b, (const double*)S.occ.p, n_atoms, np,
                              p->q0, p->q_rc, (const z64*)G.ph.p, n, L, (z64*)G.cgrid.p,
                              (double*)S.rho_out.p, (const z64*)S.wproj.p));
    }
    double eh, exc, dc;
    CHS_TRY(G.vhxc((const double*)S.rho_out.p, (const double*)S.vion.p, nullptr, nullptr,
                   (const double*)S.vhxc.p, &eh, &exc, &dc));
    double eband = 0;
    for (int i = 0; i < n_b; ++i) eband += occ[i] * eps[i];
    etot = eband - dc + eh + exc + ts;
    if (energies && it <= p->max_iter) energies[it - 1] = etot;
    if (fp64_phase) ++fp64_iters;
    if (it > 1) {
      const double de = std::abs(etot - e_prev) / n_atoms;
      small = de < p->etol ? small + 1 : 0;
      if (!fp64_phase && (de < fp32_until || (de < 10.0 * fp32_until && de > 0.9 * de_prev))) {
        fp64_phase = true;  // the FP32 filter has done its part: finish in FP64
        small = 0;
      }
      de_prev = de;
    }
    e_prev = etot;
    t_den += now() - tb;
    if (small >= 2 && fp64_phase && fp64_iters >= (mixed ? std::max(1, p->fp64_final) : 0)) {
      converged = 1;
      break;
    }
    ta = now();
    CHS_TRY(mixer.mix((const double*)S.rho_in.p, (const double*)S.rho_out.p,
                      (double*)S.rho_next.p));
    std::swap(S.rho_in, S.rho_next);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    t_den += now() - ta;
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double t1 = now();
  res->n_iter = std::min(it, p->max_iter);
  res->converged = converged;
  res->n_pw = n_pw;
  res->n_proj = P;
  res->energy = etot;
  res->mu = mu;
  res->seconds = t1 - t0;
  res->t_filter = t_filter;
  res->t_rr = t_rr;
  res->t_density = t_den;
  res->t_veff = t_veff;
  res->t_lanczos = t_lan;
  res->n_iter_fp32 = fp32_iters;
  if (eps_out) std::copy(eps.begin(), eps.end(), eps_out);
  if (rho_host) {
    CHS_HIP(ctx, hipMemcpy(rho_host, S.rho_out.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
    ctx->d2h += n3 * sizeof(double);
  }
  ctx->impl->ham.bound = false;  // the bound arrays die with S
  return converged ? CHS_OK : CHS_ERR_NOT_CONVERGED;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_density(chs_context* ctx, const double* d_psi, int n_b, const double* occ, double omega,
                double* d_rho) {
  if (!ctx || !d_psi || n_b <= 0 || !occ || !(omega > 0) || !d_rho) return CHS_ERR_INVALID_ARG;
  return density_impl(ctx, (const z64*)d_psi, n_b, occ, omega, d_rho);
}

int chs_scf_run(chs_context* ctx, const chs_scf_params* p, const double* pos, chs_scf_result* res,
                double* energies, double* eps, double* rho) {
  if (!ctx || !p || !pos || !res) return CHS_ERR_INVALID_ARG;
  if (p->n_rep <= 0 || p->n_bands <= 0 || p->max_iter <= 0 || p->cheb_degree < 0 ||
      (p->precision != CHS_C128 && p->precision != CHS_C64))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "bad SCF parameters");
  *res = chs_scf_result{};
  return scf_run_impl(ctx, p, pos, res, energies, eps, rho);
}

int chs_build_system(chs_context* ctx, const chs_scf_params* p, const double* pos, int* n_pw_out,
                     int* n_proj_out, int32_t* d_miller, double* d_kin, double* d_vion,
                     double* d_beta, double* d_d0, double* d_q) {
  if (!ctx || !p || !pos || !n_pw_out || !n_proj_out || p->n_rep <= 0) return CHS_ERR_INVALID_ARG;
  const int np = p->n_proj_per_atom;
  if (np != 0 && np != 1 && np != 4)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "projectors per atom must be 0, 1 or 4");
  const int n = p->n_grid;
  const double L = p->n_rep * p->a0, omega = L * L * L;
  const int n_atoms = 8 * p->n_rep * p->n_rep * p->n_rep;
  int n_pw = 0;
  CHS_TRY(chs_gsphere(ctx, L, p->ecut, p->kfrac, n, 0, &n_pw, nullptr, nullptr, nullptr));
  const int P = np * n_atoms;
  *n_pw_out = n_pw;
  *n_proj_out = P;
  if (!d_miller && !d_kin && !d_vion && !d_beta && !d_d0 && !d_q) return CHS_OK;
  if (!d_miller || !d_kin || !d_vion || (P > 0 && (!d_beta || !d_d0 || !d_q)))
    return CHS_ERR_INVALID_ARG;
  DevBuf idx;
  int got = 0;
  CHS_TRY(ensure(ctx, idx, (size_t)n_pw * sizeof(int32_t)));
  int rc = chs_gsphere(ctx, L, p->ecut, p->kfrac, n, n_pw, &got, d_miller, (int32_t*)idx.p, d_kin);
  release(idx);
  CHS_TRY(rc);
  GridOps G;
  CHS_TRY(G.init(ctx, n, L, omega));
  CHS_TRY(G.set_atoms(pos, n_atoms));
  CHS_TRY(G.vion(p->zval, p->rc_loc, d_vion));
  if (P > 0) {
    DevBuf dpos;
    CHS_TRY(ensure(ctx, dpos, 3 * (size_t)n_atoms * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(dpos.p, pos, 3 * (size_t)n_atoms * sizeof(double), hipMemcpyHostToDevice));
    rc = build_beta(ctx, d_miller, n_pw, L, p->kfrac, (const double*)dpos.p, n_atoms, np,
                    p->beta_rc, (z64*)d_beta);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    release(dpos);
    CHS_TRY(rc);
    std::vector<double> d0 = seeded_d0(p, P), q(P, p->q0);
    CHS_HIP(ctx, hipMemcpy(d_d0, d0.data(), P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(d_q, q.data(), P * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += 3 * (size_t)n_atoms * sizeof(double) + 2 * (size_t)P * sizeof(double);
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

}  // extern "C"
