// The SCF loop around the eigensolver hot path (run_scf, reference
// src/scf.rs:358-376; diagonalize with its k loop, :169-265): chs_scf_run and
// chs_scf_run_k on the synthetic cells, and the caller-supplied system of
// chs_scf_context_create / _run / _destroy (PROJECT_ROOT_PLAN.md:880-1029).
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

// 0.5 |b (m + k)|^2 of caller Miller triples, each product and sum rounded
// separately as in the sphere enumeration (setup.hip), so a caller's sphere
// gets the kinetic energies chs_gsphere would give it.
__global__ void k_kin_miller(const int32_t* __restrict__ miller, int n_pw, double b, double k0,
                             double k1, double k2, double* __restrict__ kin) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  const double gx = b * ((double)miller[3 * g] + k0);
  const double gy = b * ((double)miller[3 * g + 1] + k1);
  const double gz = b * ((double)miller[3 * g + 2] + k2);
  kin[g] = 0.5 * ((gx * gx + gy * gy) + gz * gz);
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

// One k-point of an SCF system: its plane waves, projectors and block.
struct ScfKpt {
  int n_pw = 0;
  double kf[3] = {0, 0, 0};
  double w = 1.0;
  DevBuf miller, kin, beta, psi, psi32, wproj;
  std::vector<double> eps;
  void free_all() {
    for (DevBuf* b : {&miller, &kin, &beta, &psi, &psi32, &wproj}) release(*b);
  }
};

// Everything the SCF loop reads: a cubic cell of edge L on an n^3 grid,
// atoms, V_ion, projectors (np per atom, atom-major, shared D0 and Q) and
// the k-points with their start blocks. Device-resident.
struct ScfSystem {
  int n = 0, n_b = 0, n_atoms = 0, np = 0, P = 0;
  double L = 0, omega = 0, n_e = 0, q0 = 0, q_rc = 1.2;
  std::vector<double> pos;  // host, 3 n_atoms
  DevBuf vion, d0, q, rho0;
  bool has_rho0 = false;
  std::vector<ScfKpt> ks;
  ~ScfSystem() {
    for (ScfKpt& k : ks) k.free_all();
    for (DevBuf* b : {&vion, &d0, &q, &rho0}) release(*b);
  }
};

// The synthetic cell of p (diamond-cubic supercell, V_loc = -zval erf(r/rc)/r,
// seeded D0, Q = q0) at the given k-points, with seeded start blocks.
static int build_synthetic(chs_context* ctx, const chs_scf_params* p, const double* pos, int n_k,
                           const double* kfrac, const double* kw, ScfSystem& S) {
  const int np = p->n_proj_per_atom;
  if (np != 0 && np != 1 && np != 4)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "projectors per atom must be 0, 1 or 4");
  S.n = p->n_grid;
  S.n_b = p->n_bands;
  S.L = p->n_rep * p->a0;
  S.omega = S.L * S.L * S.L;
  S.n_atoms = 8 * p->n_rep * p->n_rep * p->n_rep;
  S.n_e = p->zval * S.n_atoms;
  S.pos.assign(pos, pos + 3 * (size_t)S.n_atoms);
  S.np = np;
  S.P = np * S.n_atoms;
  S.q0 = p->q0;
  S.q_rc = p->q_rc;
  const size_t n3 = (size_t)S.n * S.n * S.n;
  CHS_TRY(ensure(ctx, S.vion, n3 * sizeof(double)));
  {
    GridOps G;
    CHS_TRY(G.init(ctx, S.n, S.L, S.omega));
    CHS_TRY(G.set_atoms(pos, S.n_atoms));
    CHS_TRY(G.vion(p->zval, p->rc_loc, (double*)S.vion.p));
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  DevBuf dpos, idx;
  struct Free {
    DevBuf *a, *b;
    ~Free() {
      release(*a);
      release(*b);
    }
  } fr{&dpos, &idx};
  if (S.P > 0) {
    CHS_TRY(ensure(ctx, dpos, S.pos.size() * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(dpos.p, pos, S.pos.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += S.pos.size() * sizeof(double);
    std::vector<double> d0 = seeded_d0(p, S.P), q(S.P, p->q0);
    CHS_TRY(ensure(ctx, S.d0, S.P * sizeof(double)));
    CHS_TRY(ensure(ctx, S.q, S.P * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.d0.p, d0.data(), S.P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(S.q.p, q.data(), S.P * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += 2 * (size_t)S.P * sizeof(double);
  }
  S.ks.resize(n_k);
  for (int k = 0; k < n_k; ++k) {
    ScfKpt& K = S.ks[k];
    for (int c = 0; c < 3; ++c) K.kf[c] = kfrac[3 * k + c];
    K.w = kw[k];
    int n_pw = 0;
    CHS_TRY(chs_gsphere(ctx, S.L, p->ecut, K.kf, S.n, 0, &n_pw, nullptr, nullptr, nullptr));
    if (S.n_b > n_pw) return set_error(ctx, CHS_ERR_INVALID_ARG, "more bands than plane waves");
    K.n_pw = n_pw;
    CHS_TRY(ensure(ctx, K.miller, 3 * (size_t)n_pw * sizeof(int32_t)));
    CHS_TRY(ensure(ctx, idx, (size_t)n_pw * sizeof(int32_t)));
    CHS_TRY(ensure(ctx, K.kin, (size_t)n_pw * sizeof(double)));
    CHS_TRY(chs_gsphere(ctx, S.L, p->ecut, K.kf, S.n, n_pw, &n_pw, (int32_t*)K.miller.p,
                        (int32_t*)idx.p, (double*)K.kin.p));
    if (S.P > 0) {
      CHS_TRY(ensure(ctx, K.beta, (size_t)S.P * n_pw * sizeof(z64)));
      CHS_TRY(build_beta(ctx, (const int32_t*)K.miller.p, n_pw, S.L, K.kf,
                         (const double*)dpos.p, S.n_atoms, np, p->beta_rc, (z64*)K.beta.p));
    }
    // start wavefunctions: host generator, one upload
    std::vector<double> kin(n_pw);
    CHS_HIP(ctx, hipMemcpy(kin.data(), K.kin.p, n_pw * sizeof(double), hipMemcpyDeviceToHost));
    ctx->d2h += n_pw * sizeof(double);
    const size_t ne = (size_t)S.n_b * n_pw;
    std::unique_ptr<double[]> h_psi(new double[2 * ne]);
    initial_psi_host(kin, S.n_b, p->seed_psi, h_psi.get());
    CHS_TRY(ensure(ctx, K.psi, ne * sizeof(z64)));
    CHS_HIP(ctx, hipMemcpy(K.psi.p, h_psi.get(), ne * sizeof(z64), hipMemcpyHostToDevice));
    ctx->h2d += ne * sizeof(z64);
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

// A caller-supplied system (chs_scf_system), copied to the device.
static int build_from_caller(chs_context* ctx, const chs_scf_system* sy, const chs_scf_params* p,
                             ScfSystem& S) {
  if (!sy->positions || !sy->v_ion || sy->n_atoms <= 0 || sy->n_bands <= 0 || sy->n_k <= 0 ||
      !sy->k_frac || !sy->k_weight || !sy->n_pw || !sy->miller || !(sy->n_electrons > 0))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "incomplete SCF system");
  const int np = sy->proj_per_atom;
  if (np != 0 && np != 1 && np != 4)
    return set_error(ctx, CHS_ERR_INVALID_ARG, "projectors per atom must be 0, 1 or 4");
  if (np > 0 && (!sy->beta || !sy->d0))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "projectors need beta and d0");
  if (sy->grid[0] != sy->grid[1] || sy->grid[0] != sy->grid[2])
    return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "only cubic n^3 grids (got %d x %d x %d)",
                     sy->grid[0], sy->grid[1], sy->grid[2]);
  const double a = sy->real_lattice[0];
  for (int i = 0; i < 9; ++i) {
    const double want = (i % 4 == 0) ? a : 0.0;
    if (!(a > 0) || std::abs(sy->real_lattice[i] - want) > 1e-12 * a)
      return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "only cubic cells a I are supported");
  }
  S.n = sy->grid[0];
  S.n_b = sy->n_bands;
  S.L = a;
  S.omega = a * a * a;
  S.n_atoms = sy->n_atoms;
  S.n_e = sy->n_electrons;
  S.pos.assign(sy->positions, sy->positions + 3 * (size_t)S.n_atoms);
  S.np = np;
  S.P = np * S.n_atoms;
  S.q0 = np > 0 ? sy->q0 : 0.0;
  S.q_rc = sy->q_rc > 0 ? sy->q_rc : 1.0;
  const size_t n3 = (size_t)S.n * S.n * S.n;
  CHS_TRY(ensure(ctx, S.vion, n3 * sizeof(double)));
  CHS_HIP(ctx, hipMemcpy(S.vion.p, sy->v_ion, n3 * sizeof(double), hipMemcpyHostToDevice));
  ctx->h2d += n3 * sizeof(double);
  if (sy->rho0) {
    CHS_TRY(ensure(ctx, S.rho0, n3 * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.rho0.p, sy->rho0, n3 * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += n3 * sizeof(double);
    S.has_rho0 = true;
  }
  if (S.P > 0) {
    std::vector<double> q(S.P, S.q0);
    CHS_TRY(ensure(ctx, S.d0, S.P * sizeof(double)));
    CHS_TRY(ensure(ctx, S.q, S.P * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.d0.p, sy->d0, S.P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(S.q.p, q.data(), S.P * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += 2 * (size_t)S.P * sizeof(double);
  }
  S.ks.resize(sy->n_k);
  const double b = 2.0 * M_PI / a;
  for (int k = 0; k < sy->n_k; ++k) {
    ScfKpt& K = S.ks[k];
    for (int c = 0; c < 3; ++c) K.kf[c] = sy->k_frac[3 * k + c];
    K.w = sy->k_weight[k];
    const int n_pw = sy->n_pw[k];
    if (n_pw < S.n_b || !sy->miller[k] || (S.P > 0 && !sy->beta[k]))
      return set_error(ctx, CHS_ERR_INVALID_ARG, "k-point %d: plane-wave set or projectors", k);
    for (int g = 0; g < 3 * n_pw; ++g)
      if (2 * std::abs(sy->miller[k][g]) >= S.n)
        return set_error(ctx, CHS_ERR_INVALID_ARG, "k-point %d: Miller index outside the grid", k);
    K.n_pw = n_pw;
    CHS_TRY(ensure(ctx, K.miller, 3 * (size_t)n_pw * sizeof(int32_t)));
    CHS_TRY(ensure(ctx, K.kin, (size_t)n_pw * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(K.miller.p, sy->miller[k], 3 * (size_t)n_pw * sizeof(int32_t),
                           hipMemcpyHostToDevice));
    ctx->h2d += 3 * (size_t)n_pw * sizeof(int32_t);
    hipLaunchKernelGGL(k_kin_miller, dim3(ceil_div(n_pw, 256)), dim3(256), 0, ctx->stream,
                       (const int32_t*)K.miller.p, n_pw, b, K.kf[0], K.kf[1], K.kf[2],
                       (double*)K.kin.p);
    CHS_LAUNCH_CHECK(ctx);
    if (S.P > 0) {
      CHS_TRY(ensure(ctx, K.beta, (size_t)S.P * n_pw * sizeof(z64)));
      CHS_HIP(ctx, hipMemcpy(K.beta.p, sy->beta[k], (size_t)S.P * n_pw * sizeof(z64),
                             hipMemcpyHostToDevice));
      ctx->h2d += (size_t)S.P * n_pw * sizeof(z64);
    }
    const size_t ne = (size_t)S.n_b * n_pw;
    CHS_TRY(ensure(ctx, K.psi, ne * sizeof(z64)));
    if (sy->psi0 && sy->psi0[k]) {
      CHS_HIP(ctx, hipMemcpy(K.psi.p, sy->psi0[k], ne * sizeof(z64), hipMemcpyHostToDevice));
      ctx->h2d += ne * sizeof(z64);
    } else {
      // the kinetic energies come from k_kin_miller on the context stream
      // (non-blocking: a null-stream copy would not wait for it)
      CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
      std::vector<double> kin(n_pw);
      CHS_HIP(ctx, hipMemcpy(kin.data(), K.kin.p, n_pw * sizeof(double), hipMemcpyDeviceToHost));
      ctx->d2h += n_pw * sizeof(double);
      std::unique_ptr<double[]> h_psi(new double[2 * ne]);
      initial_psi_host(kin, S.n_b, p->seed_psi, h_psi.get());
      CHS_HIP(ctx, hipMemcpy(K.psi.p, h_psi.get(), ne * sizeof(z64), hipMemcpyHostToDevice));
      ctx->h2d += ne * sizeof(z64);
    }
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

// Outputs of one SCF run.
struct ScfOut {
  double* energies = nullptr;  // max_iter
  double* drho = nullptr;      // max_iter
  double* eps = nullptr;       // n_k n_b
  double* rho = nullptr;       // n^3 (host)
  bool keep_state = false;     // final density becomes the system's start density
};

// Per iteration: V_eff[rho_in] (+ screened D), the Lanczos bound of every
// k-point, Chebyshev filter + Rayleigh-Ritz of every k-point (first
// iteration: first_passes passes; all k-points' filters in one graph launch
// when n_k > 1), one Fermi level, rho_out = sum_k w_k rho_k (+ augmentation),
// E_tot, the convergence test and Pulay mixing.
static int scf_loop(chs_context* ctx, const chs_scf_params* p, ScfSystem& S, chs_scf_result* res,
                    const ScfOut& out) {
  const int n = S.n, n_b = S.n_b, K = (int)S.ks.size(), P = S.P;
  const double L = S.L, omega = S.omega, n_e = S.n_e;
  const int n_atoms = S.n_atoms;
  const size_t n3 = (size_t)n * n * n;
  const bool mixed = p->precision == CHS_C64;
  DevBuf veff, vhxc, rho_in, rho_out, rho_next, dscr, occd;
  struct Free {
    std::vector<DevBuf*> b;
    ~Free() {
      for (DevBuf* x : b) release(*x);
    }
  } fr{{&veff, &vhxc, &rho_in, &rho_out, &rho_next, &dscr, &occd}};
  // Every exit (errors included) unbinds the context Hamiltonian (and the
  // k-point set) before the arrays it points at are freed: declared after
  // them, so destroyed first.
  struct Unbind {
    chs_context* c;
    bool kset;
    ~Unbind() {
      unbind_hamiltonian(c);
      if (kset) kset_release(c);
    }
  } unbind{ctx, K > 1};
  for (DevBuf* b : {&veff, &vhxc, &rho_in, &rho_out, &rho_next})
    CHS_TRY(ensure(ctx, *b, n3 * sizeof(double)));
  GridOps G;
  CHS_TRY(G.init(ctx, n, L, omega));
  CHS_TRY(G.set_atoms(S.pos.data(), n_atoms));
  if (P > 0) {
    CHS_TRY(ensure(ctx, dscr, P * sizeof(double)));
    CHS_HIP(ctx, hipMemcpyAsync(dscr.p, S.d0.p, P * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream));
  }
  for (ScfKpt& k : S.ks) {
    const size_t ne = (size_t)n_b * k.n_pw;
    if (mixed) CHS_TRY(ensure(ctx, k.psi32, ne * sizeof(c32)));
    if (P > 0) CHS_TRY(ensure(ctx, k.wproj, (size_t)P * n_b * sizeof(z64)));
    k.eps.clear();
  }
  // bind the Hamiltonian: one k-point directly, several as a k-point set
  // (shared V_eff and screened D, both updated in place every iteration)
  if (K == 1) {
    ScfKpt& k0 = S.ks[0];
    CHS_TRY(chs_bind_hamiltonian(ctx, n, k0.n_pw, (const int32_t*)k0.miller.p,
                                 (const double*)k0.kin.p, (const double*)veff.p));
    if (P > 0)
      CHS_TRY(chs_bind_projectors(ctx, P, (const double*)k0.beta.p, (const double*)dscr.p,
                                  (const double*)S.q.p));
  } else {
    std::vector<int> npw(K);
    std::vector<const int32_t*> mil(K);
    std::vector<const double*> kin(K), beta(K);
    for (int k = 0; k < K; ++k) {
      npw[k] = S.ks[k].n_pw;
      mil[k] = (const int32_t*)S.ks[k].miller.p;
      kin[k] = (const double*)S.ks[k].kin.p;
      beta[k] = (const double*)S.ks[k].beta.p;
    }
    CHS_TRY(chs_kset_bind(ctx, K, n, npw.data(), mil.data(), kin.data(), (const double*)veff.p, P,
                          P > 0 ? beta.data() : nullptr, (const double*)dscr.p,
                          (const double*)S.q.p));
  }
  auto select = [&](int k) -> int {
    if (K > 1) return chs_kset_select(ctx, k);
    return CHS_OK;
  };
  // launch shapes of the H|psi> kernels, tuned once per (grid, precision)
  // and context (later runs reuse them)
  if (ctx->impl->hpsi_autotune) {
    CHS_TRY(select(0));
    for (int prec : {CHS_C128, CHS_C64}) {
      if (prec == CHS_C64 && !mixed) continue;
      if (ctx->impl->vloc_tune.count(std::make_pair(n, prec))) continue;
      CHS_TRY(chs_autotune_hpsi(ctx, prec, n_b, nullptr, nullptr, nullptr, nullptr));
    }
  }
  CHS_TRY(ensure(ctx, occd, (size_t)n_b * sizeof(double)));
  PulayDev mixer;
  CHS_TRY(mixer.init(ctx, &G, std::max(1, p->pulay_depth), p->mix_alpha, p->kerker_q));
  if (S.has_rho0)
    CHS_HIP(ctx, hipMemcpyAsync(rho_in.p, S.rho0.p, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream));
  else
    hipLaunchKernelGGL(k_fill_d, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream,
                       (double*)rho_in.p, n_e / omega, n3);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));

  std::vector<double> occ((size_t)K * n_b, 0.0), occw((size_t)K * n_b), e_all((size_t)K * n_b),
      w_state((size_t)K * n_b);
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < n_b; ++i) w_state[(size_t)k * n_b + i] = S.ks[k].w;
  std::vector<double> lmax(K), b4all(4 * (size_t)K);
  std::vector<void*> fptr(K);
  double e_prev = 0, mu = 0, etot = 0, drho = 0;
  int small = 0, it = 0, converged = 0;
  bool fp64_phase = !mixed, skip_de = false;
  int fp64_iters = 0, fp32_iters = 0;
  // complex64 filter until |dE|/atom < fp32_until (default 2 etol, but not
  // below 3e-7 Ha, near the energy noise the complex64 iterates leave), or
  // until |dE| stalls (> 0.9 of the previous one) within 10 fp32_until. The
  // energy change into the first FP64 iteration spans both precisions and is
  // not counted: convergence rests on two consecutive FP64-to-FP64 changes.
  const double fp32_until = p->fp32_until > 0 ? p->fp32_until : std::max(2.0 * p->etol, 3e-7);
  double de_prev = INFINITY;
  const double t0 = now();
  double t_filter = 0, t_rr = 0, t_den = 0, t_veff = 0, t_lan = 0;
  for (it = 1; it <= p->max_iter; ++it) {
    double ta = now();
    double eh_in, exc_in;
    CHS_TRY(G.vhxc((const double*)rho_in.p, (const double*)S.vion.p, (double*)vhxc.p,
                   (double*)veff.p, nullptr, &eh_in, &exc_in, nullptr));
    if (K == 1) CHS_TRY(chs_set_veff(ctx, (const double*)veff.p));
    else CHS_TRY(chs_kset_set_veff(ctx, (const double*)veff.p));
    if (P > 0 && S.q0 != 0.0)
      CHS_TRY(screen_dion(ctx, (const double*)veff.p, n, L, n_atoms, S.np, S.q0, S.q_rc,
                          (const z64*)G.ph.p, (const double*)S.d0.p, (z64*)G.cgrid.p,
                          (double*)dscr.p));
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double tb = now();
    t_veff += tb - ta;
    const bool use32 = mixed && !fp64_phase;
    for (int k = 0; k < K; ++k) {
      CHS_TRY(select(k));
      if (p->lanczos_k > 0) {
        // complex64 filter phase: the bound's projector products on the
        // complex64 beta the filter itself uses
        CHS_TRY(lanczos_impl(ctx, p->lanczos_k, p->seed_lanczos, &lmax[k], nullptr, use32));
      } else {
        CHS_TRY(chs_grid_lambda_max(ctx, L, &lmax[k]));
      }
    }
    ta = now();
    t_lan += ta - tb;
    const int passes = it == 1 ? std::max(1, p->first_passes) : 1;
    if (use32) ++fp32_iters;
    for (int ps = 0; ps < passes; ++ps) {
      for (int k = 0; k < K; ++k)
        chs_spectral_bounds(S.ks[k].eps.data(), (int)S.ks[k].eps.size(), lmax[k], &b4all[4 * k]);
      const double tf0 = now();
      for (int k = 0; k < K; ++k) {
        ScfKpt& kp = S.ks[k];
        const size_t ne = (size_t)n_b * kp.n_pw;
        if (use32) {
          hipLaunchKernelGGL(k_z2c, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                             (const z64*)kp.psi.p, (c32*)kp.psi32.p, ne);
          CHS_LAUNCH_CHECK(ctx);
        }
        fptr[k] = use32 ? kp.psi32.p : kp.psi.p;
      }
      const int fprec = use32 ? CHS_C64 : CHS_C128;
      // final iterations of the complex64 form with the residual-based
      // filter: complex64 products, FP64 residual and assembly
      // (k-point sets: one residual filter per k-point, outside the graph)
      const bool res = mixed && fp64_phase && p->fp64_refine == 1 &&
                       (int)S.ks[0].eps.size() == n_b;
      if (res) {
        for (int k = 0; k < K; ++k) {
          CHS_TRY(select(k));
          CHS_TRY(chebyshev_filter_res_impl(ctx, CHS_C64, (const z64*)S.ks[k].psi.p, n_b,
                                            S.ks[k].eps.data(), p->cheb_degree, &b4all[4 * k],
                                            nullptr, (z64*)S.ks[k].psi.p));
        }
      } else if (K == 1)
        CHS_TRY(chebyshev_filter_impl(ctx, fprec, fptr[0], n_b, p->cheb_degree, b4all.data(),
                                      fptr[0]));
      else if (p->cheb_degree > 0)
        CHS_TRY(chs_kset_filter(ctx, fprec, fptr.data(), n_b, p->cheb_degree, b4all.data()));
      if (use32)
        for (ScfKpt& kp : S.ks) {
          const size_t ne = (size_t)n_b * kp.n_pw;
          hipLaunchKernelGGL(k_c2z, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                             (const c32*)kp.psi32.p, (z64*)kp.psi.p, ne);
          CHS_LAUNCH_CHECK(ctx);
        }
      const double tf1 = now();
      t_filter += tf1 - tf0;
      for (int k = 0; k < K; ++k) {
        ScfKpt& kp = S.ks[k];
        CHS_TRY(select(k));
        kp.eps.resize(n_b);
        // the last pass also hands over beta^H psi of the rotated block for
        // the augmentation charge; in the complex64 phase H|Y> and beta^H Y
        // come from the filter's complex64 output (Gram matrices and the
        // eigensolve stay FP64)
        CHS_TRY(rayleigh_ritz_impl(ctx, (z64*)kp.psi.p, n_b, kp.eps.data(), nullptr, nullptr,
                                   P > 0 && ps + 1 == passes ? (z64*)kp.wproj.p : nullptr,
                                   use32 ? (const c32*)kp.psi32.p : nullptr));
      }
      t_rr += now() - tf1;
    }
    tb = now();
    double ts = 0;
    for (int k = 0; k < K; ++k)
      std::copy(S.ks[k].eps.begin(), S.ks[k].eps.end(), e_all.begin() + (size_t)k * n_b);
    if (K == 1)
      mu = fermi_occupations(e_all.data(), n_b, n_e, p->smear_kt, occ.data(), &ts);
    else
      mu = fermi_occupations_w(e_all.data(), w_state.data(), K * n_b, n_e, p->smear_kt,
                               occ.data(), &ts);
    // weighted occupations w_k f_{k,i}, one host slice per k-point (each
    // stays untouched while its asynchronous uploads may still read it)
    double eband = 0;
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < n_b; ++i) {
        const size_t j = (size_t)k * n_b + i;
        occw[j] = K == 1 ? occ[j] : S.ks[k].w * occ[j];
        eband += occw[j] * S.ks[k].eps[i];
      }
    for (int k = 0; k < K; ++k) {
      ScfKpt& kp = S.ks[k];
      CHS_TRY(select(k));
      const double* ok = occw.data() + (size_t)k * n_b;
      const size_t ne = (size_t)n_b * kp.n_pw;
      if (use32 && ctx->impl->ham.st.rs_win > 0) {
        // complex64 phase: the density from the rotated block rounded to
        // complex64 (FP64 accumulation), like the filter that produced it
        hipLaunchKernelGGL(k_z2c, dim3(ceil_div(ne, 256)), dim3(256), 0, ctx->stream,
                           (const z64*)kp.psi.p, (c32*)kp.psi32.p, ne);
        CHS_LAUNCH_CHECK(ctx);
        CHS_TRY(density_c64(ctx, (const c32*)kp.psi32.p, n_b, ok, omega,
                            (double*)rho_out.p, k > 0));
      } else {
        CHS_TRY(density_impl(ctx, (const z64*)kp.psi.p, n_b, ok, omega,
                             (double*)rho_out.p, k > 0));
      }
      if (P > 0 && S.q0 != 0.0) {
        CHS_HIP(ctx, hipMemcpyAsync(occd.p, ok, n_b * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
        ctx->h2d += n_b * sizeof(double);
        CHS_TRY(augment_density(ctx, (const z64*)kp.psi.p, n_b, (const double*)occd.p, n_atoms,
                                S.np, S.q0, S.q_rc, (const z64*)G.ph.p, n, L, (z64*)G.cgrid.p,
                                (double*)rho_out.p, (const z64*)kp.wproj.p));
      }
    }
    double eh, exc, dc;
    CHS_TRY(G.vhxc((const double*)rho_out.p, (const double*)S.vion.p, nullptr, nullptr,
                   (const double*)vhxc.p, &eh, &exc, &dc));
    int rc = CHS_OK;
    drho = l1_diff(ctx, (const double*)rho_out.p, (const double*)rho_in.p, n3, &rc) * omega /
           (double)n3;
    CHS_TRY(rc);
    etot = eband - dc + eh + exc + ts;
    if (out.energies) out.energies[it - 1] = etot;
    if (out.drho) out.drho[it - 1] = drho;
    if (fp64_phase) ++fp64_iters;
    if (it > 1) {
      const double de = std::abs(etot - e_prev) / n_atoms;
      if (skip_de) {
        small = 0;
        skip_de = false;
      } else {
        small = de < p->etol ? small + 1 : 0;
      }
      if (!fp64_phase && (de < fp32_until || (de < 10.0 * fp32_until && de > 0.9 * de_prev))) {
        fp64_phase = true;  // the FP32 filter has done its part: finish in FP64
        small = 0;
        skip_de = true;
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
    CHS_TRY(mixer.mix((const double*)rho_in.p, (const double*)rho_out.p, (double*)rho_next.p));
    std::swap(rho_in, rho_next);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    t_den += now() - ta;
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double t1 = now();
  res->n_iter = std::min(it, p->max_iter);
  res->converged = converged;
  res->n_pw = S.ks[0].n_pw;
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
  res->drho = drho;
  if (out.eps)
    for (int k = 0; k < K; ++k)
      std::copy(S.ks[k].eps.begin(), S.ks[k].eps.end(), out.eps + (size_t)k * n_b);
  if (out.rho) {
    CHS_HIP(ctx, hipMemcpy(out.rho, rho_out.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
    ctx->d2h += n3 * sizeof(double);
  }
  if (out.keep_state) {
    CHS_TRY(ensure(ctx, S.rho0, n3 * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(S.rho0.p, rho_out.p, n3 * sizeof(double), hipMemcpyDeviceToDevice));
    S.has_rho0 = true;
  }
  return converged ? CHS_OK : CHS_ERR_NOT_CONVERGED;
}

static int check_solver(chs_context* ctx, const chs_scf_params* p) {
  if (p->max_iter <= 0 || p->cheb_degree < 0 ||
      (p->precision != CHS_C128 && p->precision != CHS_C64) ||
      (p->fp64_refine != 0 && p->fp64_refine != 1))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "bad SCF parameters");
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

struct chs_scf {
  chs_context* ctx = nullptr;  // null once the context is destroyed
  chs_scf_params p{};
  chs::ScfSystem sys;
};

void chs::scf_orphan(chs_scf* s) {
  for (ScfKpt& k : s->sys.ks) k.free_all();
  for (DevBuf* b : {&s->sys.vion, &s->sys.d0, &s->sys.q, &s->sys.rho0}) release(*b);
  s->sys.ks.clear();
  s->ctx = nullptr;
}

extern "C" {

int chs_density(chs_context* ctx, const double* d_psi, int n_b, const double* occ, double omega,
                double* d_rho) {
  if (!ctx || !d_psi || n_b <= 0 || !occ || !(omega > 0) || !d_rho) return CHS_ERR_INVALID_ARG;
  return density_impl(ctx, (const z64*)d_psi, n_b, occ, omega, d_rho);
}

int chs_initial_psi(chs_context* ctx, const double* d_kin, int n_pw, int n_b, uint64_t seed,
                    double* d_psi) {
  if (!ctx || !d_kin || !d_psi || n_pw <= 0 || n_b <= 0) return CHS_ERR_INVALID_ARG;
  std::vector<double> kin(n_pw);
  CHS_HIP(ctx, hipMemcpyAsync(kin.data(), d_kin, n_pw * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += n_pw * sizeof(double);
  const size_t ne = (size_t)n_b * n_pw;
  std::unique_ptr<double[]> h(new double[2 * ne]);
  initial_psi_host(kin, n_b, seed, h.get());
  CHS_HIP(ctx, hipMemcpy(d_psi, h.get(), ne * sizeof(z64), hipMemcpyHostToDevice));
  ctx->h2d += ne * sizeof(z64);
  return CHS_OK;
}

int chs_scf_run_k(chs_context* ctx, const chs_scf_params* p, const double* pos, int n_k,
                  const double* kfrac, const double* kweight, chs_scf_result* res,
                  double* energies, double* drho, double* eps, double* rho) {
  if (!ctx || !p || !pos || !res || n_k <= 0 || !kfrac || !kweight) return CHS_ERR_INVALID_ARG;
  if (p->n_rep <= 0 || p->n_bands <= 0) return set_error(ctx, CHS_ERR_INVALID_ARG, "bad SCF parameters");
  CHS_TRY(check_solver(ctx, p));
  *res = chs_scf_result{};
  ScfSystem S;
  CHS_TRY(build_synthetic(ctx, p, pos, n_k, kfrac, kweight, S));
  ScfOut out;
  out.energies = energies;
  out.drho = drho;
  out.eps = eps;
  out.rho = rho;
  return scf_loop(ctx, p, S, res, out);
}

int chs_scf_run(chs_context* ctx, const chs_scf_params* p, const double* pos, chs_scf_result* res,
                double* energies, double* eps, double* rho) {
  if (!p) return CHS_ERR_INVALID_ARG;
  const double w = 1.0;
  return chs_scf_run_k(ctx, p, pos, 1, p->kfrac, &w, res, energies, nullptr, eps, rho);
}

int chs_scf_context_create(chs_context* ctx, const chs_scf_system* sys,
                           const chs_scf_params* solver, chs_scf** out) {
  if (!ctx || !sys || !solver || !out) return CHS_ERR_INVALID_ARG;
  *out = nullptr;
  CHS_TRY(check_solver(ctx, solver));
  std::unique_ptr<chs_scf> s(new chs_scf);
  s->ctx = ctx;
  s->p = *solver;
  CHS_TRY(build_from_caller(ctx, sys, solver, s->sys));
  ctx->impl->scf_children.push_back(s.get());
  *out = s.release();
  return CHS_OK;
}

int chs_scf_context_run(chs_scf* s, chs_scf_result* res, double* energies, double* drho,
                        double* eps, double* rho) {
  if (!s || !res || !s->ctx) return CHS_ERR_INVALID_ARG;  // no context: destroyed
  *res = chs_scf_result{};
  ScfOut out;
  out.energies = energies;
  out.drho = drho;
  out.eps = eps;
  out.rho = rho;
  out.keep_state = true;
  return scf_loop(s->ctx, &s->p, s->sys, res, out);
}

int chs_scf_get_psi(chs_scf* s, int k, double* psi) {
  if (!s || !psi || !s->ctx || k < 0 || k >= (int)s->sys.ks.size()) return CHS_ERR_INVALID_ARG;
  const ScfKpt& K = s->sys.ks[k];
  const size_t bytes = (size_t)s->sys.n_b * K.n_pw * sizeof(z64);
  CHS_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  CHS_HIP(s->ctx, hipMemcpy(psi, K.psi.p, bytes, hipMemcpyDeviceToHost));
  s->ctx->d2h += bytes;
  return CHS_OK;
}

int chs_scf_context_destroy(chs_scf* s) {
  if (!s) return CHS_ERR_INVALID_ARG;
  if (s->ctx) {
    auto& ch = s->ctx->impl->scf_children;
    ch.erase(std::remove(ch.begin(), ch.end(), s), ch.end());
    (void)hipStreamSynchronize(s->ctx->stream);
  }
  delete s;
  return CHS_OK;
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
    ctx->h2d += 3 * (size_t)n_atoms * sizeof(double);
    rc = build_beta(ctx, d_miller, n_pw, L, p->kfrac, (const double*)dpos.p, n_atoms, np,
                    p->beta_rc, (z64*)d_beta);
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    release(dpos);
    CHS_TRY(rc);
    std::vector<double> d0 = seeded_d0(p, P), q(P, p->q0);
    CHS_HIP(ctx, hipMemcpy(d_d0, d0.data(), P * sizeof(double), hipMemcpyHostToDevice));
    CHS_HIP(ctx, hipMemcpy(d_q, q.data(), P * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += 2 * (size_t)P * sizeof(double);
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

}  // extern "C"
