// SCF glue for the FP64 oracle: density, V_eff (local ion + Hartree + LDA),
// screened D, total energy, Pulay/Kerker mixing and the SCF driver.
// Test infrastructure only (see orc.hpp).
//
// The reference leaves all of these as todo!() (src/scf.rs:157-159,
// :275-279, :328-334, :343-345; src/mixing.rs:24-27). Their specifications
// are PROJECT_ROOT_PLAN.md:568-611 (density), notes/plans/phase-2/
// TASKS.md:332-358 (integral of rho = N_e) and :386-440 (Pulay depth,
// convergence check); the diagonalize transition they wrap is
// src/scf.rs:169-265.
#include "scf.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace orc {

// Per-atom factorised structure phases e^{-i G.R} over the full grid.
static void atom_phases(const System& sys, std::vector<cd>& ph) {
  const int n = sys.n_grid;
  const double b = (2.0 * M_PI) / sys.L;
  ph.resize((size_t)sys.n_atoms * 3 * n);
  for (int a = 0; a < sys.n_atoms; ++a)
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < n; ++i) {
        bool nyq;
        int f = freq_of(i, n, &nyq);
        double t = -b * f * sys.pos[3 * a + c];
        ph[((size_t)a * 3 + c) * n + i] = cd(std::cos(t), std::sin(t));
      }
}

// |G|^2 on the full grid; Nyquist points flagged by a negative value.
static double grid_g2(const System& sys, int ix, int iy, int iz) {
  const int n = sys.n_grid;
  const double b = (2.0 * M_PI) / sys.L;
  bool nx, ny, nz;
  int h = freq_of(ix, n, &nx), k = freq_of(iy, n, &ny), l = freq_of(iz, n, &nz);
  if (nx || ny || nz) return -1.0;
  return b * b * (double)(h * h + k * k + l * l);
}

// rho(r) = sum_i f_i |psi_i(r)|^2 / Omega, psi_i(r) the unnormalised
// inverse FFT of the scattered coefficients (bands with f_i = 0 skipped);
// plus the augmentation
// rho_aug(G) = Omega^-1 sum_{I,p} rho_{Ip} q0 exp(-G^2 rq^2/4) e^{-iG.R_I}
// with rho_{Ip} = sum_i f_i |<beta_{Ip}|psi_i>|^2.
void density(const Model& m, const cd* psi, int n_b, const double* occ, std::vector<double>& rho) {
  const int n = m.sys.n_grid, n_pw = m.pw.n_pw;
  const size_t n3 = (size_t)n * n * n;
  rho.assign(n3, 0.0);
  const double inv_omega = 1.0 / m.sys.omega;
#pragma omp parallel
  {
    std::vector<cd> grid(n3);
    std::vector<double> acc(n3, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < n_b; ++b) {
      if (occ[b] == 0.0) continue;
      std::fill(grid.begin(), grid.end(), cd(0));
      const cd* p = psi + (size_t)b * n_pw;
      for (int g = 0; g < n_pw; ++g) grid[m.pw.fft_idx[g]] = p[g];
      fft3d(grid.data(), n, +1);
      for (size_t r = 0; r < n3; ++r) acc[r] += occ[b] * std::norm(grid[r]);
    }
#pragma omp critical
    for (size_t r = 0; r < n3; ++r) rho[r] += acc[r] * inv_omega;
  }
  const int P = m.pj.n_proj;
  if (P == 0 || m.sys.q0 == 0.0) return;
  const int np = m.sys.n_proj_per_atom;
  std::vector<double> rp(P, 0.0);
  for (int p = 0; p < P; ++p) {
    const cd* bp = m.pj.beta.data() + (size_t)p * n_pw;
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int b = 0; b < n_b; ++b) {
      const cd* x = psi + (size_t)b * n_pw;
      cd c = 0;
      for (int g = 0; g < n_pw; ++g) c += std::conj(bp[g]) * x[g];
      s += occ[b] * std::norm(c);
    }
    rp[p] = s;
  }
  std::vector<double> ra(m.sys.n_atoms, 0.0);
  for (int a = 0; a < m.sys.n_atoms; ++a)
    for (int p = 0; p < np; ++p) ra[a] += rp[a * np + p] * m.sys.q0;
  std::vector<cd> ph;
  atom_phases(m.sys, ph);
  std::vector<cd> grid(n3);
  const double rq = m.sys.q_rc;
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        size_t id = (size_t)ix + (size_t)n * (iy + (size_t)n * iz);
        double g2 = grid_g2(m.sys, ix, iy, iz);
        if (g2 < 0) { grid[id] = 0; continue; }
        cd s = 0;
        for (int a = 0; a < m.sys.n_atoms; ++a)
          s += ra[a] * ph[((size_t)a * 3 + 0) * n + ix] * ph[((size_t)a * 3 + 1) * n + iy] *
               ph[((size_t)a * 3 + 2) * n + iz];
        grid[id] = s * std::exp(-g2 * rq * rq / 4.0) / m.sys.omega;
      }
  fft3d(grid.data(), n, +1);
  for (size_t r = 0; r < n3; ++r) rho[r] += grid[r].real();
}

// D_{Ip} = D0_{Ip} + q0 sum_G V(G) e^{iG.R_I} exp(-G^2 rq^2/4), V(G) the
// forward FFT of V_eff over N^3 (the integral of V_eff against the
// normalised Gaussian augmentation function).
void screen_d(const Model& m, std::vector<double>& d) {
  const int P = m.pj.n_proj, np = m.sys.n_proj_per_atom, n = m.sys.n_grid;
  d.assign(P, 0.0);
  if (P == 0) return;
  for (int p = 0; p < P; ++p) d[p] = m.sys.dion[p];
  if (m.sys.q0 == 0.0) return;
  const size_t n3 = (size_t)n * n * n;
  std::vector<cd> vg(m.veff.begin(), m.veff.end());
  fft3d(vg.data(), n, -1);
  std::vector<cd> ph;
  atom_phases(m.sys, ph);
  const double rq = m.sys.q_rc;
  for (int a = 0; a < m.sys.n_atoms; ++a) {
    double s = 0;
#pragma omp parallel for collapse(2) reduction(+ : s) schedule(static)
    for (int iz = 0; iz < n; ++iz)
      for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
          double g2 = grid_g2(m.sys, ix, iy, iz);
          if (g2 < 0) continue;
          size_t id = (size_t)ix + (size_t)n * (iy + (size_t)n * iz);
          cd e = std::conj(ph[((size_t)a * 3 + 0) * n + ix] * ph[((size_t)a * 3 + 1) * n + iy] *
                           ph[((size_t)a * 3 + 2) * n + iz]);
          s += (vg[id] * e).real() * std::exp(-g2 * rq * rq / 4.0);
        }
    s /= (double)n3;
    for (int p = 0; p < np; ++p) d[a * np + p] += m.sys.q0 * s;
  }
}

// Fermi-Dirac occupations f_i = 2 / (1 + exp((eps_i - mu)/kT)) with mu by
// bisection so that sum f_i = n_e (the chemical-potential search of
// notes/plans/phase-2/TASKS.md:342-343). Returns mu; *ts gets the entropy
// term -T S = 2 kT sum [f ln f + (1-f) ln(1-f)] (f = f_i/2). kT = 0 gives
// integer filling of the lowest n_e/2 bands.
double fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ, double* ts) {
  *ts = 0;
  if (kt <= 0) {
    int n_occ = (int)std::lround(n_e / 2.0);
    for (int i = 0; i < n; ++i) occ[i] = i < n_occ ? 2.0 : 0.0;
    return n_occ > 0 && n_occ <= n ? eps[n_occ - 1] : 0.0;
  }
  auto fd = [&](double e, double mu) {
    double x = (e - mu) / kt;
    if (x > 200) return 0.0;
    if (x < -200) return 2.0;
    return 2.0 / (1.0 + std::exp(x));
  };
  double lo = eps[0] - 50 * kt, hi = eps[n - 1] + 50 * kt;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi), s = 0;
    for (int i = 0; i < n; ++i) s += fd(eps[i], mid);
    if (s > n_e) hi = mid; else lo = mid;
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

// Fermi-Dirac occupations of states with weights w_i (the k-point weight of
// each state, summing to one per band index over k): sum_i w_i f_i = n_e,
// -T S = 2 kT sum_i w_i [f ln f + (1-f) ln(1-f)]. With every w_i = 1 this is
// fermi_occupations above. kT = 0: states filled in ascending order, two
// electrons of weight w_i each, until n_e is reached.
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
    double x = (e - mu) / kt;
    if (x > 200) return 0.0;
    if (x < -200) return 2.0;
    return 2.0 / (1.0 + std::exp(x));
  };
  double lo = emin - 50 * kt, hi = emax + 50 * kt;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi), s = 0;
    for (int i = 0; i < n; ++i) s += w[i] * fd(eps[i], mid);
    if (s > n_e) hi = mid; else lo = mid;
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

// LDA: Slater exchange + Perdew-Zunger 1981 correlation (unpolarised).
void lda_pz(double rho, double* exc, double* vxc) {
  if (rho <= 1e-14) {
    *exc = 0;
    *vxc = 0;
    return;
  }
  const double cx = 0.75 * std::cbrt(3.0 / M_PI);
  const double r13 = std::cbrt(rho);
  const double ex = -cx * r13;
  const double vx = 4.0 / 3.0 * ex;
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * rho));
  double ec, vc;
  if (rs >= 1.0) {
    const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sq = std::sqrt(rs);
    const double den = 1.0 + b1 * sq + b2 * rs;
    ec = g / den;
    vc = ec * (1.0 + 7.0 / 6.0 * b1 * sq + 4.0 / 3.0 * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = std::log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  *exc = ex + ec;
  *vxc = vx + vc;
}

// Hartree potential and energy: rho(G) = FFT(rho)/N^3, V_H(G) = 4 pi rho(G)/G^2
// (G != 0, Nyquist excluded), E_H = Omega/2 sum 4 pi |rho(G)|^2 / G^2.
double hartree(const System& sys, const std::vector<double>& rho, std::vector<double>& vh) {
  const int n = sys.n_grid;
  const size_t n3 = (size_t)n * n * n;
  std::vector<cd> g(rho.begin(), rho.end());
  fft3d(g.data(), n, -1);
  double eh = 0;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        size_t id = (size_t)ix + (size_t)n * (iy + (size_t)n * iz);
        double g2 = grid_g2(sys, ix, iy, iz);
        if (g2 <= 0) { g[id] = 0; continue; }
        cd rg = g[id] / (double)n3;
        eh += 0.5 * sys.omega * 4.0 * M_PI * std::norm(rg) / g2;
        g[id] = 4.0 * M_PI * rg / g2;
      }
  fft3d(g.data(), n, +1);
  vh.resize(n3);
  for (size_t r = 0; r < n3; ++r) vh[r] = g[r].real();
  return eh;
}

// V_Hxc[rho] on the grid and the energies E_H, E_xc.
void vhxc(const System& sys, const std::vector<double>& rho, std::vector<double>& v, double* eh,
          double* exc) {
  const size_t n3 = rho.size();
  std::vector<double> vh;
  *eh = hartree(sys, rho, vh);
  v.resize(n3);
  double ex = 0;
  for (size_t r = 0; r < n3; ++r) {
    double e, vx;
    lda_pz(rho[r], &e, &vx);
    ex += rho[r] * e;
    v[r] = vh[r] + vx;
  }
  *exc = ex * sys.omega / (double)n3;
}

// Kerker filter K(G) = G^2 / (G^2 + q^2) applied to a real grid function.
static void kerker(const System& sys, double q, std::vector<double>& f) {
  const int n = sys.n_grid;
  const size_t n3 = (size_t)n * n * n;
  std::vector<cd> g(f.begin(), f.end());
  fft3d(g.data(), n, -1);
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        size_t id = (size_t)ix + (size_t)n * (iy + (size_t)n * iz);
        double g2 = grid_g2(sys, ix, iy, iz);
        if (g2 <= 0) { g[id] = 0; continue; }
        g[id] *= g2 / (g2 + q * q) / (double)n3;
      }
  fft3d(g.data(), n, +1);
  for (size_t r = 0; r < n3; ++r) f[r] = g[r].real();
}

// Pulay (DIIS) mixing of the density with Kerker-filtered residuals:
// minimise |sum c_i R_i|^2 subject to sum c_i = 1 over the last `depth`
// pairs, then rho_next = sum c_i rho_in_i + alpha K[sum c_i R_i].
// The spec of the reference's unimplemented DensityHistory::mix is
// plans/phase-2/PHASE_PLAN.md:221-230.
void Pulay::mix(const System& sys, const std::vector<double>& rin, const std::vector<double>& rout,
                std::vector<double>& rnext) {
  const size_t n3 = rin.size();
  std::vector<double> res(n3);
  for (size_t r = 0; r < n3; ++r) res[r] = rout[r] - rin[r];
  hin.push_back(rin);
  hres.push_back(res);
  if ((int)hin.size() > depth) {
    hin.erase(hin.begin());
    hres.erase(hres.begin());
  }
  const int m = (int)hin.size();
  std::vector<double> c(m, 0.0);
  for (;;) {
    const int k = (int)hin.size();
    const int off = m - k;
    std::vector<double> a((size_t)(k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j <= i; ++j) {
        double s = 0;
        for (size_t r = 0; r < n3; ++r) s += hres[i][r] * hres[j][r];
        a[i * (k + 1) + j] = s;
        a[j * (k + 1) + i] = s;
      }
    for (int i = 0; i < k; ++i) {
      a[i * (k + 1) + k] = 1.0;
      a[k * (k + 1) + i] = 1.0;
    }
    rhs[k] = 1.0;
    if (solve_dense(a, rhs, k + 1)) {
      std::fill(c.begin(), c.end(), 0.0);
      for (int i = 0; i < k; ++i) c[off + i] = rhs[i];
      break;
    }
    hin.erase(hin.begin());
    hres.erase(hres.begin());
    if (hin.empty()) {
      c.assign(m, 0.0);
      c[m - 1] = 1.0;
      hin.push_back(rin);
      hres.push_back(res);
      break;
    }
  }
  const int k = (int)hin.size();
  const int off = m - k;
  std::vector<double> ropt(n3, 0.0), rres(n3, 0.0);
  for (int i = 0; i < k; ++i)
    for (size_t r = 0; r < n3; ++r) {
      ropt[r] += c[off + i] * hin[i][r];
      rres[r] += c[off + i] * hres[i][r];
    }
  kerker(sys, kerker_q, rres);
  rnext.resize(n3);
  for (size_t r = 0; r < n3; ++r) rnext[r] = ropt[r] + alpha * rres[r];
}

// Gaussian elimination with partial pivoting; false when singular or
// ill-conditioned (relative pivot < 1e-14).
bool solve_dense(std::vector<double>& a, std::vector<double>& b, int n) {
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
      double f = a[r * n + col] / a[col * n + col];
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

// Seeded start: psi_{b,g} = (x + i y) / (1 + |k+G|^2), x, y ~ N(0,1) drawn
// band-major then plane-wave order from one LCG stream.
void initial_psi(const PwSet& pw, int n_b, uint64_t seed, std::vector<cd>& psi) {
  psi.resize((size_t)n_b * pw.n_pw);
  Lcg rng(seed);
  for (int b = 0; b < n_b; ++b)
    for (int g = 0; g < pw.n_pw; ++g) {
      double x = rng.normal();
      double y = rng.normal();
      psi[(size_t)b * pw.n_pw + g] = cd(x, y) / (1.0 + 2.0 * pw.kin[g]);
    }
}

void setup_model(const ScfParams& p, const double* pos, Model& m) {
  m.sys.n_rep = p.n_rep;
  m.sys.a0 = p.a0;
  m.sys.L = p.n_rep * p.a0;
  m.sys.omega = m.sys.L * m.sys.L * m.sys.L;
  m.sys.n_atoms = 8 * p.n_rep * p.n_rep * p.n_rep;
  m.sys.pos.assign(pos, pos + 3 * m.sys.n_atoms);
  m.sys.zval = p.zval;
  m.sys.rc_loc = p.rc_loc;
  m.sys.n_grid = p.n_grid;
  m.sys.n_proj_per_atom = p.n_proj_per_atom;
  m.sys.beta_rc = p.beta_rc;
  m.sys.q_rc = p.q_rc;
  m.sys.q0 = p.q0;
  m.sys.dion.assign((size_t)m.sys.n_atoms * p.n_proj_per_atom, 0.0);
  Lcg rd(p.seed_d);
  for (auto& d : m.sys.dion) d = p.d_min + (p.d_max - p.d_min) * rd.uniform();
  gsphere(m.sys.L, p.ecut, p.kfrac, p.n_grid, m.pw);
  build_projectors(m.sys, m.pw, m.pj);
  if (m.pj.n_proj) build_sinv(m.pj, m.pw.n_pw);
  m.sinv_filter = true;
}

// SCF driver: per iteration V_eff[rho_in] (+ screened D), Lanczos bound,
// Chebyshev filter + Rayleigh-Ritz (first iteration: first_passes passes),
// rho_out, E_tot, convergence |dE|/N_at < etol on two consecutive
// iterations, Pulay mixing.
int run_scf(const ScfParams& p, const double* pos, ScfResult& res, Model* model_out) {
  Model m;
  setup_model(p, pos, m);
  const int n_b = p.n_bands, n_pw = m.pw.n_pw;
  const size_t n3 = (size_t)p.n_grid * p.n_grid * p.n_grid;
  const double n_e = p.zval * m.sys.n_atoms;
  std::vector<double> occ(n_b, 0.0);
  std::vector<double> vion;
  vion_real(m.sys, vion);
  std::vector<cd> psi;
  initial_psi(m.pw, n_b, p.seed_psi, psi);
  std::vector<double> rho_in(n3, n_e / m.sys.omega), rho_out, vh, v;
  std::vector<double> eps;
  Pulay mixer;
  mixer.depth = p.pulay_depth;
  mixer.alpha = p.mix_alpha;
  mixer.kerker_q = p.kerker_q;
  res.energies.clear();
  res.n_iter = 0;
  res.converged = 0;
  res.n_pw = n_pw;
  int small = 0;
  double e_prev = 0;
  for (int it = 1; it <= p.max_iter; ++it) {
    double eh_in, exc_in;
    vhxc(m.sys, rho_in, v, &eh_in, &exc_in);
    m.veff.resize(n3);
    for (size_t r = 0; r < n3; ++r) m.veff[r] = vion[r] + v[r];
    if (m.pj.n_proj) screen_d(m, m.pj.d);
    double lmax = p.lanczos_k > 0 ? lanczos_bound(m, p.lanczos_k, p.seed_lanczos, nullptr)
                                  : grid_lambda_max(m);
    const int passes = it == 1 ? p.first_passes : 1;
    for (int ps = 0; ps < passes; ++ps) {
      Bounds bd = spectral_bounds(eps.data(), (int)eps.size(), lmax);
      std::vector<cd> y;
      double diag[4];
      int rc = chebyshev_filter(m, psi.data(), n_b, p.cheb_degree, bd, y, diag);
      if (rc) return rc;
      rc = rayleigh_ritz(m, y, n_b, eps, nullptr);
      if (rc) return 1000 + rc;
      psi.swap(y);
    }
    double ts;
    res.mu = fermi_occupations(eps.data(), n_b, n_e, p.smear_kt, occ.data(), &ts);
    density(m, psi.data(), n_b, occ.data(), rho_out);
    double eh, exc;
    std::vector<double> vout;
    vhxc(m.sys, rho_out, vout, &eh, &exc);
    double eband = 0;
    for (int i = 0; i < n_b; ++i) eband += occ[i] * eps[i];
    double dc = 0, drho = 0;
    for (size_t r = 0; r < n3; ++r) {
      dc += v[r] * rho_out[r];
      drho += std::abs(rho_out[r] - rho_in[r]);
    }
    const double dv = m.sys.omega / (double)n3;
    const double etot = eband - dc * dv + eh + exc + ts;
    res.energies.push_back(etot);
    res.drho.push_back(drho * dv);
    res.n_iter = it;
    if (it > 1) {
      double de = std::abs(etot - e_prev) / m.sys.n_atoms;
      small = de < p.etol ? small + 1 : 0;
    }
    e_prev = etot;
    if (small >= 2) {
      res.converged = 1;
      break;
    }
    std::vector<double> rnext;
    mixer.mix(m.sys, rho_in, rho_out, rnext);
    rho_in.swap(rnext);
  }
  res.eps = eps;
  res.occ = occ;
  res.rho = rho_out;
  res.psi = psi;
  if (model_out) *model_out = m;
  return 0;
}

// k-point SCF: the k loop of ScfIteration::diagonalize (src/scf.rs:169-265,
// one filter + Rayleigh-Ritz per k-point) around the same glue as run_scf.
// Every k-point has its own plane-wave sphere 0.5|k+G|^2 <= ecut, start
// block (seed_psi) and Lanczos bound (seed_lanczos); V_eff and the screened
// D are shared. One Fermi level over all (k, band) states with the k
// weights wk (sum 1); rho = sum_k wk rho_k; E = sum_k wk sum_i f eps -
// int V_Hxc[rho_in] rho_out + E_H + E_xc - TS. The batched intent is
// PROJECT_ROOT_PLAN.md:568-611.
int run_scf_k(const ScfParams& p, const double* pos, int n_k, const double* kfrac,
              const double* wk, ScfResultK& res) {
  if (n_k < 1) return 1;
  std::vector<Model> ms(n_k);
  for (int k = 0; k < n_k; ++k) {
    ScfParams pk = p;
    for (int c = 0; c < 3; ++c) pk.kfrac[c] = kfrac[3 * k + c];
    setup_model(pk, pos, ms[k]);
  }
  const Model& m0 = ms[0];
  const int n_b = p.n_bands;
  const size_t n3 = (size_t)p.n_grid * p.n_grid * p.n_grid;
  const double n_e = p.zval * m0.sys.n_atoms;
  std::vector<double> vion;
  vion_real(m0.sys, vion);
  std::vector<std::vector<cd>> psi(n_k);
  std::vector<std::vector<double>> eps(n_k);
  res.n_pw.assign(n_k, 0);
  for (int k = 0; k < n_k; ++k) {
    initial_psi(ms[k].pw, n_b, p.seed_psi, psi[k]);
    res.n_pw[k] = ms[k].pw.n_pw;
  }
  std::vector<double> rho_in(n3, n_e / m0.sys.omega), rho_out, v, dscr;
  std::vector<double> w_state((size_t)n_k * n_b), e_all((size_t)n_k * n_b),
      occ((size_t)n_k * n_b);
  for (int k = 0; k < n_k; ++k)
    for (int i = 0; i < n_b; ++i) w_state[(size_t)k * n_b + i] = wk[k];
  Pulay mixer;
  mixer.depth = p.pulay_depth;
  mixer.alpha = p.mix_alpha;
  mixer.kerker_q = p.kerker_q;
  res.energies.clear();
  res.drho.clear();
  res.n_iter = 0;
  res.converged = 0;
  int small = 0;
  double e_prev = 0;
  for (int it = 1; it <= p.max_iter; ++it) {
    double eh_in, exc_in;
    vhxc(m0.sys, rho_in, v, &eh_in, &exc_in);
    for (int k = 0; k < n_k; ++k) {
      ms[k].veff.resize(n3);
      for (size_t r = 0; r < n3; ++r) ms[k].veff[r] = vion[r] + v[r];
    }
    if (m0.pj.n_proj) {
      screen_d(ms[0], dscr);
      for (int k = 0; k < n_k; ++k) ms[k].pj.d = dscr;
    }
    for (int k = 0; k < n_k; ++k) {
      const double lmax = p.lanczos_k > 0 ? lanczos_bound(ms[k], p.lanczos_k, p.seed_lanczos, nullptr)
                                          : grid_lambda_max(ms[k]);
      const int passes = it == 1 ? p.first_passes : 1;
      for (int ps = 0; ps < passes; ++ps) {
        Bounds bd = spectral_bounds(eps[k].data(), (int)eps[k].size(), lmax);
        std::vector<cd> y;
        double diag[4];
        int rc = chebyshev_filter(ms[k], psi[k].data(), n_b, p.cheb_degree, bd, y, diag);
        if (rc) return rc;
        rc = rayleigh_ritz(ms[k], y, n_b, eps[k], nullptr);
        if (rc) return 1000 + rc;
        psi[k].swap(y);
      }
      std::copy(eps[k].begin(), eps[k].end(), e_all.begin() + (size_t)k * n_b);
    }
    double ts;
    res.mu = fermi_occupations_w(e_all.data(), w_state.data(), n_k * n_b, n_e, p.smear_kt,
                                 occ.data(), &ts);
    rho_out.assign(n3, 0.0);
    double eband = 0;
    for (int k = 0; k < n_k; ++k) {
      std::vector<double> ok(n_b), rk;
      for (int i = 0; i < n_b; ++i) {
        ok[i] = wk[k] * occ[(size_t)k * n_b + i];
        eband += ok[i] * eps[k][i];
      }
      density(ms[k], psi[k].data(), n_b, ok.data(), rk);
      for (size_t r = 0; r < n3; ++r) rho_out[r] += rk[r];
    }
    double eh, exc;
    std::vector<double> vout;
    vhxc(m0.sys, rho_out, vout, &eh, &exc);
    double dc = 0, drho = 0;
    for (size_t r = 0; r < n3; ++r) {
      dc += v[r] * rho_out[r];
      drho += std::abs(rho_out[r] - rho_in[r]);
    }
    const double dv = m0.sys.omega / (double)n3;
    const double etot = eband - dc * dv + eh + exc + ts;
    res.energies.push_back(etot);
    res.drho.push_back(drho * dv);
    res.n_iter = it;
    if (it > 1) {
      double de = std::abs(etot - e_prev) / m0.sys.n_atoms;
      small = de < p.etol ? small + 1 : 0;
    }
    e_prev = etot;
    if (small >= 2) {
      res.converged = 1;
      break;
    }
    std::vector<double> rnext;
    mixer.mix(m0.sys, rho_in, rho_out, rnext);
    rho_in.swap(rnext);
  }
  res.eps = e_all;
  res.rho = rho_out;
  return 0;
}

}  // namespace orc
