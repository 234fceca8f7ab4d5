// Chebyshev filter, spectral bounds, Lanczos bound and Rayleigh-Ritz for the
// FP64 oracle. Test infrastructure only (see orc.hpp).
#include "eigensolver.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

static double frob(const cd* x, size_t n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < (long)n; ++i) s += std::norm(x[i]);
  return std::sqrt(s);
}

// Operator A = H, or S^{-1} H for the generalized (USPP) problem.
void Model::apply_a(const cd* x, int n_b, cd* ax, cd* sx_opt) const {
  apply_h(pw, sys.n_grid, veff.data(), pj.n_proj ? &pj : nullptr, x, n_b, ax, sx_opt);
  if (pj.n_proj && sinv_filter) apply_sinv(pj, pw.n_pw, ax, n_b);
}

// Reference src/eigensolver/chebyshev.rs:214-254 (compute_spectral_bounds):
// lambda_max = 0.5 gmax^2 + (Vmax - Vmin) with gmax the grid corner, or an
// externally supplied upper bound (the Lanczos estimate); eps_cut =
// lambda_max/3 without eigenvalues, else min(e_last + 0.2 (e_last - e0),
// 0.95 lambda_max).
Bounds spectral_bounds(const double* eig, int n_eig, double lambda_max) {
  Bounds b;
  b.lambda_max = lambda_max;
  if (n_eig <= 0) {
    b.eps_cut = lambda_max / 3.0;
  } else {
    const double e0 = eig[0], el = eig[n_eig - 1];
    b.eps_cut = std::min(el + 0.2 * (el - e0), lambda_max * 0.95);
  }
  b.center = 0.5 * (b.lambda_max + b.eps_cut);
  b.half_width = 0.5 * (b.lambda_max - b.eps_cut);
  return b;
}

double grid_lambda_max(const Model& m) {
  const double b = (2.0 * M_PI) / m.sys.L;
  const double gmax = b * std::sqrt(3.0) * (m.sys.n_grid / 2);
  double vmin = m.veff[0], vmax = m.veff[0];
  for (double v : m.veff) {
    vmin = std::min(vmin, v);
    vmax = std::max(vmax, v);
  }
  return 0.5 * gmax * gmax + (vmax - vmin);
}

// k-step Lanczos upper bound (Zhou & Li 2011, eq. 2.5; the filtering upper
// bound of Zhou et al. 2014 eq. 12), run in the S inner product for the
// generalized problem. Start vector: complex N(0,1) per plane wave from the
// LCG with `seed`. Returns b_up = max eig(T_k) + ||f_k||_S; *ritz_min gets
// min eig(T_k) for eq. 13 lower bounds.
double lanczos_bound(const Model& m, int k, uint64_t seed, double* ritz_min) {
  const int n = m.pw.n_pw;
  std::vector<cd> v(n), sv(n), vp(n, cd(0)), svp(n, cd(0)), u(n), w(n), f(n), sf(n);
  Lcg rng(seed);
  for (int g = 0; g < n; ++g) {
    double re = rng.normal();
    double im = rng.normal();
    v[g] = cd(re, im);
  }
  const bool gen = m.pj.n_proj > 0;
  auto s_apply = [&](const cd* x, cd* y) {
    std::vector<cd> tmp(n);
    apply_h(m.pw, m.sys.n_grid, m.veff.data(), &m.pj, x, 1, tmp.data(), y);
  };
  if (gen) s_apply(v.data(), sv.data()); else sv = v;
  double nrm = 0;
  for (int g = 0; g < n; ++g) nrm += (std::conj(v[g]) * sv[g]).real();
  nrm = std::sqrt(nrm);
  for (int g = 0; g < n; ++g) { v[g] /= nrm; sv[g] /= nrm; }
  std::vector<double> alpha(k), beta(k);
  double b = 0;
  for (int j = 0; j < k; ++j) {
    apply_h(m.pw, m.sys.n_grid, m.veff.data(), gen ? &m.pj : nullptr, v.data(), 1, u.data(), nullptr);
    w = u;
    if (gen) apply_sinv(m.pj, n, w.data(), 1);
    double a = 0;
    for (int g = 0; g < n; ++g) a += (std::conj(v[g]) * u[g]).real();
    for (int g = 0; g < n; ++g) {
      f[g] = w[g] - a * v[g] - b * vp[g];
      sf[g] = (gen ? u[g] : w[g]) - a * sv[g] - b * svp[g];
    }
    double bn = 0;
    for (int g = 0; g < n; ++g) bn += (std::conj(f[g]) * sf[g]).real();
    bn = std::sqrt(std::max(bn, 0.0));
    alpha[j] = a;
    beta[j] = bn;
    vp = v;
    svp = sv;
    if (bn > 0)
      for (int g = 0; g < n; ++g) { v[g] = f[g] / bn; sv[g] = sf[g] / bn; }
    b = bn;
  }
  std::vector<cd> t((size_t)k * k, cd(0));
  for (int j = 0; j < k; ++j) {
    t[(size_t)j * k + j] = alpha[j];
    if (j + 1 < k) {
      t[(size_t)(j + 1) * k + j] = beta[j];
      t[(size_t)j * k + j + 1] = beta[j];
    }
  }
  std::vector<double> ev;
  std::vector<cd> z;
  jacobi_eigh(t, k, ev, z);
  if (ritz_min) *ritz_min = ev[0];
  return ev[k - 1] + beta[k - 1];
}

// Degree-m Chebyshev filter, src/eigensolver/chebyshev.rs:626-816:
// Y1 = (A Y0 - c Y0)/e; Y_k = 2 (A Y_{k-1} - c Y_{k-1})/e - Y_{k-2};
// after each step ||Y_k||_F / ||Y_{k-1}||_F > 10 is ChebyshevDiverged
// (chebyshev.rs:546-563, loop :764-767). Returns 0 or -k (the failing step)
// with the norms in *diag. On success y holds Y_m.
int chebyshev_filter(const Model& m, const cd* x, int n_b, int degree, const Bounds& bd,
                     std::vector<cd>& y, double* diag) {
  const size_t ne = (size_t)n_b * m.pw.n_pw;
  std::vector<cd> a(x, x + ne), bb(ne), c(ne), ay(ne);
  if (degree == 0) {
    y = a;
    return 0;
  }
  const double inv_e = 1.0 / bd.half_width, cc = bd.center;
  m.apply_a(a.data(), n_b, ay.data(), nullptr);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)ne; ++i) bb[i] = (ay[i] - cc * a[i]) * inv_e;
  for (int k = 2; k <= degree; ++k) {
    m.apply_a(bb.data(), n_b, ay.data(), nullptr);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)ne; ++i) c[i] = 2.0 * (ay[i] - cc * bb[i]) * inv_e - a[i];
    const double np = frob(bb.data(), ne), nc = frob(c.data(), ne);
    const double growth = np > 0 ? nc / np : 1.0;
    if (growth > 10.0) {
      if (diag) { diag[0] = k; diag[1] = np; diag[2] = nc; diag[3] = growth; }
      return -k;
    }
    a.swap(bb);
    bb.swap(c);
  }
  y = bb;
  return 0;
}

// Rayleigh-Ritz with the Cholesky reduction of subspace_diag.rs:229-391
// (the semantics the reference's GPU rayleigh_ritz.rs:48-245 intends).
// On return y is rotated in place to the Ritz vectors; hy/sy (if non-null)
// are rotated consistently.
int rayleigh_ritz(const Model& m, std::vector<cd>& y, int n_b, std::vector<double>& eps,
                  std::vector<cd>* hy_out) {
  const int n_pw = m.pw.n_pw;
  const size_t ne = (size_t)n_b * n_pw;
  std::vector<cd> hy(ne), sy(ne);
  apply_h(m.pw, m.sys.n_grid, m.veff.data(), m.pj.n_proj ? &m.pj : nullptr, y.data(), n_b,
          hy.data(), sy.data());
  std::vector<cd> hs, ss, v;
  subspace_matrix(y.data(), hy.data(), n_pw, n_b, hs);
  subspace_matrix(y.data(), sy.data(), n_pw, n_b, ss);
  int info = gen_eigh(hs, ss, n_b, eps, v);
  if (info) return info;
  std::vector<cd> out(ne);
  rotate(y.data(), v, n_pw, n_b, out.data());
  y.swap(out);
  if (hy_out) {
    hy_out->resize(ne);
    rotate(hy.data(), v, n_pw, n_b, hy_out->data());
  }
  return 0;
}

}  // namespace orc
