// Band-by-band CG stack for the FP64 oracle: closed-form line search, CG
// helpers, TPA + USPP preconditioner, band CG minimiser.
// Test infrastructure only (see orc.hpp).
#include "bandcg.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace orc {

static cd dot(const cd* a, const cd* b, int n) {
  cd s = 0;
  for (int i = 0; i < n; ++i) s += std::conj(a[i]) * b[i];
  return s;
}

// src/eigensolver/line_search.rs:57-198 (line_search_2d_quadratic),
// CASTEP electronic_ideal_step_size.
LineSearchResult line_search_2d_quadratic(double dot_hpsi_psi, double dot_hpsi_dir,
                                          double dot_hdir_dir, double dot_sdir_dir) {
  const double a = dot_hpsi_psi, b = -2.0 * dot_hpsi_dir, c = dot_hdir_dir, d = dot_sdir_dir;
  double step, ev;
  int status;
  const LineSearchResult fail{0.0, a, 1.0, -1};
  if (std::abs(d) > DBL_MIN) {
    const double ad = a * d, bd = b * d;
    if (std::abs(bd) < 1e-30) return LineSearchResult{0.0, a, 1.0, -2};
    const double tmp = ad * ad - 2.0 * ad * c + c * c + b * bd;
    if (tmp < 0.0) return fail;
    const double det = std::sqrt(tmp);
    if (std::abs(bd) < DBL_MIN) return fail;
    const double r1 = (-ad + c + det) / bd, r2 = (-ad + c - det) / bd;
    auto eval = [&](double r, double* e) {
      const double den = 1.0 + d * r * r;
      if (den > 0.0) {
        *e = (a + b * r + c * r * r) / den;
        return true;
      }
      return false;
    };
    double e1, e2;
    const bool ok1 = eval(r1, &e1), ok2 = eval(r2, &e2);
    double r, e;
    if (ok1 && ok2) {
      if (e1 < e2) { r = r1; e = e1; } else { r = r2; e = e2; }
    } else if (ok1 && !ok2 && e1 < a) {
      r = r1; e = e1;
    } else if (!ok1 && ok2 && e2 < a) {
      r = r2; e = e2;
    } else {
      return fail;
    }
    step = -r;
    ev = e;
    status = 0;
  } else {
    if (c > 0.0) {
      const double r = -b / (2.0 * c);
      step = r;
      ev = (a + b * r + c * r * r) / (1.0 + d * r * r);
      status = 0;
    } else {
      step = 0.0;
      ev = a;
      status = -1;
    }
  }
  if (step > 15.0) {
    const double s = 15.0;
    ev = (a + b * s + c * s * s) / (1.0 + d * s * s);
    step = s;
    status = 1;
  } else if (step < 0.0) {
    status = -1;
  }
  const double t = 1.0 + d * step * step;
  LineSearchResult r;
  r.step_size = step;
  r.eigenvalue = ev;
  if (t > 0.0) {
    r.norm = std::sqrt(t);
    r.status = status;
  } else {
    r.norm = 0.0;
    r.status = 2;
  }
  return r;
}

// src/eigensolver/cg_helpers.rs:49-82 (compute_residual). Returns 0, or -1
// when |<psi|S|psi>| <= DBL_EPSILON (the reference asserts).
int compute_residual(const cd* psi, const cd* hpsi, const cd* spsi, int n, cd* r, double* eps) {
  if (n == 0) {
    *eps = NAN;
    return 0;
  }
  const double he = dot(psi, hpsi, n).real(), se = dot(psi, spsi, n).real();
  if (!(std::abs(se) > DBL_EPSILON)) return -1;
  const double e = he / se;
  for (int i = 0; i < n; ++i) r[i] = hpsi[i] - e * spsi[i];
  *eps = e;
  return 0;
}

// cg_helpers.rs:97-103
double residual_bare_norm(const cd* r, int n) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += std::norm(r[i]);
  return std::sqrt(s);
}

// cg_helpers.rs:121-136. Returns -1 when <psi|S|psi> < -1e-14.
int s_normalize(const cd* psi, const cd* spsi, int n, cd* out) {
  const double so = dot(psi, spsi, n).real();
  if (!(so > -1e-14)) return -1;
  if (std::abs(so - 1.0) < 1e-14) {
    std::copy(psi, psi + n, out);
    return 0;
  }
  const double nrm = std::sqrt(so);
  for (int i = 0; i < n; ++i) out[i] = psi[i] / nrm;
  return 0;
}

// cg_helpers.rs:161-216: modified Gram-Schmidt in the S metric using the
// real part of the overlap, then S-normalise. Returns -1 when the final
// <psi|S|psi> is not positive. With complex vectors the real part leaves
// i Im<psi_j|S psi> psi_j behind; full_overlap removes the whole overlap.
int s_orthogonalize_against(const cd* psi, const cd* spsi, int n, int n_conv, const cd* conv_psi,
                            const cd* conv_spsi, cd* out_psi, cd* out_spsi, bool full_overlap) {
  std::vector<cd> p(psi, psi + n), s(spsi, spsi + n);
  for (int j = 0; j < n_conv; ++j) {
    const cd* pj = conv_psi + (size_t)j * n;
    const cd* sj = conv_spsi + (size_t)j * n;
    if (full_overlap) {
      const cd ov = dot(pj, s.data(), n);
      for (int i = 0; i < n; ++i) p[i] -= ov * pj[i];
      for (int i = 0; i < n; ++i) s[i] -= ov * sj[i];
      continue;
    }
    const double ov = dot(pj, s.data(), n).real();
    for (int i = 0; i < n; ++i) p[i] -= ov * pj[i];
    for (int i = 0; i < n; ++i) s[i] -= ov * sj[i];
  }
  const double so = dot(p.data(), s.data(), n).real();
  if (!(so > 0.0)) return -1;
  const double nrm = std::sqrt(so);
  for (int i = 0; i < n; ++i) {
    out_psi[i] = p[i] / nrm;
    out_spsi[i] = s[i] / nrm;
  }
  return 0;
}

// src/eigensolver/uspp_preconditioner.rs:148-180 (compute_tpa_diag).
void tpa_diag(const double* kin, int n, double* t) {
  if (n == 0) return;
  double sum = 0;
  for (int i = 0; i < n; ++i) sum += kin[i];
  const double mean = sum / n;
  const double inv = mean > 0.0 ? 1.0 / mean : 1.0;
  for (int i = 0; i < n; ++i) {
    const double x = kin[i] * inv;
    const double tmp = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
    t[i] = 1.0 / (1.0 + 16.0 * std::pow(x, 4) / tmp);
  }
}

// General dense inverse via Gauss-Jordan with full pivoting (the reference
// uses faer's full_piv_lu().inverse(), uspp_preconditioner.rs:222-244).
static bool cinverse(std::vector<cd>& a, int n) {
  std::vector<int> rowp(n), colp(n);
  std::vector<bool> used(n, false);
  std::vector<int> ipiv(n, 0), indxr(n), indxc(n);
  for (int i = 0; i < n; ++i) {
    double big = 0;
    int irow = -1, icol = -1;
    for (int j = 0; j < n; ++j)
      if (ipiv[j] != 1)
        for (int k = 0; k < n; ++k)
          if (ipiv[k] == 0 && std::abs(a[(size_t)k * n + j]) >= big) {
            big = std::abs(a[(size_t)k * n + j]);
            irow = j;
            icol = k;
          }
    if (irow < 0 || big == 0.0) return false;
    ++ipiv[icol];
    if (irow != icol)
      for (int l = 0; l < n; ++l) std::swap(a[(size_t)l * n + irow], a[(size_t)l * n + icol]);
    indxr[i] = irow;
    indxc[i] = icol;
    const cd pivinv = 1.0 / a[(size_t)icol * n + icol];
    a[(size_t)icol * n + icol] = 1.0;
    for (int l = 0; l < n; ++l) a[(size_t)l * n + icol] *= pivinv;
    for (int ll = 0; ll < n; ++ll)
      if (ll != icol) {
        const cd dum = a[(size_t)icol * n + ll];
        a[(size_t)icol * n + ll] = 0.0;
        for (int l = 0; l < n; ++l) a[(size_t)l * n + ll] -= a[(size_t)l * n + icol] * dum;
      }
  }
  for (int l = n - 1; l >= 0; --l)
    if (indxr[l] != indxc[l])
      for (int k = 0; k < n; ++k) std::swap(a[(size_t)indxr[l] * n + k], a[(size_t)indxc[l] * n + k]);
  return true;
}

// uspp_preconditioner.rs:56-73 (new), :187-214 (C = beta^H T^-1 beta),
// :222-244 (R = (-Q^-1 - C)^-1). beta is n_pw x n_proj row-major.
int UsppPrecond::init(const cd* beta_g, int n_pw_, int n_proj_, const cd* q_matrix, const double* kin) {
  n_pw = n_pw_;
  n_proj = n_proj_;
  tpa.resize(n_pw);
  tpa_diag(kin, n_pw, tpa.data());
  beta.assign(beta_g, beta_g + (size_t)n_pw * n_proj);
  c.assign((size_t)n_proj * n_proj, cd(0));
  for (int p = 0; p < n_proj; ++p)
    for (int q = 0; q < n_proj; ++q) {
      cd s = 0;
      for (int g = 0; g < n_pw; ++g)
        s += std::conj(beta[(size_t)g * n_proj + p]) * (beta[(size_t)g * n_proj + q] * tpa[g]);
      c[(size_t)q * n_proj + p] = s;
    }
  std::vector<cd> qi(q_matrix, q_matrix + (size_t)n_proj * n_proj);
  // q_matrix is row-major (ndarray); convert to column-major.
  for (int i = 0; i < n_proj; ++i)
    for (int j = 0; j < n_proj; ++j) qi[(size_t)j * n_proj + i] = q_matrix[(size_t)i * n_proj + j];
  if (!cinverse(qi, n_proj)) return -1;
  r.assign((size_t)n_proj * n_proj, cd(0));
  for (size_t i = 0; i < r.size(); ++i) r[i] = -qi[i] - c[i];
  if (!cinverse(r, n_proj)) return -2;
  return 0;
}

// uspp_preconditioner.rs:80-131 (apply).
void UsppPrecond::apply(const cd* res, cd* out) const {
  std::vector<cd> tr(n_pw), w(n_proj, cd(0)), wr(n_proj, cd(0));
  for (int g = 0; g < n_pw; ++g) tr[g] = res[g] * tpa[g];
  for (int p = 0; p < n_proj; ++p) {
    cd s = 0;
    for (int g = 0; g < n_pw; ++g) s += std::conj(beta[(size_t)g * n_proj + p]) * tr[g];
    w[p] = s;
  }
  for (int i = 0; i < n_proj; ++i) {
    cd s = 0;
    for (int j = 0; j < n_proj; ++j) s += r[(size_t)j * n_proj + i] * w[j];
    wr[i] = s;
  }
  for (int g = 0; g < n_pw; ++g) {
    cd s = 0;
    for (int p = 0; p < n_proj; ++p) s += beta[(size_t)g * n_proj + p] * wr[p];
    out[g] = tr[g] + s * tpa[g];
  }
}

// src/eigensolver/band_cg.rs:157-437 (band_cg_minimize), CASTEP
// electronic_find_eigenstate. apply_hs(x, hx, sx) applies H and S to one
// band; the preconditioner is `pre(res, out)`.
CgResult band_cg_minimize(const cd* psi_init, int n, const PrecondFn& pre, int n_conv,
                          const cd* conv_psi, const cd* conv_spsi, int max_steps, double tol,
                          const HsFn& apply_hs, bool full_overlap) {
  std::vector<cd> h0(n), s0(n);
  apply_hs(psi_init, h0.data(), s0.data());
  CgResult out;
  const double so = dot(psi_init, s0.data(), n).real();
  if (!(so > 0.0)) {
    out.status = -1;
    return out;
  }
  const double inv = 1.0 / std::sqrt(so);
  std::vector<cd> psi(n), hpsi(n), spsi(n);
  for (int i = 0; i < n; ++i) {
    psi[i] = psi_init[i] * inv;
    hpsi[i] = h0[i] * inv;
    spsi[i] = s0[i] * inv;
  }
  double eig = dot(psi.data(), hpsi.data(), n).real();
  std::vector<cd> r(n), g(n), sg(n), go(n), sgo(n), d(n), d_old, hd(n), sd(n);
  double tmp_e;
  compute_residual(psi.data(), hpsi.data(), spsi.data(), n, r.data(), &tmp_e);
  double beta_old = 0;
  int step_type = 0;  // 0 = SD, 1 = CG
  for (int step = 1; step <= max_steps; ++step) {
    pre(r.data(), g.data());
    if (residual_bare_norm(g.data(), n) < DBL_EPSILON * 1000.0) {
      out.psi = psi;
      out.eigenvalue = eig;
      out.n_steps = step - 1;
      out.converged = 1;
      out.residual_norm = residual_bare_norm(r.data(), n);
      out.step_type = step_type;
      return out;
    }
    const double prev = eig;
    std::vector<cd> htmp(n);
    apply_hs(g.data(), htmp.data(), sg.data());
    if (s_orthogonalize_against(g.data(), sg.data(), n, n_conv, conv_psi, conv_spsi, go.data(),
                                sgo.data(), full_overlap) != 0) {
      out.status = -2;
      return out;
    }
    if (step <= 2) {
      step_type = 0;
      d = go;
      beta_old = dot(go.data(), r.data(), n).real();
    } else {
      step_type = 1;
      const double beta = dot(go.data(), r.data(), n).real();
      const double gamma = std::abs(beta_old) > 1e-300 ? beta / beta_old : 0.0;
      for (int i = 0; i < n; ++i) d[i] = gamma * d_old[i] - go[i];
      beta_old = beta;
    }
    apply_hs(d.data(), hd.data(), sd.data());
    const cd ov = dot(psi.data(), sd.data(), n);
    if (std::abs(ov) > 1e-30) {
      for (int i = 0; i < n; ++i) {
        d[i] -= ov * psi[i];
        hd[i] -= ov * hpsi[i];
        sd[i] -= ov * spsi[i];
      }
    }
    d_old = d;
    const double rn = residual_bare_norm(r.data(), n);
    if (rn < tol * 0.01 && rn < 1e-12 * std::max(std::abs(eig), 1.0)) {
      out.psi = psi;
      out.eigenvalue = eig;
      out.n_steps = step - 1;
      out.converged = 1;
      out.residual_norm = rn;
      out.step_type = step_type;
      return out;
    }
    const LineSearchResult ls =
        line_search_2d_quadratic(eig, dot(hpsi.data(), d.data(), n).real(),
                                 dot(hd.data(), d.data(), n).real(), dot(sd.data(), d.data(), n).real());
    const double s = ls.step_size;
    const double in = ls.norm > 0.0 ? 1.0 / ls.norm : 1.0;
    for (int i = 0; i < n; ++i) {
      psi[i] = (psi[i] + s * d[i]) * in;
      hpsi[i] = (hpsi[i] + s * hd[i]) * in;
      spsi[i] = (spsi[i] + s * sd[i]) * in;
    }
    const double e_new = dot(psi.data(), hpsi.data(), n).real();
    compute_residual(psi.data(), hpsi.data(), spsi.data(), n, r.data(), &tmp_e);
    if (std::abs(e_new - prev) < tol) {
      out.psi = psi;
      out.eigenvalue = e_new;
      out.n_steps = step;
      out.converged = 1;
      out.residual_norm = residual_bare_norm(r.data(), n);
      out.step_type = step_type;
      return out;
    }
    eig = e_new;
  }
  out.psi = psi;
  out.eigenvalue = eig;
  out.n_steps = max_steps;
  out.converged = 0;
  out.residual_norm = residual_bare_norm(r.data(), n);
  out.step_type = step_type;
  return out;
}

}  // namespace orc
