// C ABI of the FP64 oracle. Test infrastructure only (see orc.hpp).
#include "oracle.h"

#include <algorithm>
#include <cstring>

#include "bandcg.hpp"
#include "scf.hpp"

using namespace orc;

#include <omp.h>

void orc_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
void orc_downsample(const double* fine, int nf, int nw, double* out) {
  orc::downsample(fine, nf, nw, out);
}
int orc_get_max_threads(void) { return omp_get_max_threads(); }

static ScfParams to_params(const orc_params* p) {
  ScfParams s;
  s.n_rep = p->n_rep;
  s.a0 = p->a0;
  s.zval = p->zval;
  s.rc_loc = p->rc_loc;
  s.ecut = p->ecut;
  s.n_grid = p->n_grid;
  s.n_bands = p->n_bands;
  for (int i = 0; i < 3; ++i) s.kfrac[i] = p->kfrac[i];
  s.n_proj_per_atom = p->n_proj_per_atom;
  s.beta_rc = p->beta_rc;
  s.q_rc = p->q_rc;
  s.q0 = p->q0;
  s.d_min = p->d_min;
  s.d_max = p->d_max;
  s.seed_d = p->seed_d;
  s.seed_psi = p->seed_psi;
  s.seed_lanczos = p->seed_lanczos;
  s.cheb_degree = p->cheb_degree;
  s.first_passes = p->first_passes;
  s.lanczos_k = p->lanczos_k;
  s.pulay_depth = p->pulay_depth;
  s.mix_alpha = p->mix_alpha;
  s.kerker_q = p->kerker_q;
  s.etol = p->etol;
  s.smear_kt = p->smear_kt;
  s.max_iter = p->max_iter;
  return s;
}

static cd* C(double* x) { return reinterpret_cast<cd*>(x); }
static const cd* C(const double* x) { return reinterpret_cast<const cd*>(x); }

extern "C" {

void orc_default_params(orc_params* p) {
  ScfParams s;
  std::memset(p, 0, sizeof(*p));
  p->n_rep = s.n_rep;
  p->a0 = s.a0;
  p->zval = s.zval;
  p->rc_loc = s.rc_loc;
  p->ecut = s.ecut;
  p->n_grid = s.n_grid;
  p->n_bands = s.n_bands;
  p->n_proj_per_atom = s.n_proj_per_atom;
  p->beta_rc = s.beta_rc;
  p->q_rc = s.q_rc;
  p->q0 = s.q0;
  p->d_min = s.d_min;
  p->d_max = s.d_max;
  p->seed_d = s.seed_d;
  p->seed_psi = s.seed_psi;
  p->seed_lanczos = s.seed_lanczos;
  p->cheb_degree = s.cheb_degree;
  p->first_passes = s.first_passes;
  p->lanczos_k = s.lanczos_k;
  p->pulay_depth = s.pulay_depth;
  p->mix_alpha = s.mix_alpha;
  p->kerker_q = s.kerker_q;
  p->etol = s.etol;
  p->smear_kt = s.smear_kt;
  p->max_iter = s.max_iter;
}

double orc_lcg_uniforms(uint64_t seed, int n, double* out) {
  Lcg r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
  return n ? out[n - 1] : 0.0;
}

void orc_lcg_normals(uint64_t seed, int n, double* out) {
  Lcg r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.normal();
}

void orc_synth_positions(int n_rep, double a0, double sigma, uint64_t seed, double* pos) {
  synth_positions(n_rep, a0, sigma, seed, pos);
}

int orc_gsphere(double L, double ecut, const double kfrac[3], int n_grid, int max_pw,
                int32_t* miller, int32_t* fft_idx, double* kin) {
  PwSet pw;
  int n = gsphere(L, ecut, kfrac, n_grid, pw);
  if (n < 0) return n;
  if (miller && n <= max_pw) {
    std::copy(pw.miller.begin(), pw.miller.end(), miller);
    std::copy(pw.fft_idx.begin(), pw.fft_idx.end(), fft_idx);
    std::copy(pw.kin.begin(), pw.kin.end(), kin);
  }
  return n;
}

void orc_pw_coords_to_fft_indices(const int32_t* miller, int n_pw, int ngx, int ngy, int ngz,
                                  int32_t* idx) {
  pw_coords_to_fft_indices(miller, n_pw, ngx, ngy, ngz, idx);
}

void orc_fft3d(double* grid, int n, int sign) { fft3d(C(grid), n, sign); }

struct OrcModel {
  ScfParams p;
  Model m;
  std::vector<double> vion;
};

void* orc_model_create(const orc_params* p, const double* pos) {
  auto* om = new OrcModel;
  om->p = to_params(p);
  setup_model(om->p, pos, om->m);
  vion_real(om->m.sys, om->vion);
  om->m.veff = om->vion;
  return om;
}

void orc_model_destroy(void* m) { delete static_cast<OrcModel*>(m); }
int orc_model_n_pw(void* m) { return static_cast<OrcModel*>(m)->m.pw.n_pw; }
int orc_model_n_proj(void* m) { return static_cast<OrcModel*>(m)->m.pj.n_proj; }

void orc_model_get_pw(void* mm, int32_t* miller, int32_t* fft_idx, double* kin) {
  auto& pw = static_cast<OrcModel*>(mm)->m.pw;
  std::copy(pw.miller.begin(), pw.miller.end(), miller);
  std::copy(pw.fft_idx.begin(), pw.fft_idx.end(), fft_idx);
  std::copy(pw.kin.begin(), pw.kin.end(), kin);
}

void orc_model_get_vion(void* mm, double* vion) {
  auto* om = static_cast<OrcModel*>(mm);
  std::copy(om->vion.begin(), om->vion.end(), vion);
}

void orc_model_get_beta(void* mm, double* beta) {
  auto& pj = static_cast<OrcModel*>(mm)->m.pj;
  std::copy(pj.beta.begin(), pj.beta.end(), C(beta));
}

void orc_model_get_d(void* mm, double* d) {
  auto& pj = static_cast<OrcModel*>(mm)->m.pj;
  std::copy(pj.d.begin(), pj.d.end(), d);
}

void orc_model_set_d(void* mm, const double* d) {
  auto& pj = static_cast<OrcModel*>(mm)->m.pj;
  std::copy(d, d + pj.n_proj, pj.d.begin());
}

void orc_model_set_veff(void* mm, const double* veff) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  std::copy(veff, veff + m.veff.size(), m.veff.begin());
}

void orc_model_set_sinv_filter(void* mm, int on) {
  static_cast<OrcModel*>(mm)->m.sinv_filter = on != 0;
}

void orc_initial_psi(void* mm, int n_b, uint64_t seed, double* psi) {
  std::vector<cd> v;
  initial_psi(static_cast<OrcModel*>(mm)->m.pw, n_b, seed, v);
  std::copy(v.begin(), v.end(), C(psi));
}

void orc_apply_h(void* mm, const double* psi, int n_b, double* hpsi, double* spsi) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  apply_h(m.pw, m.sys.n_grid, m.veff.data(), m.pj.n_proj ? &m.pj : nullptr, C(psi), n_b, C(hpsi),
          spsi ? C(spsi) : nullptr);
}

void orc_apply_sinv(void* mm, double* x, int n_b) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  apply_sinv(m.pj, m.pw.n_pw, C(x), n_b);
}

double orc_grid_lambda_max(void* mm) { return grid_lambda_max(static_cast<OrcModel*>(mm)->m); }

double orc_lanczos_bound(void* mm, int k, uint64_t seed, double* ritz_min) {
  return lanczos_bound(static_cast<OrcModel*>(mm)->m, k, seed, ritz_min);
}

void orc_spectral_bounds(const double* eig, int n_eig, double lambda_max, double* out4) {
  Bounds b = spectral_bounds(eig, n_eig, lambda_max);
  out4[0] = b.lambda_max;
  out4[1] = b.eps_cut;
  out4[2] = b.center;
  out4[3] = b.half_width;
}

int orc_chebyshev_filter(void* mm, const double* psi, int n_b, int degree, const double* b4,
                         double* out, double* diag4) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  Bounds b{b4[0], b4[1], b4[2], b4[3]};
  std::vector<cd> y;
  int rc = chebyshev_filter(m, C(psi), n_b, degree, b, y, diag4);
  if (rc == 0) std::copy(y.begin(), y.end(), C(out));
  return rc;
}

int orc_rayleigh_ritz(void* mm, double* psi, int n_b, double* eps, double* hpsi_out) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  const size_t ne = (size_t)n_b * m.pw.n_pw;
  std::vector<cd> y(C(psi), C(psi) + ne), hy;
  std::vector<double> e;
  int rc = rayleigh_ritz(m, y, n_b, e, hpsi_out ? &hy : nullptr);
  if (rc) return rc;
  std::copy(y.begin(), y.end(), C(psi));
  std::copy(e.begin(), e.end(), eps);
  if (hpsi_out) std::copy(hy.begin(), hy.end(), C(hpsi_out));
  return 0;
}

void orc_density(void* mm, const double* psi, int n_b, const double* occ, double* rho) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  std::vector<double> r;
  density(m, C(psi), n_b, occ, r);
  std::copy(r.begin(), r.end(), rho);
}

void orc_vhxc(void* mm, const double* rho, double* v, double* eh, double* exc) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  const size_t n3 = m.veff.size();
  std::vector<double> r(rho, rho + n3), vv;
  vhxc(m.sys, r, vv, eh, exc);
  std::copy(vv.begin(), vv.end(), v);
}

void orc_screen_d(void* mm, double* d) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  std::vector<double> dd;
  screen_d(m, dd);
  std::copy(dd.begin(), dd.end(), d);
}

void orc_lda_pz(int n, const double* rho, double* exc, double* vxc) {
  for (int i = 0; i < n; ++i) lda_pz(rho[i], exc + i, vxc + i);
}

int orc_cholesky_lower(double* a, int n) {
  std::vector<cd> v(C(a), C(a) + (size_t)n * n);
  int rc = cholesky_lower(v, n);
  std::copy(v.begin(), v.end(), C(a));
  return rc;
}

void orc_jacobi_eigh(double* a, int n, double* w, double* v) {
  std::vector<cd> m(C(a), C(a) + (size_t)n * n), vv;
  std::vector<double> ww;
  jacobi_eigh(m, n, ww, vv);
  std::copy(ww.begin(), ww.end(), w);
  std::copy(vv.begin(), vv.end(), C(v));
}

void orc_subspace_matrix(const double* x, const double* y, int n_pw, int n_b, double* m) {
  std::vector<cd> mm;
  subspace_matrix(C(x), C(y), n_pw, n_b, mm);
  std::copy(mm.begin(), mm.end(), C(m));
}

int orc_gen_eigh(const double* h, const double* s, int n, double* eps, double* v) {
  std::vector<cd> hh(C(h), C(h) + (size_t)n * n), ss(C(s), C(s) + (size_t)n * n), vv;
  std::vector<double> e;
  int rc = gen_eigh(hh, ss, n, e, vv);
  if (rc) return rc;
  std::copy(e.begin(), e.end(), eps);
  std::copy(vv.begin(), vv.end(), C(v));
  return 0;
}

double orc_fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ,
                             double* ts) {
  return fermi_occupations(eps, n, n_e, kt, occ, ts);
}

int orc_scf_run(const orc_params* p, const double* pos, int* n_iter, int* converged,
                double* energies, double* drho, double* eps, double* rho) {
  ScfParams s = to_params(p);
  ScfResult r;
  int rc = run_scf(s, pos, r, nullptr);
  if (rc) return rc;
  *n_iter = r.n_iter;
  *converged = r.converged;
  std::copy(r.energies.begin(), r.energies.end(), energies);
  std::copy(r.drho.begin(), r.drho.end(), drho);
  std::copy(r.eps.begin(), r.eps.end(), eps);
  if (rho) std::copy(r.rho.begin(), r.rho.end(), rho);
  return 0;
}

int orc_scf_run_ex(const orc_params* p, const double* pos, int* n_iter, int* converged,
                   double* energies, double* drho, double* eps, double* rho, double* psi,
                   double* occ, double* mu) {
  ScfParams s = to_params(p);
  ScfResult r;
  int rc = run_scf(s, pos, r, nullptr);
  if (rc) return rc;
  *n_iter = r.n_iter;
  *converged = r.converged;
  std::copy(r.energies.begin(), r.energies.end(), energies);
  std::copy(r.drho.begin(), r.drho.end(), drho);
  std::copy(r.eps.begin(), r.eps.end(), eps);
  if (rho) std::copy(r.rho.begin(), r.rho.end(), rho);
  if (psi) std::copy(r.psi.begin(), r.psi.end(), C(psi));
  if (occ) std::copy(r.occ.begin(), r.occ.end(), occ);
  if (mu) *mu = r.mu;
  return 0;
}

int orc_scf_run_k(const orc_params* p, const double* pos, int n_k, const double* kfrac,
                  const double* wk, int* n_iter, int* converged, double* energies, double* drho,
                  double* eps, double* rho, int* n_pw, double* mu) {
  ScfParams s = to_params(p);
  ScfResultK r;
  int rc = run_scf_k(s, pos, n_k, kfrac, wk, r);
  if (rc) return rc;
  *n_iter = r.n_iter;
  *converged = r.converged;
  std::copy(r.energies.begin(), r.energies.end(), energies);
  std::copy(r.drho.begin(), r.drho.end(), drho);
  std::copy(r.eps.begin(), r.eps.end(), eps);
  if (rho) std::copy(r.rho.begin(), r.rho.end(), rho);
  std::copy(r.n_pw.begin(), r.n_pw.end(), n_pw);
  *mu = r.mu;
  return 0;
}

double orc_fermi_occupations_w(const double* eps, const double* w, int n, double n_e, double kt,
                               double* occ, double* ts) {
  return fermi_occupations_w(eps, w, n, n_e, kt, occ, ts);
}

struct OrcPulay {
  System sys;
  Pulay mixer;
};

void* orc_pulay_create(int n_grid, double L, int depth, double alpha, double kerker_q) {
  auto* h = new OrcPulay;
  h->sys.n_grid = n_grid;
  h->sys.L = L;
  h->sys.omega = L * L * L;
  h->mixer.depth = depth;
  h->mixer.alpha = alpha;
  h->mixer.kerker_q = kerker_q;
  return h;
}

void orc_pulay_mix(void* h, const double* rin, const double* rout, double* rnext) {
  auto* p = static_cast<OrcPulay*>(h);
  const size_t n3 = (size_t)p->sys.n_grid * p->sys.n_grid * p->sys.n_grid;
  std::vector<double> a(rin, rin + n3), b(rout, rout + n3), c;
  p->mixer.mix(p->sys, a, b, c);
  std::copy(c.begin(), c.end(), rnext);
}

void orc_pulay_destroy(void* h) { delete static_cast<OrcPulay*>(h); }

void orc_line_search(double a, double hd, double dd, double sd, double* out3, int* status) {
  LineSearchResult r = line_search_2d_quadratic(a, hd, dd, sd);
  out3[0] = r.step_size;
  out3[1] = r.eigenvalue;
  out3[2] = r.norm;
  *status = r.status;
}

int orc_compute_residual(const double* psi, const double* hpsi, const double* spsi, int n,
                         double* r, double* eps) {
  return compute_residual(C(psi), C(hpsi), C(spsi), n, C(r), eps);
}

double orc_residual_bare_norm(const double* r, int n) { return residual_bare_norm(C(r), n); }

int orc_s_normalize(const double* psi, const double* spsi, int n, double* out) {
  return s_normalize(C(psi), C(spsi), n, C(out));
}

int orc_s_orthogonalize_against(const double* psi, const double* spsi, int n, int n_conv,
                                const double* conv_psi, const double* conv_spsi, double* out_psi,
                                double* out_spsi) {
  return s_orthogonalize_against(C(psi), C(spsi), n, n_conv, C(conv_psi), C(conv_spsi),
                                 C(out_psi), C(out_spsi));
}

void orc_tpa_diag(const double* kin, int n, double* t) { tpa_diag(kin, n, t); }

int orc_uspp_precond_apply(const double* beta, int n_pw, int n_proj, const double* q,
                           const double* kin, const double* res, int n_vec, double* out,
                           double* r_matrix, double* c_matrix) {
  UsppPrecond pc;
  int rc = pc.init(C(beta), n_pw, n_proj, C(q), kin);
  if (rc) return rc;
  for (int v = 0; v < n_vec; ++v)
    pc.apply(C(res) + (size_t)v * n_pw, C(out) + (size_t)v * n_pw);
  if (r_matrix) std::copy(pc.r.begin(), pc.r.end(), C(r_matrix));
  if (c_matrix) std::copy(pc.c.begin(), pc.c.end(), C(c_matrix));
  return 0;
}

// band_cg_minimize for every band of a block, band i against the converged
// set {(psi_j, S psi_j), j < i} of the block as handed in, with the model's
// H and S. precond 0: TPA only (beta = 0, Q = I); 1: the full USPP
// preconditioner with the model's projectors and Q; + 256: the lower-band
// S-orthogonalisation removes the complex overlap (not only its real part).
int orc_band_cg_model(void* mm, double* psi_io, int n_b, int max_steps, double tol, int precond,
                      double* eps, int* n_steps, int* converged) {
  auto& m = static_cast<OrcModel*>(mm)->m;
  const bool full = (precond & 256) != 0;
  precond &= 255;
  const int n = m.pw.n_pw;
  const Projectors* pj = m.pj.n_proj ? &m.pj : nullptr;
  std::vector<cd> in(C(psi_io), C(psi_io) + (size_t)n_b * n), sin((size_t)n_b * n), tmp(n);
  for (int b = 0; b < n_b; ++b)
    apply_h(m.pw, m.sys.n_grid, m.veff.data(), pj, in.data() + (size_t)b * n, 1, tmp.data(),
            sin.data() + (size_t)b * n);
  UsppPrecond pc;
  int rc;
  if (precond == 1 && pj && m.sys.q0 != 0.0) {
    const int P = pj->n_proj;
    std::vector<cd> beta_rm((size_t)n * P), q((size_t)P * P, cd(0));
    for (int p = 0; p < P; ++p) {
      for (int g = 0; g < n; ++g) beta_rm[(size_t)g * P + p] = pj->beta[(size_t)p * n + g];
      q[(size_t)p * P + p] = pj->q[p];
    }
    rc = pc.init(beta_rm.data(), n, P, q.data(), m.pw.kin.data());
  } else {
    std::vector<cd> beta0(n, cd(0)), q1(1, cd(1));
    rc = pc.init(beta0.data(), n, 1, q1.data(), m.pw.kin.data());
  }
  if (rc) return rc;
  HsFn hs = [&](const cd* x, cd* hx, cd* sx) {
    apply_h(m.pw, m.sys.n_grid, m.veff.data(), pj, x, 1, hx, sx);
  };
  PrecondFn pre = [&](const cd* r, cd* o) { pc.apply(r, o); };
  for (int b = 0; b < n_b; ++b) {
    CgResult res = band_cg_minimize(in.data() + (size_t)b * n, n, pre, b, in.data(), sin.data(),
                                    max_steps, tol, hs, full);
    if (res.status) return res.status;
    std::copy(res.psi.begin(), res.psi.end(), C(psi_io) + (size_t)b * n);
    eps[b] = res.eigenvalue;
    n_steps[b] = res.n_steps;
    converged[b] = res.converged;
  }
  return 0;
}

int orc_band_cg_dense(const double* psi_init, int n, const double* hmat, const double* smat,
                      const double* beta, int n_proj, const double* q, const double* kin,
                      int n_conv, const double* conv_psi, const double* conv_spsi, int max_steps,
                      double tol, double* psi_out, double* out3, int* n_steps, int* converged,
                      int* step_type) {
  UsppPrecond pc;
  int rc = pc.init(C(beta), n, n_proj, C(q), kin);
  if (rc) return rc;
  const cd* H = C(hmat);
  const cd* S = C(smat);
  HsFn hs = [&](const cd* x, cd* hx, cd* sx) {
    for (int i = 0; i < n; ++i) {
      cd a = 0, b = 0;
      for (int j = 0; j < n; ++j) {
        a += H[(size_t)j * n + i] * x[j];
        b += S[(size_t)j * n + i] * x[j];
      }
      hx[i] = a;
      sx[i] = b;
    }
  };
  PrecondFn pre = [&](const cd* r, cd* o) { pc.apply(r, o); };
  CgResult res = band_cg_minimize(C(psi_init), n, pre, n_conv, C(conv_psi), C(conv_spsi),
                                  max_steps, tol, hs);
  if (res.status) return res.status;
  std::copy(res.psi.begin(), res.psi.end(), C(psi_out));
  out3[0] = res.eigenvalue;
  out3[1] = res.residual_norm;
  out3[2] = 0;
  *n_steps = res.n_steps;
  *converged = res.converged;
  *step_type = res.step_type;
  return 0;
}

}  // extern "C"
