/* C ABI of the FP64 CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY: called from tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg, as the checker. The product library never
 * uses it. Complex arrays are interleaved (re, im) doubles; wavefunction
 * blocks are band-major (psi[b * n_pw + g]); grids are ix-fastest. */
#ifndef CHS_ORACLE_H
#define CHS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int32_t n_rep;
  double a0, zval, rc_loc, ecut;
  int32_t n_grid, n_bands;
  double kfrac[3];
  int32_t n_proj_per_atom;
  double beta_rc, q_rc, q0, d_min, d_max;
  uint64_t seed_d, seed_psi, seed_lanczos;
  int32_t cheb_degree, first_passes, lanczos_k, pulay_depth;
  double mix_alpha, kerker_q, etol, smear_kt;
  int32_t max_iter;
} orc_params;

void orc_default_params(orc_params* p);
/* OpenMP threads of the oracle's parallel loops (the CPU baseline's 1-thread
 * and all-core timings) */
void orc_set_num_threads(int n);
int orc_get_max_threads(void);

/* generator and setup */
double orc_lcg_uniforms(uint64_t seed, int n, double* out);
void orc_lcg_normals(uint64_t seed, int n, double* out);
void orc_synth_positions(int n_rep, double a0, double sigma, uint64_t seed, double* pos);
int orc_gsphere(double L, double ecut, const double kfrac[3], int n_grid, int max_pw,
                int32_t* miller, int32_t* fft_idx, double* kin);
void orc_pw_coords_to_fft_indices(const int32_t* miller, int n_pw, int ngx, int ngy, int ngz,
                                  int32_t* idx);
void orc_fft3d(double* grid, int n, int sign);
/* downsample_array_to_wave_grid (src/scf.rs:413-454), cubic grids */
void orc_downsample(const double* fine, int nf, int nw, double* out);

/* model: cell + plane waves + projectors + V_eff */
void* orc_model_create(const orc_params* p, const double* pos);
void orc_model_destroy(void* m);
int orc_model_n_pw(void* m);
int orc_model_n_proj(void* m);
void orc_model_get_pw(void* m, int32_t* miller, int32_t* fft_idx, double* kin);
void orc_model_get_vion(void* m, double* vion);
void orc_model_get_beta(void* m, double* beta);  /* P x n_pw complex */
void orc_model_get_d(void* m, double* d);
void orc_model_set_d(void* m, const double* d);
void orc_model_set_veff(void* m, const double* veff);
void orc_model_set_sinv_filter(void* m, int on);
void orc_initial_psi(void* m, int n_b, uint64_t seed, double* psi);
void orc_apply_h(void* m, const double* psi, int n_b, double* hpsi, double* spsi);
void orc_apply_sinv(void* m, double* x, int n_b);
double orc_grid_lambda_max(void* m);
double orc_lanczos_bound(void* m, int k, uint64_t seed, double* ritz_min);
void orc_spectral_bounds(const double* eig, int n_eig, double lambda_max, double* out4);
int orc_chebyshev_filter(void* m, const double* psi, int n_b, int degree, const double* bounds4,
                         double* out, double* diag4);
int orc_rayleigh_ritz(void* m, double* psi, int n_b, double* eps, double* hpsi_out);
void orc_density(void* m, const double* psi, int n_b, const double* occ, double* rho);
double orc_fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ,
                             double* ts);
void orc_vhxc(void* m, const double* rho, double* v, double* eh, double* exc);
void orc_screen_d(void* m, double* d);
void orc_lda_pz(int n, const double* rho, double* exc, double* vxc);

/* dense linear algebra (column-major complex) */
int orc_cholesky_lower(double* a, int n);
void orc_jacobi_eigh(double* a, int n, double* w, double* v);
void orc_subspace_matrix(const double* x, const double* y, int n_pw, int n_b, double* m);
int orc_gen_eigh(const double* h, const double* s, int n, double* eps, double* v);

/* full SCF: energies[max_iter], drho[max_iter], eps[n_bands], rho[n^3] */
int orc_scf_run(const orc_params* p, const double* pos, int* n_iter, int* converged,
                double* energies, double* drho, double* eps, double* rho);

/* the same with the final block (psi, n_bands x n_pw complex), occupations
 * (n_bands) and chemical potential (each may be NULL) */
int orc_scf_run_ex(const orc_params* p, const double* pos, int* n_iter, int* converged,
                   double* energies, double* drho, double* eps, double* rho, double* psi,
                   double* occ, double* mu);
/* k-point SCF over n_k k-points (kfrac 3 n_k, weights wk summing to 1):
 * eps[n_k * n_bands] k-major, n_pw[n_k] */
int orc_scf_run_k(const orc_params* p, const double* pos, int n_k, const double* kfrac,
                  const double* wk, int* n_iter, int* converged, double* energies, double* drho,
                  double* eps, double* rho, int* n_pw, double* mu);
double orc_fermi_occupations_w(const double* eps, const double* w, int n, double n_e, double kt,
                               double* occ, double* ts);
/* Pulay/Kerker density mixer on a cubic n^3 grid of edge L */
void* orc_pulay_create(int n_grid, double L, int depth, double alpha, double kerker_q);
void orc_pulay_mix(void* h, const double* rin, const double* rout, double* rnext);
void orc_pulay_destroy(void* h);

/* band-CG stack (src/eigensolver/{line_search,cg_helpers,uspp_preconditioner,band_cg}.rs) */
void orc_line_search(double a, double hd, double dd, double sd, double* out3, int* status);
int orc_compute_residual(const double* psi, const double* hpsi, const double* spsi, int n,
                         double* r, double* eps);
double orc_residual_bare_norm(const double* r, int n);
int orc_s_normalize(const double* psi, const double* spsi, int n, double* out);
int orc_s_orthogonalize_against(const double* psi, const double* spsi, int n, int n_conv,
                                const double* conv_psi, const double* conv_spsi, double* out_psi,
                                double* out_spsi);
void orc_tpa_diag(const double* kin, int n, double* t);
int orc_uspp_precond_apply(const double* beta, int n_pw, int n_proj, const double* q,
                           const double* kin, const double* res, int n_vec, double* out,
                           double* r_matrix, double* c_matrix);
/* band CG of every band of a block against the lower bands of the block as
 * handed in, with the model's H/S; precond 0 = TPA, 1 = TPA + USPP term */
int orc_band_cg_model(void* m, double* psi_io, int n_b, int max_steps, double tol, int precond,
                      double* eps, int* n_steps, int* converged);
/* band CG with dense H and S (n x n column-major complex) and the
 * TPA/USPP preconditioner built from beta (n x n_proj row-major), q, kin. */
int orc_band_cg_dense(const double* psi_init, int n, const double* hmat, const double* smat,
                      const double* beta, int n_proj, const double* q, const double* kin,
                      int n_conv, const double* conv_psi, const double* conv_spsi, int max_steps,
                      double tol, double* psi_out, double* out3, int* n_steps, int* converged,
                      int* step_type);

#ifdef __cplusplus
}
#endif
#endif
