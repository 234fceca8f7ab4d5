// Oracle SCF declarations. Test infrastructure only.
#pragma once
#include "eigensolver.hpp"

namespace orc {

struct ScfParams {
  int n_rep = 1;
  double a0 = 10.26;
  double zval = 4.0, rc_loc = 1.0;
  double ecut = 15.0;
  int n_grid = 36;
  int n_bands = 20;
  double kfrac[3] = {0, 0, 0};
  int n_proj_per_atom = 0;
  double beta_rc = 1.3, q_rc = 1.2, q0 = 0.0, d_min = 0.5, d_max = 2.0;
  uint64_t seed_d = 3, seed_psi = 2, seed_lanczos = 4;
  int cheb_degree = 10;
  int first_passes = 4;
  int lanczos_k = 10;      // 0: reference grid-corner lambda_max
  int pulay_depth = 6;
  double mix_alpha = 0.5, kerker_q = 0.8;
  double smear_kt = 0.01;  // Fermi-Dirac kT (Ha); 0 = integer filling
  double etol = 1e-6;
  int max_iter = 60;
};

struct ScfResult {
  int n_iter = 0, converged = 0, n_pw = 0;
  double mu = 0;
  std::vector<double> energies, drho, eps, occ, rho;
  std::vector<cd> psi;
};

// k-point SCF result: eigenvalues k-major (eps[k * n_bands + i]).
struct ScfResultK {
  int n_iter = 0, converged = 0;
  double mu = 0;
  std::vector<int> n_pw;
  std::vector<double> energies, drho, eps, rho;
};

struct Pulay {
  int depth = 6;
  double alpha = 0.5, kerker_q = 0.8;
  std::vector<std::vector<double>> hin, hres;
  void mix(const System& sys, const std::vector<double>& rin, const std::vector<double>& rout,
           std::vector<double>& rnext);
};

bool solve_dense(std::vector<double>& a, std::vector<double>& b, int n);
void lda_pz(double rho, double* exc, double* vxc);
double hartree(const System& sys, const std::vector<double>& rho, std::vector<double>& vh);
void vhxc(const System& sys, const std::vector<double>& rho, std::vector<double>& v, double* eh,
          double* exc);
void density(const Model& m, const cd* psi, int n_b, const double* occ, std::vector<double>& rho);
double fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ, double* ts);
double fermi_occupations_w(const double* eps, const double* w, int n, double n_e, double kt,
                           double* occ, double* ts);
void screen_d(const Model& m, std::vector<double>& d);
void initial_psi(const PwSet& pw, int n_b, uint64_t seed, std::vector<cd>& psi);
void setup_model(const ScfParams& p, const double* pos, Model& m);
int run_scf(const ScfParams& p, const double* pos, ScfResult& res, Model* model_out);
int run_scf_k(const ScfParams& p, const double* pos, int n_k, const double* kfrac,
              const double* wk, ScfResultK& res);

}  // namespace orc
