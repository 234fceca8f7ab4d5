// Oracle eigensolver and SCF declarations. Test infrastructure only.
#pragma once
#include "orc.hpp"

namespace orc {

// One Hamiltonian instance: cell, plane waves, V_eff on the grid, projectors.
struct Model {
  System sys;
  PwSet pw;
  Projectors pj;
  std::vector<double> veff;   // n^3
  bool sinv_filter = true;    // filter S^{-1}H for the generalized problem
  void apply_a(const cd* x, int n_b, cd* ax, cd* sx_opt) const;
};

struct Bounds {
  double lambda_max, eps_cut, center, half_width;
};

Bounds spectral_bounds(const double* eig, int n_eig, double lambda_max);
double grid_lambda_max(const Model& m);
double lanczos_bound(const Model& m, int k, uint64_t seed, double* ritz_min);
int chebyshev_filter(const Model& m, const cd* x, int n_b, int degree, const Bounds& bd,
                     std::vector<cd>& y, double* diag);
int rayleigh_ritz(const Model& m, std::vector<cd>& y, int n_b, std::vector<double>& eps,
                  std::vector<cd>* hy_out);

void subspace_matrix(const cd* x, const cd* y, int n_pw, int n_b, std::vector<cd>& m);
int gen_eigh(const std::vector<cd>& h, const std::vector<cd>& s, int n, std::vector<double>& eps,
             std::vector<cd>& vmat);
void rotate(const cd* x, const std::vector<cd>& v, int n_pw, int n_b, cd* out);

}  // namespace orc
