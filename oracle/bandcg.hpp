// Oracle band-CG declarations. Test infrastructure only.
#pragma once
#include <functional>

#include "orc.hpp"

namespace orc {

struct LineSearchResult {
  double step_size, eigenvalue, norm;
  int status;
};

LineSearchResult line_search_2d_quadratic(double dot_hpsi_psi, double dot_hpsi_dir,
                                          double dot_hdir_dir, double dot_sdir_dir);
int compute_residual(const cd* psi, const cd* hpsi, const cd* spsi, int n, cd* r, double* eps);
double residual_bare_norm(const cd* r, int n);
int s_normalize(const cd* psi, const cd* spsi, int n, cd* out);
// full_overlap: project out the complex overlap instead of its real part
// (an option of this build; the reference keeps the real part)
int s_orthogonalize_against(const cd* psi, const cd* spsi, int n, int n_conv, const cd* conv_psi,
                            const cd* conv_spsi, cd* out_psi, cd* out_spsi,
                            bool full_overlap = false);
void tpa_diag(const double* kin, int n, double* t);

struct UsppPrecond {
  int n_pw = 0, n_proj = 0;
  std::vector<double> tpa;
  std::vector<cd> beta;  // n_pw x n_proj row-major
  std::vector<cd> c, r;  // n_proj^2 column-major
  int init(const cd* beta_g, int n_pw, int n_proj, const cd* q_matrix, const double* kin);
  void apply(const cd* res, cd* out) const;
};

struct CgResult {
  std::vector<cd> psi;
  double eigenvalue = 0;
  int n_steps = 0, converged = 0;
  double residual_norm = 0;
  int step_type = 0;  // 0 = SD, 1 = CG
  int status = 0;     // 0 ok, <0 contract violation (the reference panics)
};

using HsFn = std::function<void(const cd*, cd*, cd*)>;
using PrecondFn = std::function<void(const cd*, cd*)>;

CgResult band_cg_minimize(const cd* psi_init, int n, const PrecondFn& pre, int n_conv,
                          const cd* conv_psi, const cd* conv_spsi, int max_steps, double tol,
                          const HsFn& apply_hs, bool full_overlap = false);

}  // namespace orc
