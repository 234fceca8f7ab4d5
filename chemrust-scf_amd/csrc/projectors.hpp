// USPP projector setup and SCF glue (projectors.hip).
#pragma once
#include "impl.hpp"

namespace chs {

// beta (P x n_pw, projector-major) for n_atoms atoms with np projectors each.
int build_beta(chs_context* ctx, const int32_t* miller, int n_pw, double L, const double* kf,
               const double* d_pos, int n_atoms, int np, double s, z64* beta);
// rho += augmentation charge of the bound projectors for a c128 block psi;
// w (optional): beta^H psi already at hand (P x n_b).
int augment_density(chs_context* ctx, const z64* psi, int n_b, const double* d_occ, int n_atoms,
                    int np, double q0, double rq, const z64* d_ph, int n, double L, z64* cgrid,
                    double* rho, const z64* w = nullptr);
// d_out = D0 + q0 * integral of V_eff against the Gaussian augmentation.
int screen_dion(chs_context* ctx, const double* veff, int n, double L, int n_atoms, int np,
                double q0, double rq, const z64* d_ph, const double* d_d0, z64* cgrid,
                double* d_out);

}  // namespace chs
