// Device-side SCF helpers shared by scf.hip and scf_driver.hip.
#pragma once
#include <vector>

#include "impl.hpp"

namespace chs {

struct GridOps {
  chs_context* ctx = nullptr;
  int n = 0;
  size_t n3 = 0;
  double b = 0, omega = 0;
  DevBuf cgrid;  // complex scratch grid
  DevBuf ph;     // per-atom 1D phase tables [atom][axis][i] = e^{-i b f(i) x_a}
  int n_atoms = 0;
  int init(chs_context* c, int n, double L, double omega);
  int set_atoms(const double* pos, int n_atoms);
  int vion(double zval, double rc, double* d_vion);
  int vhxc(const double* rho, const double* vion, double* vhxc_out, double* veff,
           const double* vin, double* eh, double* exc, double* dc);
  int kerker(double* f, double q);
  ~GridOps() {
    release(cgrid);
    release(ph);
  }
};

struct PulayDev {
  chs_context* ctx = nullptr;
  GridOps* g = nullptr;
  int depth = 6;
  double alpha = 0.5, q = 0.8;
  DevBuf mem;              // (2*depth + 2) grids: rho_in/res pairs + scratch
  std::vector<int> slots;  // history, oldest first
  std::vector<double> gram;
  int init(chs_context* c, GridOps* g, int depth, double alpha, double q);
  int mix(const double* rho_in, const double* rho_out, double* rho_next);
  double* rin(int slot);
  double* res(int slot);
  ~PulayDev() { release(mem); }
};

double dot_real(chs_context* ctx, const double* a, const double* b, size_t n, int* rc);
int reduce2(chs_context* ctx, double* out2);
// integral-free L1 distance sum_i |a_i - b_i| (FP64, fixed reduction order)
double l1_diff(chs_context* ctx, const double* a, const double* b, size_t n, int* rc);
double fermi_occupations(const double* eps, int n, double n_e, double kt, double* occ, double* ts);
// Occupations of n states with weights w (k-point weights): sum w_i f_i = n_e
double fermi_occupations_w(const double* eps, const double* w, int n, double n_e, double kt,
                           double* occ, double* ts);
// rho(r) = sum_b occ_b |psi_b(r)|^2 / omega for a c128 band-major block
// (added to rho with `accumulate`)
int density_impl(chs_context* ctx, const z64* psi, int n_b, const double* occ_host, double omega,
                 double* d_rho, bool accumulate = false);
// the same from a complex64 block (FP64 accumulation; zero-window stick tables)
int density_c64(chs_context* ctx, const c32* psi, int n_b, const double* occ_host, double omega,
                double* d_rho, bool accumulate = false);

}  // namespace chs
