// Synthetic cell, plane-wave sphere, index map, ionic local potential and
// USPP projectors for the FP64 oracle. Test infrastructure only (see orc.hpp).
#include "orc.hpp"

#include <cmath>
#include <stdexcept>

namespace orc {

double Lcg::normal() {
  double u1 = 1.0 - uniform();
  double u2 = uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// Diamond-cubic basis in units of a0.
static const double kBasis[8][3] = {
    {0.00, 0.00, 0.00}, {0.00, 0.50, 0.50}, {0.50, 0.00, 0.50}, {0.50, 0.50, 0.00},
    {0.25, 0.25, 0.25}, {0.25, 0.75, 0.75}, {0.75, 0.25, 0.75}, {0.75, 0.75, 0.25}};

// Atoms are emitted replica-major (ix, iy, iz), then basis order; each
// Cartesian component is displaced by sigma * N(0,1) drawn in that order.
void synth_positions(int n_rep, double a0, double sigma, uint64_t seed, double* pos) {
  Lcg rng(seed);
  int a = 0;
  for (int ix = 0; ix < n_rep; ++ix)
    for (int iy = 0; iy < n_rep; ++iy)
      for (int iz = 0; iz < n_rep; ++iz)
        for (int b = 0; b < 8; ++b) {
          const double cell[3] = {(double)ix, (double)iy, (double)iz};
          for (int c = 0; c < 3; ++c) {
            double d = sigma * rng.normal();
            pos[3 * a + c] = a0 * (cell[c] + kBasis[b][c]) + d;
          }
          ++a;
        }
}

// Reference: src/scf.rs:387-401 (pw_coords_to_fft_indices): negative
// Miller indices wrap by +n; linear index ix + ngx*(iy + ngy*iz).
void pw_coords_to_fft_indices(const int32_t* miller, int n_pw, int ngx, int ngy,
                              int ngz, int32_t* idx) {
  for (int g = 0; g < n_pw; ++g) {
    int h = miller[3 * g], k = miller[3 * g + 1], l = miller[3 * g + 2];
    int ix = h >= 0 ? h : h + ngx;
    int iy = k >= 0 ? k : k + ngy;
    int iz = l >= 0 ? l : l + ngz;
    idx[g] = ix + ngx * (iy + ngy * iz);
  }
}

// G-sphere 0.5|k+G|^2 <= ecut, G = (2 pi / L) (h, k, l), lexicographic
// (h slowest). The kinetic energy per plane wave is the per-PW form of
// tests/phase0_gate1_consistency.rs:187-197 (the reference's GPU path indexes
// a grid-ordered array by PW index instead; that defect is not reproduced).
// The arithmetic order is fixed and unfused so that the sphere test is
// reproducible bit for bit on any host or device.
int gsphere(double L, double ecut, const double kfrac[3], int n_grid, PwSet& out) {
  const double b = (2.0 * M_PI) / L;
  const int hmax = (int)std::ceil(std::sqrt(2.0 * ecut) / b) + 1;
  out.miller.clear();
  out.kin.clear();
  out.gcart.clear();
  for (int h = -hmax; h <= hmax; ++h)
    for (int k = -hmax; k <= hmax; ++k)
      for (int l = -hmax; l <= hmax; ++l) {
        // (gx*gx + gy*gy) + gz*gz, each product rounded separately; the
        // library is built with -ffp-contract=off.
        double gx = b * ((double)h + kfrac[0]);
        double gy = b * ((double)k + kfrac[1]);
        double gz = b * ((double)l + kfrac[2]);
        double g2 = (gx * gx + gy * gy) + gz * gz;
        double e = 0.5 * g2;
        if (e <= ecut) {
          if (2 * std::abs(h) >= n_grid || 2 * std::abs(k) >= n_grid ||
              2 * std::abs(l) >= n_grid)
            return -1;
          out.miller.push_back(h);
          out.miller.push_back(k);
          out.miller.push_back(l);
          out.kin.push_back(e);
          out.gcart.push_back(gx);
          out.gcart.push_back(gy);
          out.gcart.push_back(gz);
        }
      }
  out.n_pw = (int)out.kin.size();
  out.fft_idx.resize(out.n_pw);
  pw_coords_to_fft_indices(out.miller.data(), out.n_pw, n_grid, n_grid, n_grid,
                           out.fft_idx.data());
  return out.n_pw;
}

// Ionic local potential on the real-space grid.
// v(G) per atom for V(r) = -Z erf(r/rc)/r:  -4 pi Z exp(-G^2 rc^2/4) / (Omega G^2),
// and at G = 0 the non-Coulomb limit  pi Z rc^2 / Omega.  Nyquist planes are
// zero. V(r) = sum_G V(G) e^{iG.r} (unnormalised inverse FFT, real part).
void vion_real(const System& sys, std::vector<double>& vr) {
  const int n = sys.n_grid;
  const size_t n3 = (size_t)n * n * n;
  const double b = (2.0 * M_PI) / sys.L;
  const double omega = sys.omega;
  const double z = sys.zval, rc = sys.rc_loc;
  // Per-atom 1D phase tables e^{-i b f x_a} for every axis frequency f.
  std::vector<cd> ph((size_t)sys.n_atoms * 3 * n);
  for (int a = 0; a < sys.n_atoms; ++a)
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < n; ++i) {
        bool nyq;
        int f = freq_of(i, n, &nyq);
        double t = -b * f * sys.pos[3 * a + c];
        ph[((size_t)a * 3 + c) * n + i] = cd(std::cos(t), std::sin(t));
      }
  std::vector<cd> grid(n3);
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        bool nx, ny, nz;
        int h = freq_of(ix, n, &nx), k = freq_of(iy, n, &ny), l = freq_of(iz, n, &nz);
        size_t id = (size_t)ix + (size_t)n * (iy + (size_t)n * iz);
        if (nx || ny || nz) {
          grid[id] = 0;
          continue;
        }
        double g2 = b * b * (double)(h * h + k * k + l * l);
        double v;
        if (h == 0 && k == 0 && l == 0)
          v = M_PI * z * rc * rc / omega;
        else
          v = -4.0 * M_PI * z * std::exp(-g2 * rc * rc / 4.0) / (omega * g2);
        cd s = 0;
        for (int a = 0; a < sys.n_atoms; ++a)
          s += ph[((size_t)a * 3 + 0) * n + ix] * ph[((size_t)a * 3 + 1) * n + iy] *
               ph[((size_t)a * 3 + 2) * n + iz];
        grid[id] = v * s;
      }
  fft3d(grid.data(), n, +1);
  vr.resize(n3);
  for (size_t i = 0; i < n3; ++i) vr[i] = grid[i].real();
}

// Projectors beta_{I,p}(k+G) = Omega^{-1/2} f_p(k+G) e^{-i(k+G).R_I} for
// p = (l=0), (l=1, x), (l=1, y), (l=1, z), from the Gaussian Fourier
// transforms of r^l exp(-r^2/s^2) Y_lm (real harmonics):
//   f_0 = Y00 pi^{3/2} s^3 exp(-G^2 s^2/4)
//   f_a = sqrt(3/4pi) (-i s^2 G_a / 2) pi^{3/2} s^3 exp(-G^2 s^2/4).
// Storage is projector-major (beta[p * n_pw + g]), the reference's per-ion
// (n_exp x n_pw) row-major block (src/eigensolver/vnl_data.rs:13-21, :79-80)
// concatenated over ions.
void build_projectors(const System& sys, const PwSet& pw, Projectors& pj) {
  const int np = sys.n_proj_per_atom;
  const int P = np * sys.n_atoms;
  pj.n_proj = P;
  pj.beta.assign((size_t)P * pw.n_pw, cd(0));
  pj.d.assign(P, 0.0);
  pj.q.assign(P, sys.q0);
  if (P == 0) return;
  const double s = sys.beta_rc;
  const double pref = std::pow(M_PI, 1.5) * s * s * s / std::sqrt(sys.omega);
  const double y00 = 1.0 / std::sqrt(4.0 * M_PI);
  const double y1 = std::sqrt(3.0 / (4.0 * M_PI));
  for (int a = 0; a < sys.n_atoms; ++a) {
    for (int p = 0; p < np; ++p) pj.d[a * np + p] = sys.dion[a * np + p];
#pragma omp parallel for schedule(static)
    for (int g = 0; g < pw.n_pw; ++g) {
      const double gx = pw.gcart[3 * g], gy = pw.gcart[3 * g + 1], gz = pw.gcart[3 * g + 2];
      const double g2 = gx * gx + gy * gy + gz * gz;
      const double rad = pref * std::exp(-g2 * s * s / 4.0);
      const double t = -(gx * sys.pos[3 * a] + gy * sys.pos[3 * a + 1] + gz * sys.pos[3 * a + 2]);
      const cd phase(std::cos(t), std::sin(t));
      const cd mi(0.0, -0.5 * s * s);
      pj.beta[(size_t)(a * np + 0) * pw.n_pw + g] = y00 * rad * phase;
      if (np > 1) {
        pj.beta[(size_t)(a * np + 1) * pw.n_pw + g] = y1 * rad * mi * gx * phase;
        pj.beta[(size_t)(a * np + 2) * pw.n_pw + g] = y1 * rad * mi * gy * phase;
        pj.beta[(size_t)(a * np + 3) * pw.n_pw + g] = y1 * rad * mi * gz * phase;
      }
    }
  }
}

}  // namespace orc
