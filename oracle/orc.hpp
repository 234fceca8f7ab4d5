// FP64 CPU oracle for the chemrust-scf SCF eigensolver hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by the
// pytest suite, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg,
// and by nothing else. The product (libchs_gpu.so) never links, loads or
// calls it, and has no CPU fallback.
//
// It restates, in plain FP64 C++, the algorithms of the reference
// (TonyWu20/chemrust-scf) on the path from ScfIteration::diagonalize down to
// the device layer, with the defects the reference itself documents fixed
// (kinetic energy per plane wave, Rayleigh-Ritz operand layout). Every
// function cites the reference file:line it follows.
//
// Complex numbers are interleaved (re, im) doubles, as in the reference's
// CudaComplex (src/device/mod.rs:24-43). Wavefunction blocks are band-major
// (psi[b * n_pw + g]), the reference's ColumnDistributed layout
// (src/layout.rs:48-53). FFT grids are ix-fastest: idx = ix + n*(iy + n*iz)
// (src/scf.rs:385-399).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>
#include <string>

namespace orc {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Seeded generator: the reference's 48-bit LCG
// (src/eigensolver/subspace_diag.rs:434-447) plus Box-Muller.
// ---------------------------------------------------------------------------
struct Lcg {
  uint64_t s;
  explicit Lcg(uint64_t seed) : s(seed) {}
  double uniform() {
    s = (s * 25214903917ull + 11ull) & 0xFFFFFFFFFFFFull;
    return (double)s / 281474976710656.0;
  }
  // One standard normal from two uniforms: sqrt(-2 ln(1-u1)) cos(2 pi u2).
  double normal();
};

// ---------------------------------------------------------------------------
// Synthetic system description. Both this oracle and the GPU library build
// the same cell from the same numbers.
// ---------------------------------------------------------------------------
struct System {
  int n_rep = 1;          // supercell repetitions per axis
  double a0 = 10.26;      // diamond-cubic lattice constant (bohr)
  double L = 10.26;       // cubic supercell edge (bohr)
  double omega = 0;       // cell volume
  int n_atoms = 0;
  std::vector<double> pos;  // 3 * n_atoms, Cartesian bohr
  double zval = 4.0;      // ionic charge per atom
  double rc_loc = 1.0;    // V_loc(r) = -zval * erf(r / rc_loc) / r
  int n_grid = 36;
  // USPP part (n_proj_per_atom == 0 means local potential only).
  int n_proj_per_atom = 0;  // 4: l=0 (1) + l=1 (3)
  double beta_rc = 1.3;     // beta_l(r) = r^l exp(-r^2 / beta_rc^2)
  double q_rc = 1.2;        // Gaussian augmentation radius
  double q0 = 0.0;          // Q_ij = delta_ij * q0 (integrated charge)
  std::vector<double> dion; // n_atoms * n_proj_per_atom, diagonal D0 (Ha)
};

void synth_positions(int n_rep, double a0, double sigma, uint64_t seed, double* pos);

// Plane-wave sphere at one k-point: 0.5 |k+G|^2 <= ecut, Miller triples in
// lexicographic (h, k, l) order, h slowest.
struct PwSet {
  int n_pw = 0;
  std::vector<int32_t> miller;  // 3 * n_pw
  std::vector<int32_t> fft_idx; // n_pw
  std::vector<double> kin;      // 0.5 |k+G|^2
  std::vector<double> gcart;    // 3 * n_pw, k+G Cartesian
};
int gsphere(double L, double ecut, const double kfrac[3], int n_grid, PwSet& out);
void pw_coords_to_fft_indices(const int32_t* miller, int n_pw, int ngx, int ngy,
                              int ngz, int32_t* idx);

// ---------------------------------------------------------------------------
// FFT: unnormalised complex transforms on an n^3 grid, ix fastest.
// sign = +1: inverse (e^{+iGr}), sign = -1: forward (e^{-iGr}).
// ---------------------------------------------------------------------------
void fft1d(cd* x, int n, int stride, int sign);
void fft3d(cd* grid, int n, int sign);
// fine (nf^3) -> wave (nw^3) real array by G-space truncation (scf.rs:413-454)
void downsample(const double* fine, int nf, int nw, double* out);

// Signed frequency of FFT index i on an n-point axis; the Nyquist index
// (i == n/2 for even n) is reported through *nyq.
inline int freq_of(int i, int n, bool* nyq) {
  *nyq = (n % 2 == 0) && (i == n / 2);
  return (i <= n / 2) ? i : i - n;
}

// ---------------------------------------------------------------------------
// Hamiltonian pieces
// ---------------------------------------------------------------------------
struct Projectors {
  int n_proj = 0;            // total projectors P
  std::vector<cd> beta;      // P * n_pw, projector-major: beta[p*n_pw + g]
  std::vector<double> d;     // P, diagonal D (screened or bare)
  std::vector<double> q;     // P, diagonal Q
  std::vector<cd> sinv_m;    // P*P: M = (Q^-1 + B)^-1, B = beta^H beta (S^-1 Woodbury)
};

void vion_real(const System& sys, std::vector<double>& vr);
void build_projectors(const System& sys, const PwSet& pw, Projectors& pj);
void build_sinv(Projectors& pj, int n_pw);

// H|psi> for n_b bands. With `spsi` non-null, also S|psi>.
void apply_h(const PwSet& pw, int n, const double* veff, const Projectors* pj,
             const cd* psi, int n_b, cd* hpsi, cd* spsi);
// S^{-1} x in place via the Woodbury identity.
void apply_sinv(const Projectors& pj, int n_pw, cd* x, int n_b);

// Dense Hermitian helpers (column-major n x n).
int cholesky_lower(std::vector<cd>& a, int n);  // returns 0 or failing column+1
void jacobi_eigh(std::vector<cd>& a, int n, std::vector<double>& w,
                 std::vector<cd>& v);

}  // namespace orc
