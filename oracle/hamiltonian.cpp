// H|psi>, S|psi> and S^{-1} for the FP64 oracle. Test infrastructure only.
#include "orc.hpp"

#include <algorithm>
#include <cmath>

#include <omp.h>

namespace orc {

// Local part, reference src/eigensolver/chebyshev.rs:279-365
// (apply_v_loc_hamiltonian): T|psi> with T = 0.5|k+G|^2 per plane wave
// (per-PW form, tests/phase0_gate1_consistency.rs:187-197), then scatter to
// the grid, unnormalised inverse FFT, multiply by V_eff(r), unnormalised
// forward FFT and gather with 1/N^3 (chebyshev.rs:126-142).
// Nonlocal part, chebyshev.rs:416-500 and the operator form of
// tests/phase0_gate1_consistency.rs:320-385:  H += sum beta D beta^H,
// S = 1 + sum beta Q beta^H, with D and Q diagonal per projector.
void apply_h(const PwSet& pw, int n, const double* veff, const Projectors* pj,
             const cd* psi, int n_b, cd* hpsi, cd* spsi) {
  const int n_pw = pw.n_pw;
  const size_t n3 = (size_t)n * n * n;
  const double inv = 1.0 / (double)n3;
  auto one_band = [&](int b, std::vector<cd>& grid) {
    const cd* p = psi + (size_t)b * n_pw;
    cd* hp = hpsi + (size_t)b * n_pw;
    std::fill(grid.begin(), grid.end(), cd(0));
    for (int g = 0; g < n_pw; ++g) grid[pw.fft_idx[g]] = p[g];
    fft3d(grid.data(), n, +1);
    for (size_t r = 0; r < n3; ++r) grid[r] *= veff[r];
    fft3d(grid.data(), n, -1);
    for (int g = 0; g < n_pw; ++g) hp[g] = pw.kin[g] * p[g] + grid[pw.fft_idx[g]] * inv;
  };
  if (n_b < omp_get_max_threads()) {
    // few bands: parallelise inside each FFT instead
    std::vector<cd> grid(n3);
    for (int b = 0; b < n_b; ++b) one_band(b, grid);
  } else {
#pragma omp parallel
    {
      std::vector<cd> grid(n3);
#pragma omp for schedule(dynamic, 1)
      for (int b = 0; b < n_b; ++b) one_band(b, grid);
    }
  }
  const int P = pj ? pj->n_proj : 0;
  if (spsi) std::copy(psi, psi + (size_t)n_b * n_pw, spsi);
  if (P == 0) return;
  // C = beta^H psi  (P x n_b)
  std::vector<cd> c((size_t)P * n_b);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < n_b; ++b)
    for (int p = 0; p < P; ++p) {
      const cd* bp = pj->beta.data() + (size_t)p * n_pw;
      const cd* x = psi + (size_t)b * n_pw;
      cd s = 0;
      for (int g = 0; g < n_pw; ++g) s += std::conj(bp[g]) * x[g];
      c[(size_t)b * P + p] = s;
    }
#pragma omp parallel for schedule(static)
  for (int b = 0; b < n_b; ++b) {
    cd* hp = hpsi + (size_t)b * n_pw;
    cd* sp = spsi ? spsi + (size_t)b * n_pw : nullptr;
    for (int p = 0; p < P; ++p) {
      const cd* bp = pj->beta.data() + (size_t)p * n_pw;
      const cd cd_ = pj->d[p] * c[(size_t)b * P + p];
      const cd cq = pj->q[p] * c[(size_t)b * P + p];
      for (int g = 0; g < n_pw; ++g) hp[g] += bp[g] * cd_;
      if (sp)
        for (int g = 0; g < n_pw; ++g) sp[g] += bp[g] * cq;
    }
  }
}

// S^{-1} = 1 - beta M beta^H with M = (Q^{-1} + beta^H beta)^{-1} (Woodbury).
// The reference applies the same identity inside its preconditioner
// (src/eigensolver/uspp_preconditioner.rs:80-131 with T = 1).
void build_sinv(Projectors& pj, int n_pw) {
  const int P = pj.n_proj;
  pj.sinv_m.assign((size_t)P * P, cd(0));
  if (P == 0) return;
  // Q = 0: S = 1, and M = (Q^{-1} + beta^H beta)^{-1} -> 0 in the limit
  bool any_q = false;
  for (int i = 0; i < P; ++i) any_q = any_q || pj.q[i] != 0.0;
  if (!any_q) return;
  std::vector<cd> a((size_t)P * P);
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < P; ++j)
    for (int i = 0; i < P; ++i) {
      const cd* bi = pj.beta.data() + (size_t)i * n_pw;
      const cd* bj = pj.beta.data() + (size_t)j * n_pw;
      cd s = 0;
      for (int g = 0; g < n_pw; ++g) s += std::conj(bi[g]) * bj[g];
      a[(size_t)j * P + i] = s;
    }
  for (int i = 0; i < P; ++i) a[(size_t)i * P + i] += 1.0 / pj.q[i];
  // Hermitian positive definite: invert through Cholesky.
  std::vector<cd> l = a;
  if (cholesky_lower(l, P) != 0) return;
  // Linv by forward substitution, then M = Linv^H Linv.
  std::vector<cd> li((size_t)P * P, cd(0));
  for (int col = 0; col < P; ++col) {
    li[(size_t)col * P + col] = 1.0 / l[(size_t)col * P + col];
    for (int i = col + 1; i < P; ++i) {
      cd s = 0;
      for (int k = col; k < i; ++k) s += l[(size_t)k * P + i] * li[(size_t)col * P + k];
      li[(size_t)col * P + i] = -s / l[(size_t)i * P + i];
    }
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < P; ++j)
    for (int i = 0; i < P; ++i) {
      cd s = 0;
      for (int k = std::max(i, j); k < P; ++k)
        s += std::conj(li[(size_t)i * P + k]) * li[(size_t)j * P + k];
      pj.sinv_m[(size_t)j * P + i] = s;
    }
}

void apply_sinv(const Projectors& pj, int n_pw, cd* x, int n_b) {
  const int P = pj.n_proj;
  if (P == 0) return;
#pragma omp parallel
  {
    std::vector<cd> c(P), mc(P);
#pragma omp for schedule(static)
    for (int b = 0; b < n_b; ++b) {
      cd* xb = x + (size_t)b * n_pw;
      for (int p = 0; p < P; ++p) {
        const cd* bp = pj.beta.data() + (size_t)p * n_pw;
        cd s = 0;
        for (int g = 0; g < n_pw; ++g) s += std::conj(bp[g]) * xb[g];
        c[p] = s;
      }
      for (int i = 0; i < P; ++i) {
        cd s = 0;
        for (int j = 0; j < P; ++j) s += pj.sinv_m[(size_t)j * P + i] * c[j];
        mc[i] = s;
      }
      for (int p = 0; p < P; ++p) {
        const cd* bp = pj.beta.data() + (size_t)p * n_pw;
        for (int g = 0; g < n_pw; ++g) xb[g] -= bp[g] * mc[p];
      }
    }
  }
}

}  // namespace orc
