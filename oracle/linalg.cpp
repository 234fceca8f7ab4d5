// Dense FP64 linear algebra and Rayleigh-Ritz for the oracle.
// Test infrastructure only (see orc.hpp).
#include "orc.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace orc {

// In-place lower Cholesky A = L L^H of a Hermitian matrix, column-major.
// Only the lower triangle is read; the strict upper triangle is zeroed.
// Returns 0, or j+1 when the pivot of column j is not positive (the
// reference panics there: src/eigensolver/subspace_diag.rs:291).
int cholesky_lower(std::vector<cd>& a, int n) {
  for (int j = 0; j < n; ++j) {
    double d = a[(size_t)j * n + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(a[(size_t)k * n + j]);
    if (!(d > 0.0)) return j + 1;
    const double ljj = std::sqrt(d);
    a[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      cd s = a[(size_t)j * n + i];
      for (int k = 0; k < j; ++k) s -= a[(size_t)k * n + i] * std::conj(a[(size_t)k * n + j]);
      a[(size_t)j * n + i] = s / ljj;
    }
    for (int i = 0; i < j; ++i) a[(size_t)j * n + i] = 0;
  }
  return 0;
}

// Cyclic complex Jacobi eigensolver for a Hermitian matrix (column-major).
// Rotation for pair (p, q) with a_pq = |a_pq| e: U = [[c, s e], [-s e*, c]],
// t = tan(theta) from tau = (a_qq - a_pp) / (2|a_pq|). Eigenvalues ascending,
// eigenvectors in the columns of v (the ordering of the reference's
// self_adjoint_evd, src/eigensolver/subspace_diag.rs:332-363).
void jacobi_eigh(std::vector<cd>& a, int n, std::vector<double>& w, std::vector<cd>& v) {
  v.assign((size_t)n * n, cd(0));
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  auto A = [&](int i, int j) -> cd& { return a[(size_t)j * n + i]; };
  double scale = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) scale += std::norm(A(i, j));
  scale = std::sqrt(scale);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < j; ++i) off += std::norm(A(i, j));
    if (std::sqrt(2.0 * off) <= 1e-15 * scale || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cd apq = A(p, q);
        const double mag = std::abs(apq);
        if (mag == 0.0 || mag < 1e-300) continue;
        const cd e = apq / mag;
        const double app = A(p, p).real(), aqq = A(q, q).real();
        const double tau = (aqq - app) / (2.0 * mag);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        // A <- A U (columns p, q)
        for (int k = 0; k < n; ++k) {
          const cd akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * std::conj(e) * akq;
          A(k, q) = s * e * akp + c * akq;
        }
        // A <- U^H A (rows p, q)
        for (int k = 0; k < n; ++k) {
          const cd apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * e * aqk;
          A(q, k) = s * std::conj(e) * apk + c * aqk;
        }
        A(p, q) = 0;
        A(q, p) = 0;
        A(p, p) = A(p, p).real();
        A(q, q) = A(q, q).real();
        for (int k = 0; k < n; ++k) {
          const cd vkp = v[(size_t)p * n + k], vkq = v[(size_t)q * n + k];
          v[(size_t)p * n + k] = c * vkp - s * std::conj(e) * vkq;
          v[(size_t)q * n + k] = s * e * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return A(x, x).real() < A(y, y).real(); });
  w.resize(n);
  std::vector<cd> vs((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    w[k] = A(ord[k], ord[k]).real();
    std::copy(v.begin() + (size_t)ord[k] * n, v.begin() + (size_t)(ord[k] + 1) * n,
              vs.begin() + (size_t)k * n);
  }
  v.swap(vs);
}

// Subspace matrix X^H Y over band-major blocks, Hermitised as (M + M^H)/2
// (src/eigensolver/subspace_diag.rs:71-130 build_h_sub / :146-196 build_s_sub).
void subspace_matrix(const cd* x, const cd* y, int n_pw, int n_b, std::vector<cd>& m) {
  m.assign((size_t)n_b * n_b, cd(0));
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int j = 0; j < n_b; ++j)
    for (int i = 0; i < n_b; ++i) {
      if (i > j) continue;
      const cd* xi = x + (size_t)i * n_pw;
      const cd* yj = y + (size_t)j * n_pw;
      cd s = 0;
      for (int g = 0; g < n_pw; ++g) s += std::conj(xi[g]) * yj[g];
      m[(size_t)j * n_b + i] = s;
    }
  for (int j = 0; j < n_b; ++j)
    for (int i = j + 1; i < n_b; ++i) m[(size_t)j * n_b + i] = std::conj(m[(size_t)i * n_b + j]);
  for (int i = 0; i < n_b; ++i) m[(size_t)i * n_b + i] = m[(size_t)i * n_b + i].real();
}

// Cholesky-reduction generalized eigensolve H v = eps S v
// (src/eigensolver/subspace_diag.rs:229-391): S = L L^H, Linv = L^{-1},
// B = Linv H Linv^H (Hermitised), B U = U diag(eps) ascending,
// V = Linv^H U. Returns 0 or the failing Cholesky column + 1.
int gen_eigh(const std::vector<cd>& h, const std::vector<cd>& s, int n, std::vector<double>& eps,
             std::vector<cd>& vmat) {
  std::vector<cd> l = s;
  int info = cholesky_lower(l, n);
  if (info) return info;
  std::vector<cd> li((size_t)n * n, cd(0));
  for (int col = 0; col < n; ++col) {
    li[(size_t)col * n + col] = 1.0 / l[(size_t)col * n + col];
    for (int i = col + 1; i < n; ++i) {
      cd acc = 0;
      for (int k = col; k < i; ++k) acc += l[(size_t)k * n + i] * li[(size_t)col * n + k];
      li[(size_t)col * n + i] = -acc / l[(size_t)i * n + i];
    }
  }
  // T = Linv H ; B = T Linv^H
  std::vector<cd> t((size_t)n * n, cd(0)), b((size_t)n * n, cd(0));
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd acc = 0;
      for (int k = 0; k <= i; ++k) acc += li[(size_t)k * n + i] * h[(size_t)j * n + k];
      t[(size_t)j * n + i] = acc;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd acc = 0;
      for (int k = 0; k <= j; ++k) acc += t[(size_t)k * n + i] * std::conj(li[(size_t)k * n + j]);
      b[(size_t)j * n + i] = acc;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) {
      cd v = 0.5 * (b[(size_t)j * n + i] + std::conj(b[(size_t)i * n + j]));
      b[(size_t)j * n + i] = v;
      b[(size_t)i * n + j] = std::conj(v);
    }
  std::vector<cd> u;
  jacobi_eigh(b, n, eps, u);
  // V = Linv^H U
  vmat.assign((size_t)n * n, cd(0));
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd acc = 0;
      for (int k = i; k < n; ++k) acc += std::conj(li[(size_t)i * n + k]) * u[(size_t)j * n + k];
      vmat[(size_t)j * n + i] = acc;
    }
  return 0;
}

// Rotate band-major block: out_i = sum_j V_ji x_j (subspace_diag.rs:380-388).
void rotate(const cd* x, const std::vector<cd>& v, int n_pw, int n_b, cd* out) {
#pragma omp parallel for schedule(static)
  for (int g0 = 0; g0 < n_pw; g0 += 256) {
    const int g1 = std::min(n_pw, g0 + 256);
    for (int i = 0; i < n_b; ++i) {
      cd* o = out + (size_t)i * n_pw;
      for (int g = g0; g < g1; ++g) o[g] = 0;
      for (int j = 0; j < n_b; ++j) {
        const cd vji = v[(size_t)i * n_b + j];
        const cd* xj = x + (size_t)j * n_pw;
        for (int g = g0; g < g1; ++g) o[g] += vji * xj[g];
      }
    }
  }
}

}  // namespace orc
