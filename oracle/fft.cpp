// Mixed-radix (4, 2, 3, generic) complex FFT for the FP64 oracle.
// Test infrastructure only (see orc.hpp).
//
// Conventions follow the reference's cuFFT use (src/eigensolver/chebyshev.rs
// :331,:346 and src/device/fft.rs:177-222): both directions unnormalised,
// inverse = e^{+i 2 pi jk/n}, forward = e^{-i 2 pi jk/n}; the 1/N^3 is applied
// by the caller (the gather at chebyshev.rs:136-137).
#include "orc.hpp"

#include <cmath>
#include <map>
#include <mutex>

namespace orc {

namespace {

struct Plan {
  int n;
  std::vector<int> fac;
  std::vector<cd> roots[2];  // [0]: sign -1, [1]: sign +1; roots[t] = e^{s 2 pi i t/n}
};

Plan make_plan(int n) {
  Plan p;
  p.n = n;
  int m = n;
  while (m % 4 == 0) { p.fac.push_back(4); m /= 4; }
  while (m % 2 == 0) { p.fac.push_back(2); m /= 2; }
  while (m % 3 == 0) { p.fac.push_back(3); m /= 3; }
  for (int f = 5; m > 1 && f <= 16; f += 2)
    while (m % f == 0) { p.fac.push_back(f); m /= f; }
  if (m > 1) p.fac.clear();  // unsupported length: fft1d/fft3d leave data untouched
  for (int s = 0; s < 2; ++s) {
    p.roots[s].resize(n);
    double sg = s ? 1.0 : -1.0;
    for (int t = 0; t < n; ++t) {
      double a = 2.0 * M_PI * (double)t / (double)n;
      p.roots[s][t] = cd(std::cos(a), sg * std::sin(a));
    }
  }
  return p;
}

const Plan& get_plan(int n) {
  static std::mutex mu;
  static std::map<int, Plan> plans;
  std::lock_guard<std::mutex> lk(mu);
  auto it = plans.find(n);
  if (it == plans.end()) it = plans.emplace(n, make_plan(n)).first;
  return it->second;
}

// Decimation in time: out[0..n) = DFT of in[0], in[s], in[2s], ...
void rec(const cd* in, cd* out, int n, int istride, const int* fac, const cd* roots,
         int N) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int r = fac[0];
  const int m = n / r;
  for (int j = 0; j < r; ++j) rec(in + (size_t)j * istride, out + (size_t)j * m, m, istride * r, fac + 1, roots, N);
  const int step = N / n;  // w_n^t = roots[t * step], t < n
  cd tmp[16], res[16], wr[16 * 16];  // radices are <= 16 (make_plan)
  for (int j = 0; j < r; ++j)
    for (int q = 0; q < r; ++q) wr[j * r + q] = roots[((j * q) % r) * (N / r)];
  for (int k = 0; k < m; ++k) {
    for (int j = 0; j < r; ++j) tmp[j] = out[(size_t)j * m + k] * roots[j * k * step];
    for (int q = 0; q < r; ++q) {
      cd s = tmp[0];
      for (int j = 1; j < r; ++j) s += tmp[j] * wr[j * r + q];
      res[q] = s;
    }
    for (int q = 0; q < r; ++q) out[k + (size_t)q * m] = res[q];
  }
}

}  // namespace

void fft1d(cd* x, int n, int stride, int sign) {
  const Plan& p = get_plan(n);
  if (p.fac.empty()) return;
  std::vector<cd> in(n), out(n);
  for (int i = 0; i < n; ++i) in[i] = x[(size_t)i * stride];
  rec(in.data(), out.data(), n, 1, p.fac.data(), p.roots[sign > 0 ? 1 : 0].data(), n);
  for (int i = 0; i < n; ++i) x[(size_t)i * stride] = out[i];
}

void fft3d(cd* g, int n, int sign) {
  const Plan& p = get_plan(n);
  if (p.fac.empty()) return;
  const cd* roots = p.roots[sign > 0 ? 1 : 0].data();
  const size_t nn = (size_t)n * n;
#pragma omp parallel
  {
    std::vector<cd> in(n), out(n);
    // x lines (contiguous)
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nn; ++line) {
      cd* base = g + (size_t)line * n;
      rec(base, out.data(), n, 1, p.fac.data(), roots, n);
      for (int i = 0; i < n; ++i) base[i] = out[i];
    }
    // y lines (stride n)
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nn; ++line) {
      const size_t iz = line / n, ix = line % n;
      cd* base = g + iz * nn + ix;
      for (int i = 0; i < n; ++i) in[i] = base[(size_t)i * n];
      rec(in.data(), out.data(), n, 1, p.fac.data(), roots, n);
      for (int i = 0; i < n; ++i) base[(size_t)i * n] = out[i];
    }
    // z lines (stride n^2)
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nn; ++line) {
      cd* base = g + line;
      for (int i = 0; i < n; ++i) in[i] = base[(size_t)i * nn];
      rec(in.data(), out.data(), n, 1, p.fac.data(), roots, n);
      for (int i = 0; i < n; ++i) base[(size_t)i * nn] = out[i];
    }
  }
}

// Fine -> wave grid by G-space truncation, the reference's
// downsample_array_to_wave_grid (src/scf.rs:413-454): forward FFT of the fine
// real array, each wave-grid point takes the fine coefficient of its
// frequency (f = i for i <= n/2, i - n above: the Nyquist plane of an even
// wave grid takes the fine grid's +n/2 coefficient), inverse FFT on the wave
// grid, divided by the fine grid's point count; the real part is the result.
void downsample(const double* fine, int nf, int nw, double* out) {
  const size_t nf3 = (size_t)nf * nf * nf, nw3 = (size_t)nw * nw * nw;
  std::vector<cd> g(nf3), w(nw3);
  for (size_t i = 0; i < nf3; ++i) g[i] = cd(fine[i], 0.0);
  fft3d(g.data(), nf, -1);
  auto f2i = [](int i, int n, int m) {  // wave index i (grid n) -> fine index (grid m)
    const int f = i <= n / 2 ? i : i - n;
    return f >= 0 ? f : m + f;
  };
  for (int iz = 0; iz < nw; ++iz)
    for (int iy = 0; iy < nw; ++iy)
      for (int ix = 0; ix < nw; ++ix)
        w[(size_t)ix + (size_t)nw * (iy + (size_t)nw * iz)] =
            g[(size_t)f2i(ix, nw, nf) + (size_t)nf * (f2i(iy, nw, nf) + (size_t)nf * f2i(iz, nw, nf))];
  fft3d(w.data(), nw, +1);
  for (size_t i = 0; i < nw3; ++i) out[i] = w[i].real() / (double)nf3;
}

}  // namespace orc
