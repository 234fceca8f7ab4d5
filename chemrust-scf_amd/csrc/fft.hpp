// Register codelets and LDS-staged line FFTs for gfx950.
//
// A length-N line transform (N = N1 * N2) runs as two passes over LDS:
//   pass 1: per (line, n2) one thread loads x[n1*N2 + n2] (n1 < N1) into
//           registers, does a DFT_N1, multiplies by w_N^{n2*k1} and stores
//           back in place;
//   pass 2: per (line, k1) one thread loads x[k1*N2 + n2] (n2 < N2), does a
//           DFT_N2 and stores X[k1 + N1*k2] (natural order).
// Codelets DFT_R (R = 2, 3, 4 directly; 6, 8, 9, 12, 16 as R = A*B
// Cooley-Tukey with constant twiddles) are fully unrolled in registers.
// S = +1 is the inverse transform e^{+i 2 pi jk/N}, S = -1 the forward
// e^{-i 2 pi jk/N}; neither is normalised.
#pragma once
#include "common.hpp"
#include "roots.inc"

namespace chs {

template <int R> struct Roots;
#define CHS_ROOTS(R)                                                                     \
  template <> struct Roots<R> {                                                          \
    static __device__ __forceinline__ double re(int j) { return kRootRe##R[j]; }         \
    static __device__ __forceinline__ double im(int j) { return kRootIm##R[j]; }         \
  };
CHS_ROOTS(6)
CHS_ROOTS(8)
CHS_ROOTS(9)
CHS_ROOTS(12)
CHS_ROOTS(16)
#undef CHS_ROOTS

// i * S * z
template <int S, typename T> __device__ __forceinline__ cplx<T> rot_i(cplx<T> z) {
  return S > 0 ? mk(-z.y, z.x) : mk(z.y, -z.x);
}

template <int R, int S, typename T> struct Dft;

template <int S, typename T> struct Dft<1, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {}
};

template <int S, typename T> struct Dft<2, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    cplx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <int S, typename T> struct Dft<3, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    const T h = (T)0.8660254037844386;  // sqrt(3)/2
    cplx<T> t1 = a[1] + a[2];
    cplx<T> t2 = mk(a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y);
    cplx<T> d = a[1] - a[2];
    cplx<T> t3 = rot_i<S>(mk(h * d.x, h * d.y));
    a[0] = a[0] + t1;
    a[1] = t2 + t3;
    a[2] = t2 - t3;
  }
};

template <int S, typename T> struct Dft<4, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    cplx<T> s02 = a[0] + a[2], d02 = a[0] - a[2];
    cplx<T> s13 = a[1] + a[3], d13 = rot_i<S>(a[1] - a[3]);
    a[0] = s02 + s13;
    a[2] = s02 - s13;
    a[1] = d02 + d13;
    a[3] = d02 - d13;
  }
};

// R = A * B: X[k1 + A*k2] = sum_{n2} w_B^{n2 k2} w_R^{n2 k1} sum_{n1} w_A^{n1 k1} x[B n1 + n2]
template <int A, int B, int S, typename T> struct DftCT {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    constexpr int R = A * B;
    cplx<T> y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      cplx<T> t[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n1] = a[B * n1 + n2];
      Dft<A, S, T>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        const int e = (n2 * k1) % R;
        if (e == 0) {
          y[k1 * B + n2] = t[k1];
        } else if ((4 * e) % R == 0) {
          // w = (i S)^q, q = 4 e / R: an exact rotation (a multiply by the
          // table's 0 and 1 entries gives the same values, but the compiler
          // keeps x * 0, which is not 0 for an infinite or NaN x)
          const int q = 4 * e / R;
          y[k1 * B + n2] = q == 1 ? rot_i<S>(t[k1]) : q == 2 ? mk(-t[k1].x, -t[k1].y)
                                                             : rot_i<-S>(t[k1]);
        } else {
          const cplx<T> w = mk((T)Roots<R>::re(e), (T)(S * Roots<R>::im(e)));
          y[k1 * B + n2] = t[k1] * w;
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      cplx<T> t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[k1 * B + n2];
      Dft<B, S, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) a[k1 + A * k2] = t[k2];
    }
  }
};

template <int S, typename T> struct Dft<6, S, T> : DftCT<2, 3, S, T> {};
template <int S, typename T> struct Dft<8, S, T> : DftCT<2, 4, S, T> {};
template <int S, typename T> struct Dft<9, S, T> : DftCT<3, 3, S, T> {};
template <int S, typename T> struct Dft<12, S, T> : DftCT<4, 3, S, T> {};
template <int S, typename T> struct Dft<16, S, T> : DftCT<4, 4, S, T> {};

// Factorisation N = N1 * N2 used for each supported line length.
template <int N> struct Split;
#define CHS_SPLIT(N, A, B)            \
  template <> struct Split<N> {       \
    static constexpr int N1 = A;      \
    static constexpr int N2 = B;      \
  };
CHS_SPLIT(8, 2, 4)
CHS_SPLIT(12, 3, 4)
CHS_SPLIT(16, 4, 4)
CHS_SPLIT(18, 2, 9)
CHS_SPLIT(24, 4, 6)
CHS_SPLIT(32, 4, 8)
CHS_SPLIT(36, 4, 9)
CHS_SPLIT(48, 4, 12)
CHS_SPLIT(64, 8, 8)
CHS_SPLIT(72, 8, 9)
CHS_SPLIT(96, 8, 12)
CHS_SPLIT(108, 12, 9)
CHS_SPLIT(128, 8, 16)
CHS_SPLIT(144, 12, 12)
#undef CHS_SPLIT

// Zero window of a plane line. The wave-function sphere on a grid at least
// twice its diameter leaves every line index in [LO, HI) = [Z0 N2, Z1 N2)
// empty: the first-stage inputs n1 in [Z0, Z1) are zeros, and the outputs
// there are not needed. Host code checks the stick table against the window
// before it launches a kernel that relies on it. K2A..K2B: the second-stage
// output blocks k2 whose indices k1 + N1 k2 (any k1 < N1) all fall inside.
// RS: compact-plane row stride for such a table (every column outside the
// window plus one spare, = 1 mod 8 to spread column accesses over the banks).
template <int N> struct Window {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int Z0 = (N1 + 3) / 4, Z1 = N1 - Z0;
  static constexpr int LO = Z0 * N2, HI = Z1 * N2;
  static constexpr int K2A = (LO + N1 - 1) / N1, K2B = HI >= N1 ? (HI - N1) / N1 + 1 : 0;
  static constexpr int RS0 = N - (HI > LO ? HI - LO : 0) + 1;
  static constexpr int RS = RS0 + ((9 - RS0 % 8) % 8);
  static constexpr bool zero_in(int n1) { return Z0 <= n1 && n1 < Z1; }
  static constexpr bool skip_out(int k2) { return K2A <= k2 && k2 < K2B; }
  // Row order of the plane pass: slot t takes row 32 (t / 32) + (U t mod 32)
  // (t below the last whole block of 32; the rest in order), U = N2 / RS
  // mod 32, so the groups of a wavefront (consecutive t) start their compact
  // rows N2 words apart modulo the 32 eight-byte bank pairs and a row load
  // or store by a wavefront touches every bank pair once.
  static constexpr int inv32(int a) {
    int x = 1;
    for (int i = 0; i < 5; ++i) x = (x * (2 - a * x)) & 31;
    return x;
  }
  static constexpr int U = (inv32(RS) * N2) & 31;
  static constexpr int N32 = N & ~31;
  static __device__ __forceinline__ int row(int t) {
    return (U & 1) && t < N32 ? (t & ~31) | ((U * t) & 31) : t;
  }
};

// Twiddle w_N^t = e^{S 2 pi i t / N}; tw holds the S = +1 table (t < N).
template <int S, typename T>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T>* tw, int t) {
  cplx<T> w = tw[t];
  return S > 0 ? w : mk(w.x, -w.y);
}

// FFT of `nlines` lines of length N held in LDS. Element j of line l lives at
// lds[off(l) + j * es]. Called by every thread of the block (has barriers);
// starts and ends with a barrier-consistent state. tw: length-N root table
// (global memory or LDS).
template <int N, int S, typename T, typename Off>
__device__ __forceinline__ void lds_fft_lines(cplx<T>* lds, int nlines, Off off, int es,
                                              const cplx<T>* tw) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  const int nt = blockDim.x;
  __syncthreads();
  for (int t = threadIdx.x; t < nlines * N2; t += nt) {
    const int l = t / N2, n2 = t - l * N2;
    cplx<T>* base = lds + off(l) + n2 * es;
    cplx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) a[n1] = base[n1 * N2 * es];
    Dft<N1, S, T>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = a[k1] * twiddle<S>(tw, n2 * k1);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) base[k1 * N2 * es] = a[k1];
  }
  __syncthreads();
  // pass 2 in rounds of whole lines, so that no round overwrites inputs of
  // a later round of the same line
  const int lines_per_round = nt / N1 > 0 ? nt / N1 : 1;
  for (int l0 = 0; l0 < nlines; l0 += lines_per_round) {
    const int t = threadIdx.x;
    const int l = l0 + t / N1, k1 = t % N1;
    const bool act = (t < lines_per_round * N1) && (l < nlines);
    cplx<T> a[N2];
    cplx<T>* base = act ? lds + off(l) : lds;
    if (act) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) a[n2] = base[(k1 * N2 + n2) * es];
      Dft<N2, S, T>::run(a);
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) base[(k1 + N1 * k2) * es] = a[k2];
    }
    __syncthreads();
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Lane-group line FFT: one line of length N = N1 * N2 in LDS is transformed
// by the N2 lanes of a group that lives inside one wavefront, so the two
// exchange points need only wavefront-scope ordering, not a workgroup
// barrier. Lane n2 does the DFT_N1 of x[n1*N2 + n2] and its twiddles in
// registers; lane k1 (and k1 + N2, ...) then does the DFT_N2 of row k1 and
// hands X[k1 + N1*k2] to `post(k, v)`, whose result is stored in place.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct PostId {
  template <typename T> __device__ __forceinline__ cplx<T> operator()(int, cplx<T> v) const {
    return v;
  }
};

template <int N, int S, typename T, typename Post>
__device__ __forceinline__ void group_fft_line(cplx<T>* base, int es, int lane,
                                               const cplx<T>* __restrict__ tw, Post post) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int R = (N1 + N2 - 1) / N2;
  {
    cplx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) a[n1] = base[(n1 * N2 + lane) * es];
    Dft<N1, S, T>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = a[k1] * twiddle<S>(tw, lane * k1);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) base[(k1 * N2 + lane) * es] = a[k1];
  }
  wave_sync();
  cplx<T> b[R][N2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[r][n2] = base[(k1 * N2 + n2) * es];
      Dft<N2, S, T>::run(b[r]);
    }
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int k = k1 + N1 * k2;
        base[k * es] = post(k, b[r][k2]);
      }
    }
  }
  wave_sync();
}

// N = 108 = 12 x 9: the second stage of a line has N1 = 12 rows k1 of 9
// points for the 9 lanes of a group, so rows k1 = 9, 10, 11 would take a
// second round on 3 lanes at the cost of a full one. Those three 9-point
// transforms are instead spread over all 9 lanes, lane = 3 t + j doing row
// k1 = 9 + t as a 3 x 3 Cooley-Tukey (n = 3 a + b, k = c + 3 d):
//   X[c + 3 d] = sum_b w_3^{b d} w_9^{b c} sum_a w_3^{a c} x[3 a + b]
// with one exchange through the group's scratch between the two DFT_3s
// (lane j = b first, then lane j = c). Complex64 only: the complex128 plane
// kernel holds one workgroup per CU, and its extra exchange latency costs
// more than the instructions saved (6.3 against 5.8 ms per launch).
template <int N, typename T> struct Mid9Tail {
  static constexpr bool on = Split<N>::N1 == 12 && Split<N>::N2 == 9 && sizeof(T) == 4;
  static constexpr int RR = on ? 1 : (Split<N>::N1 + Split<N>::N2 - 1) / Split<N>::N2;
  // Exchange tiles in the group scratch: the value lane (t, x) hands to lane
  // (t, y) sits at slot E + 3 t + x + 17 y, so the writes (fixed y) and the
  // reads (fixed x) of the 7 groups of a wavefront spread evenly over the
  // LDS banks (with the group scratch stride of 137 slots).
  static constexpr int E1 = 0, E2 = 48;
  static __device__ __forceinline__ int at(int t, int x, int y) { return 3 * t + x + 17 * y; }
};

// One x-row of the plane pass: inverse transform from the compact row (only
// active columns stored; column x sits at compact index act_of[x], or is
// absent), V_eff multiply, forward transform back into the compact row, with
// two exchanges through the group's N-element scratch `sc` and everything
// else in registers:
//   inverse: lane n2 does DFT_N1 over x[n1 N2 + n2] and twiddles; lane k1
//            then DFT_N2 over n2, giving y[k1 + N1 k2];
//   y *= V;  lane k1 does the forward DFT_N2 over k2 and twiddles
//            w_N^{-k1 m2}; lane m2 then DFT_N1 over k1, giving
//            Z[m2 + N2 m1], stored for the active columns only.
// A row plan gives what is the same for every row of a lane: the compact
// indices it reads (in) and writes (out) and its twiddles w1(k1) =
// w_N^{+lane k1}, w2(r, m2) = w_N^{-k1 m2} with k1 = lane + r N2.
// RowPlanRegs loads them once per group into registers; RowPlanLoad reads
// them per row (for launch shapes whose register budget is tight).
template <int N, typename T> struct RowPlanRegs {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int RR = Mid9Tail<N, T>::RR;
  int in_idx[N1];
  int out_idx[N1];
  cplx<T> tw1[N1];
  cplx<T> tw2[RR][N2];
  cplx<T> tw9[2];  // Mid9Tail: w_9^{j c}, c = 1, 2
  cplx<T> twt[3];  // Mid9Tail: w_N^{-k1 m2}, k1 = 9 + t, m2 = j + 3 q
  __device__ __forceinline__ RowPlanRegs(const int* __restrict__ act_of,
                                         const cplx<T>* __restrict__ tw, int lane) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      in_idx[n1] = act_of[n1 * N2 + lane];
      out_idx[n1] = act_of[lane + N2 * n1];
      tw1[n1] = twiddle<1>(tw, (lane * n1) % N);
    }
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
      for (int m2 = 0; m2 < N2; ++m2) {
        const int k1 = lane + r * N2;
        tw2[r][m2] = twiddle<-1>(tw, k1 < N1 ? (k1 * m2) % N : 0);
      }
    if constexpr (Mid9Tail<N, T>::on) {
      const int t = lane / 3, j = lane - 3 * t;
#pragma unroll
      for (int c = 1; c < 3; ++c) tw9[c - 1] = twiddle<1>(tw, (N / 9) * j * c);
#pragma unroll
      for (int q = 0; q < 3; ++q) twt[q] = twiddle<-1>(tw, (N2 + t) * (j + 3 * q));
    }
  }
  // Completes the twiddle loads at this point: an empty asm statement takes
  // every twiddle in use as a register operand, so the compiler waits for
  // them here and the waits it places in a later loop count only newer loads.
  __device__ __forceinline__ void settle() {
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) asm volatile("" : "+v"(tw1[k1].x), "+v"(tw1[k1].y));
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
      for (int m2 = 1; m2 < N2; ++m2) asm volatile("" : "+v"(tw2[r][m2].x), "+v"(tw2[r][m2].y));
    if constexpr (Mid9Tail<N, T>::on) {
#pragma unroll
      for (int c = 0; c < 2; ++c) asm volatile("" : "+v"(tw9[c].x), "+v"(tw9[c].y));
#pragma unroll
      for (int q = 0; q < 3; ++q) asm volatile("" : "+v"(twt[q].x), "+v"(twt[q].y));
    }
  }
  __device__ __forceinline__ int in(int n1) const { return in_idx[n1]; }
  __device__ __forceinline__ int out(int m1) const { return out_idx[m1]; }
  __device__ __forceinline__ cplx<T> w1(int k1) const { return tw1[k1]; }
  __device__ __forceinline__ cplx<T> w2(int r, int m2) const { return tw2[r][m2]; }
  __device__ __forceinline__ cplx<T> w9(int c) const { return tw9[c - 1]; }
  __device__ __forceinline__ cplx<T> wt(int q) const { return twt[q]; }
};

// Per-lane twiddle table of RowPlanLds in LDS: row `lane` (N2 rows, pitch P)
// holds w_N^{lane k}, k < P - 1 (every w1 and, conjugated, every w2 of the
// lane); the odd pitch spreads a group's rows over the banks.
template <int N> struct LaneTw {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int P = (N1 > N2 ? N1 : N2) + 1;
  static constexpr int SIZE = N2 * P;
  template <typename T>
  static __device__ __forceinline__ void fill(cplx<T>* twl, const cplx<T>* __restrict__ tw) {
    for (int i = threadIdx.x; i < SIZE; i += blockDim.x) {
      const int l = i / P, k = i - l * P;
      twl[i] = tw[(l * k) % N];
    }
  }
};

// Row plan with the row twiddles read from the LaneTw table (one base
// register per lane, the entries at immediate offsets) instead of held in
// registers: about 38 registers fewer than RowPlanRegs, which lets the
// 512-thread complex64 plane kernel keep V_eff one row ahead within the 128
// registers of four wavefronts per SIMD. The Mid9Tail twiddles stay in
// registers. Single-round middle stages only (Mid9Tail<N, T>::RR == 1).
template <int N, typename T> struct RowPlanLds {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static_assert(Mid9Tail<N, T>::RR == 1, "RowPlanLds needs a single-round middle stage");
  int in_idx[N1];
  int out_idx[N1];
  const cplx<T>* t1;
  cplx<T> tw9[2];
  cplx<T> twt[3];
  __device__ __forceinline__ RowPlanLds(const int* __restrict__ act_of, const cplx<T>* twl,
                                        const cplx<T>* __restrict__ tw, int lane) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      in_idx[n1] = act_of[n1 * N2 + lane];
      out_idx[n1] = act_of[lane + N2 * n1];
    }
    t1 = twl + lane * LaneTw<N>::P;
    if constexpr (Mid9Tail<N, T>::on) {
      const int t = lane / 3, j = lane - 3 * t;
#pragma unroll
      for (int c = 1; c < 3; ++c) tw9[c - 1] = twiddle<1>(tw, (N / 9) * j * c);
#pragma unroll
      for (int q = 0; q < 3; ++q) twt[q] = twiddle<-1>(tw, (N2 + t) * (j + 3 * q));
    }
  }
  __device__ __forceinline__ void settle() {
    if constexpr (Mid9Tail<N, T>::on) {
#pragma unroll
      for (int c = 0; c < 2; ++c) asm volatile("" : "+v"(tw9[c].x), "+v"(tw9[c].y));
#pragma unroll
      for (int q = 0; q < 3; ++q) asm volatile("" : "+v"(twt[q].x), "+v"(twt[q].y));
    }
  }
  __device__ __forceinline__ int in(int n1) const { return in_idx[n1]; }
  __device__ __forceinline__ int out(int m1) const { return out_idx[m1]; }
  __device__ __forceinline__ cplx<T> w1(int k1) const { return t1[k1]; }
  __device__ __forceinline__ cplx<T> w2(int, int m2) const { return conj(t1[m2]); }
  __device__ __forceinline__ cplx<T> w9(int c) const { return tw9[c - 1]; }
  __device__ __forceinline__ cplx<T> wt(int q) const { return twt[q]; }
};

template <int N, typename T> struct RowPlanLoad {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  const int* act_of;
  const cplx<T>* tw;
  int lane;
  __device__ __forceinline__ RowPlanLoad(const int* a, const cplx<T>* t, int l)
      : act_of(a), tw(t), lane(l) {}
  __device__ __forceinline__ int in(int n1) const { return act_of[n1 * N2 + lane]; }
  __device__ __forceinline__ int out(int m1) const { return act_of[lane + N2 * m1]; }
  __device__ __forceinline__ cplx<T> w1(int k1) const { return twiddle<1>(tw, lane * k1); }
  __device__ __forceinline__ cplx<T> w2(int r, int m2) const {
    return twiddle<-1>(tw, (lane + r * N2) * m2);
  }
  __device__ __forceinline__ cplx<T> w9(int c) const {
    return twiddle<1>(tw, (N / 9) * (lane - 3 * (lane / 3)) * c);
  }
  __device__ __forceinline__ cplx<T> wt(int q) const {
    const int t = lane / 3;
    return twiddle<-1>(tw, (N2 + t) * (lane - 3 * t + 3 * q));
  }
};

// Row scratch of row_apply_v, one per lane group, STRIDE elements apart. The
// first exchange uses it as [k1][n2] with row pitch PA, the second as
// [m2][k1] with pitch PB; the pitches and the group stride spread the four
// exchange patterns of the groups of a wavefront over the LDS banks.
template <int N, typename T> struct RowScratch {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int PA = sizeof(T) == 4 ? N2 + 1 : N2;
  static constexpr int PB = N1 + 1;
  static constexpr int SPAN = N1 * PA > N2 * PB ? N1 * PA : N2 * PB;
  static constexpr int STRIDE = SPAN + (sizeof(T) == 4 ? 17 : 20);
};

// One line of a plane (element j at line[j ES], ES compile-time) by the N2
// lanes of a group, exchanging through the group's scratch tile sc (the row
// scratch of row_apply_v, [k1][n2] with pitch RowScratch::PA) instead of
// through the line itself: read along the line, the second pass would hit an
// LDS bank pair up to four times per 32 lanes. ZIN: the zero window's inputs
// are not read; ZOUT: its outputs are not stored (Window<N>).
template <int N, int S, int ES, bool ZIN, bool ZOUT, typename T>
__device__ __forceinline__ void group_fft_line_x(cplx<T>* line, cplx<T>* sc, int lane,
                                                 const cplx<T>* __restrict__ tw) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int R = (N1 + N2 - 1) / N2, PA = RowScratch<N, T>::PA;
  using W = Window<N>;
  {
    cplx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1)
      a[n1] = ZIN && W::zero_in(n1) ? mk((T)0, (T)0) : line[(n1 * N2 + lane) * ES];
    Dft<N1, S, T>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = a[k1] * twiddle<S>(tw, lane * k1);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) sc[k1 * PA + lane] = a[k1];
  }
  wave_sync();
  using MT = Mid9Tail<N, T>;
#pragma unroll
  for (int r = 0; r < MT::RR; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
      cplx<T> b[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[n2] = sc[k1 * PA + n2];
      Dft<N2, S, T>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2)
        if (!(ZOUT && W::skip_out(k2))) line[(k1 + N1 * k2) * ES] = b[k2];
    }
  }
  if constexpr (MT::on) {  // rows k1 = 9 + t by lanes 3 t + j (Mid9Tail)
    const int t = lane / 3, j = lane - 3 * t, k1 = N2 + t;
    cplx<T> y[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = sc[k1 * PA + 3 * a + j];
    Dft<3, S, T>::run(y);
    y[1] = y[1] * twiddle<S>(tw, (N / 9) * j);
    y[2] = y[2] * twiddle<S>(tw, (N / 9) * 2 * j);
    wave_sync();  // every read of the [k1][n2] tile precedes the exchange writes
#pragma unroll
    for (int c = 0; c < 3; ++c) sc[MT::E1 + MT::at(t, j, c)] = y[c];
    wave_sync();
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) y[bb] = sc[MT::E1 + MT::at(t, bb, j)];
    Dft<3, S, T>::run(y);  // X[k2 = j + 3 d]
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // k2 in [3 d, 3 d + 3): skipped as a whole, kept as a whole, or tested
      const bool all_skip = ZOUT && W::K2A <= 3 * d && 3 * d + 3 <= W::K2B;
      const bool none_skip = !ZOUT || 3 * d + 3 <= W::K2A || 3 * d >= W::K2B;
      if (all_skip) continue;
      const int k2 = j + 3 * d;
      if (none_skip || !W::skip_out(k2)) line[(k1 + N1 * k2) * ES] = y[d];
    }
  }
  wave_sync();  // the tile is reused by the group's next line
}

// V_eff of one row as a lane of a group uses it: v[r][k2] = V(x) at
// x = k1 + N1 k2, k1 = lane + r N2 (entries of k1 >= N1 unused). The plane
// pass loads the next row's values while it transforms the current one, so
// the L2 latency of the loads is spent on the row's FFT work.
template <int N, typename T> struct RowV {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int RR = Mid9Tail<N, T>::RR;
  T v[RR][N2];
  T vt[3];  // Mid9Tail: V(x) at x = 9 + t + N1 (j + 3 d), d < 3
  __device__ __forceinline__ void load(const T* __restrict__ vrow, int lane) {
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int k1 = lane + r * N2;
      if (k1 < N1) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[r][k2] = vrow[k1 + N1 * k2];
      }
    }
    if constexpr (Mid9Tail<N, T>::on) {
      const int t = lane / 3, j = lane - 3 * t;
#pragma unroll
      for (int d = 0; d < 3; ++d) vt[d] = vrow[N2 + t + N1 * (j + 3 * d)];
    }
  }
};

// PR: the stick table fits the zero window (Window<N>): the window's inputs
// are not read and its outputs not stored. PF: V_eff comes from vv (this
// row's values in registers), and vv is refilled from next_vrow (when given)
// once the multiply has used it, so the loads run during the rest of this
// row and the start of the next; otherwise V_eff is read from vrow at the
// multiply.
template <int N, typename T, bool PR, bool PF, typename Plan>
__device__ __forceinline__ void row_apply_v(cplx<T>* crow, cplx<T>* sc, const Plan& rp,
                                            RowV<N, T>& vv, const T* __restrict__ vrow,
                                            const T* __restrict__ next_vrow, int lane) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int R = (N1 + N2 - 1) / N2;
  using W = Window<N>;
  {
    cplx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      if (PR && W::zero_in(n1)) {
        a[n1] = mk((T)0, (T)0);
      } else {
        const int ai = rp.in(n1);
        a[n1] = ai >= 0 ? crow[ai] : mk((T)0, (T)0);
      }
    }
    Dft<N1, 1, T>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = a[k1] * rp.w1(k1);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) sc[k1 * RowScratch<N, T>::PA + lane] = a[k1];
  }
  wave_sync();
  using MT = Mid9Tail<N, T>;
  constexpr int RR = MT::RR;
  cplx<T> b[RR][N2];
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[r][n2] = sc[k1 * RowScratch<N, T>::PA + n2];
      Dft<N2, 1, T>::run(b[r]);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        T v;
        if constexpr (PF)
          v = vv.v[r][k2];
        else
          v = vrow[k1 + N1 * k2];
        b[r][k2] = mk(b[r][k2].x * v, b[r][k2].y * v);
      }
      Dft<N2, -1, T>::run(b[r]);
#pragma unroll
      for (int m2 = 1; m2 < N2; ++m2) b[r][m2] = b[r][m2] * rp.w2(r, m2);
    }
  }
  // Mid9Tail: rows k1 = 9 + t by lanes 3 t + j; y ends as Z[m2 = j + 3 q]
  cplx<T> y[3];
  int tk1 = 0, tj = 0;
  if constexpr (MT::on) {
    const int t = lane / 3, j = lane - 3 * t, k1 = N2 + t;
    tk1 = k1;
    tj = j;
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = sc[k1 * RowScratch<N, T>::PA + 3 * a + j];
    Dft<3, 1, T>::run(y);
    y[1] = y[1] * rp.w9(1);
    y[2] = y[2] * rp.w9(2);
    wave_sync();  // every read of the [k1][n2] tile precedes the exchange writes
#pragma unroll
    for (int c = 0; c < 3; ++c) sc[MT::E1 + MT::at(t, j, c)] = y[c];
    wave_sync();
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) y[bb] = sc[MT::E1 + MT::at(t, bb, j)];
    Dft<3, 1, T>::run(y);  // X[j + 3 d]
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T v;
      if constexpr (PF)
        v = vv.vt[d];
      else
        v = vrow[k1 + N1 * (j + 3 * d)];
      y[d] = mk(y[d].x * v, y[d].y * v);
    }
    // forward 3 x 3 with the roles swapped: DFT_3 over d, w_9^{-j p}, then
    // (lane j = p) DFT_3 over the lanes' index c
    Dft<3, -1, T>::run(y);
    y[1] = y[1] * conj(rp.w9(1));
    y[2] = y[2] * conj(rp.w9(2));
#pragma unroll
    for (int p = 0; p < 3; ++p) sc[MT::E2 + MT::at(t, j, p)] = y[p];
    wave_sync();
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c] = sc[MT::E2 + MT::at(t, c, j)];
    Dft<3, -1, T>::run(y);  // Z[j + 3 q]
#pragma unroll
    for (int q = 0; q < 3; ++q) y[q] = y[q] * rp.wt(q);
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
#pragma unroll
      for (int m2 = 0; m2 < N2; ++m2) sc[m2 * RowScratch<N, T>::PB + k1] = b[r][m2];
    }
  }
  if constexpr (MT::on) {
#pragma unroll
    for (int q = 0; q < 3; ++q) sc[(tj + 3 * q) * RowScratch<N, T>::PB + tk1] = y[q];
  }
  if constexpr (PF)
    if (next_vrow) vv.load(next_vrow, lane);
  wave_sync();
  {
    cplx<T> c[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) c[k1] = sc[lane * RowScratch<N, T>::PB + k1];
    Dft<N1, -1, T>::run(c);
#pragma unroll
    for (int m1 = 0; m1 < N1; ++m1) {
      if (PR && W::zero_in(m1)) continue;
      const int ai = rp.out(m1);
      if (ai >= 0) crow[ai] = c[m1];
    }
  }
  wave_sync();
}

// Inverse x-transform of one compact row for the density: the first half of
// row_apply_v (zero-window inputs not read), each output y[x] then adding
// f |y[x]|^2 to orow[x] in FP64. A row belongs to one group, an x to one
// lane of it, so the update needs no atomics.
template <int N, typename T, typename Plan>
__device__ __forceinline__ void row_density(const cplx<T>* crow, cplx<T>* sc, const Plan& rp,
                                            double* __restrict__ orow, double f, int lane) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int R = (N1 + N2 - 1) / N2, PA = RowScratch<N, T>::PA;
  using W = Window<N>;
  {
    cplx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      if (W::zero_in(n1)) {
        a[n1] = mk((T)0, (T)0);
      } else {
        const int ai = rp.in(n1);
        a[n1] = ai >= 0 ? crow[ai] : mk((T)0, (T)0);
      }
    }
    Dft<N1, 1, T>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = a[k1] * rp.w1(k1);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) sc[k1 * PA + lane] = a[k1];
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k1 = lane + r * N2;
    if (k1 < N1) {
      cplx<T> b[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[n2] = sc[k1 * PA + n2];
      Dft<N2, 1, T>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const double re = (double)b[k2].x, im = (double)b[k2].y;
        orow[k1 + N1 * k2] += f * (re * re + im * im);
      }
    }
  }
  wave_sync();
}

// Group geometry of a block: groups of N2 lanes packed into each wavefront.
template <int N> struct Groups {
  static constexpr int N2 = Split<N>::N2;
  static constexpr int PER_WAVE = 64 / N2;
  __device__ __forceinline__ static int count() { return (int)(blockDim.x >> 6) * PER_WAVE; }
  // group id (or -1 for lanes outside any group) and lane within the group
  __device__ __forceinline__ static int id(int* lane) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int g = l / N2;
    *lane = l - g * N2;
    return g < PER_WAVE ? w * PER_WAVE + g : -1;
  }
};

// Dispatch a functor templated on the supported line length N.
#define CHS_DISPATCH_N(n, ...)                 \
  switch (n) {                                  \
    case 8: { constexpr int NN = 8; __VA_ARGS__; } break;     \
    case 12: { constexpr int NN = 12; __VA_ARGS__; } break;   \
    case 16: { constexpr int NN = 16; __VA_ARGS__; } break;   \
    case 18: { constexpr int NN = 18; __VA_ARGS__; } break;   \
    case 24: { constexpr int NN = 24; __VA_ARGS__; } break;   \
    case 32: { constexpr int NN = 32; __VA_ARGS__; } break;   \
    case 36: { constexpr int NN = 36; __VA_ARGS__; } break;   \
    case 48: { constexpr int NN = 48; __VA_ARGS__; } break;   \
    case 64: { constexpr int NN = 64; __VA_ARGS__; } break;   \
    case 72: { constexpr int NN = 72; __VA_ARGS__; } break;   \
    case 96: { constexpr int NN = 96; __VA_ARGS__; } break;   \
    case 108: { constexpr int NN = 108; __VA_ARGS__; } break; \
    case 128: { constexpr int NN = 128; __VA_ARGS__; } break; \
    case 144: { constexpr int NN = 144; __VA_ARGS__; } break; \
    default: return set_error(ctx, CHS_ERR_FFT, "unsupported FFT length %d", (int)(n)); \
  }

inline bool fft_length_supported(int n) {
  switch (n) {
    case 8: case 12: case 16: case 18: case 24: case 32: case 36: case 48:
    case 64: case 72: case 96: case 108: case 128: case 144:
      return true;
    default:
      return false;
  }
}

}  // namespace chs
