// Fused local-potential H|psi> for the SCF eigensolver hot path.
//
// Reference: apply_v_loc_hamiltonian (src/eigensolver/chebyshev.rs:279-365)
// runs init_kinetic, zero_buffer, scatter_pw_to_grid, a batched cuFFT
// inverse, veff_multiply, a batched cuFFT forward and gather_add_kinetic over
// n_b full N^3 grids. Here the same operator is three kernels that never
// materialise a full grid:
//   A  k_stick_inv : per (band, SB sticks): scatter the sticks' plane waves
//                    into LDS, inverse z-FFT, write the stick buffer
//                    sbuf[b][z][stick] (coalesced over sticks);
//   B  k_plane     : per (band, z-plane): sticks -> LDS plane in compact
//                    form [iy][active ix]; inverse y-FFT on the active
//                    columns; per row: inverse x-FFT read straight from the
//                    compact row, x V_eff(r) and the forward x-FFT in
//                    registers, active outputs back (row_apply_v); forward
//                    y-FFT; sticks back to sbuf (in place);
//   C  k_stick_fwd : per (band, SB sticks): forward z-FFT, gather the plane
//                    waves, x 1/N^3, + 0.5|k+G|^2 psi, and the Chebyshev
//                    three-term update of chebyshev.rs:700-762 in the same
//                    epilogue, with per-block norm partial sums for the
//                    divergence guard (:546-563).
// Every 1D line is transformed by a lane group inside one wavefront
// (group_fft_line), so workgroup barriers are needed only between phases;
// SB (16, 32 or 64) and the plane kernel's launch shape are tuned per grid.
#include <algorithm>
#include <cmath>

#include "fft.hpp"
#include "impl.hpp"
#include "scf_dev.hpp"

namespace chs {

// Sticks per block of the z passes (SB) and the plane kernel's row-buffer
// count are tuned per (grid, precision) by vloc_autotune; defaults SB = 32,
// all row buffers that fit.

struct StickArgs {
  int n_st, n_act, n_pw, rs;  // rs: row stride of the compact plane (= 1 mod 8)
  int ld;                     // stick-buffer row stride: n_st rounded up to 32 (256-B rows)
  int win;                    // 1: no stick or active column in the zero window (rs = Window RS)
  const int32_t* iy;
  const int32_t* act;
  const int32_t* pw0;
  const int32_t* cnt;
  const int32_t* iz0;
  const int32_t* act_ix;
  const int32_t* act_of;  // N: compact index of x column, or -1
  const int32_t* pw_siz;  // n_pw: (stick << 8) | z index of the plane wave
};

template <typename T> __device__ __forceinline__ cplx<T> zero_c() { return mk((T)0, (T)0); }

// Workgroup of the z passes: whole wavefronts enough for one lane group per
// stick of the block, so every stick's line is transformed in one round
// (16, 32, 64 sticks of a 108-point grid: 3, 5, 10 wavefronts of 7 groups).
template <int N, int SB> struct StickTpb {
  static constexpr int W = (SB + Groups<N>::PER_WAVE - 1) / Groups<N>::PER_WAVE;
  static constexpr int value = W * 64;
  static_assert(value <= 1024, "z-pass workgroup too large");
};

// --- A: scatter + inverse z-FFT -------------------------------------------
template <int N, typename T, int SB>
__global__ void __launch_bounds__((StickTpb<N, SB>::value)) k_stick_inv(const cplx<T>* __restrict__ psi, StickArgs st,
                                                   cplx<T>* __restrict__ sbuf,
                                                   const cplx<T>* __restrict__ tw) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[SB * LS];
  const int s0 = blockIdx.x * SB;
  const int b = blockIdx.y;
  const int ns = min(SB, st.n_st - s0);
  for (int i = threadIdx.x; i < SB * LS; i += blockDim.x) lds[i] = zero_c<T>();
  __syncthreads();
  // the block's sticks hold one contiguous plane-wave range
  const cplx<T>* pb = psi + (size_t)b * st.n_pw;
  const int p_end = st.pw0[s0 + ns - 1] + st.cnt[s0 + ns - 1];
  for (int gi = st.pw0[s0] + threadIdx.x; gi < p_end; gi += blockDim.x) {
    const int si = st.pw_siz[gi];
    lds[((si >> 8) - s0) * LS + (si & 255)] = pb[gi];
  }
  __syncthreads();
  int lane;
  const int g = Groups<N>::id(&lane), ng = Groups<N>::count();
  if (g >= 0)
    for (int sl = g; sl < ns; sl += ng) group_fft_line<N, 1>(lds + sl * LS, 1, lane, tw, PostId());
  __syncthreads();
  cplx<T>* out = sbuf + (size_t)b * N * st.ld + s0;
  for (int t = threadIdx.x; t < N * SB; t += blockDim.x) {
    const int z = t / SB, sl = t - z * SB;
    if (sl < ns) out[(size_t)z * st.ld + sl] = lds[sl * LS + z];
  }
}

// occ * |psi(r)|^2 accumulated into the block's own density plane row
struct AccRho {
  double* row;
  double f;
  __device__ __forceinline__ z64 operator()(int k, z64 c) const {
    row[k] += f * (c.x * c.x + c.y * c.y);
    return c;
  }
};

// compact plane <-> one full x row, by the N2 lanes of a group
template <int N, typename T>
__device__ __forceinline__ void expand_row(const cplx<T>* cmp_row, cplx<T>* rb, const StickArgs& st,
                                           int lane) {
  constexpr int N2 = Split<N>::N2;
  for (int x = lane; x < N; x += N2) rb[x] = zero_c<T>();
  wave_sync();
  for (int a = lane; a < st.n_act; a += N2) rb[st.act_ix[a]] = cmp_row[a];
  wave_sync();
}

// --- B: plane pass with V_eff ----------------------------------------------
// Block order: workgroups are dealt to the eight XCDs round-robin by launch
// index (physical block p runs on XCD p % 8); logical block
// L = (p % 8) (P / 8) + p / 8 gives every XCD a contiguous range of
// (z-plane, band) pairs, band fastest, so each XCD's L2 holds only its
// N / 8 planes of V_eff while all bands stream through them.
// PR: the stick table leaves the zero window of every line empty (st.win):
// the compact plane has the compile-time row stride Window<N>::RS, the
// y-lines exchange through the groups' row scratch (group_fft_line_x, by the
// first nrb groups) without the window's inputs (inverse) or outputs
// (forward), and the rows skip them too and run in the bank-spreading order
// Window<N>::row.
template <int N, typename T, int TPB, bool PR>
__global__ void __launch_bounds__(TPB) k_plane(cplx<T>* __restrict__ sbuf, StickArgs st,
                                               const T* __restrict__ veff, int nrb, int n_b,
                                               const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  using W = Window<N>;
  constexpr int RSW = W::RS;
  const int P = gridDim.x, p = blockIdx.x;
  const int L = (p & 7) * (P >> 3) + (p >> 3);
  if (L >= N * n_b) return;  // grid padding (whole block)
  const int z = L / n_b, b = L - z * n_b;
  const int RS = PR ? RSW : st.rs;
  cplx<T>* cmp = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* rowbuf = cmp + N * RS;
  int* act_of = reinterpret_cast<int*>(rowbuf + nrb * RowScratch<N, T>::STRIDE);
  // LT: the rows use RowPlanLds, whose twiddle table follows act_of (N is even,
  // so the table starts 8-byte aligned)
  constexpr bool LT = sizeof(T) == 4 && TPB == 512;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(act_of + N);
  cplx<T>* plane = sbuf + ((size_t)b * N + z) * st.ld;
  int lane;
  const int g = Groups<N>::id(&lane), ng = Groups<N>::count();
  const bool rows = g >= 0 && g < nrb;  // the groups that transform rows
  const T* vz = veff + (size_t)z * N * N;
  // Rows t = g, g + nrb, ... With PF each row's V_eff is loaded during the
  // work before it (the first row's from here on, the others' during the
  // row before); that holds 12 more registers per lane across the row, which
  // the 256-thread complex64 shape has (two wavefronts per SIMD, 256
  // registers each), the 512-thread one with the LDS twiddle plan, and the
  // 1024-thread shape and complex128 do not.
  constexpr bool PF = sizeof(T) == 4 && TPB <= 512;
  RowV<N, T> vv;
  if (PF && rows) vv.load(vz + (PR ? W::row(g) : g) * N, lane);
  for (int i = threadIdx.x; i < N * RS; i += blockDim.x) cmp[i] = zero_c<T>();
  for (int i = threadIdx.x; i < N; i += blockDim.x) act_of[i] = st.act_of[i];
  if (LT) LaneTw<N>::fill(twl, tw);
  __syncthreads();
  for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) cmp[st.iy[s] * RS + st.act[s]] = plane[s];
  __syncthreads();
  cplx<T>* sc = rowbuf + (g >= 0 ? g : 0) * RowScratch<N, T>::STRIDE;
  if constexpr (PR) {
    if (g >= 0 && g < nrb)
      for (int a = g; a < st.n_act; a += nrb)
        group_fft_line_x<N, 1, RSW, true, false>(cmp + a, sc, lane, tw);
  } else {
    if (g >= 0)
      for (int a = g; a < st.n_act; a += ng) group_fft_line<N, 1>(cmp + a, RS, lane, tw, PostId());
  }
  __syncthreads();
  if (rows) {
    auto rows_pass = [&](auto& rp) {
      // PF: the plan's global loads complete before the loop, whose waits
      // then count only the V_eff loads
      if (PF) rp.settle();
      for (int t = g; t < N; t += nrb) {
        const int iy = PR ? W::row(t) : t, tn = t + nrb;
        const T* vn = PF && tn < N ? vz + (PR ? W::row(tn) : tn) * N : nullptr;
        row_apply_v<N, T, PR, PF>(cmp + iy * RS, sc, rp, vv, vz + iy * N, vn, lane);
      }
    };
    if constexpr (LT) {
      RowPlanLds<N, T> rp(act_of, twl, tw, lane);
      rows_pass(rp);
    } else if (TPB <= 512) {
      RowPlanRegs<N, T> rp(act_of, tw, lane);
      rows_pass(rp);
    } else {
      const RowPlanLoad<N, T> rp(act_of, tw, lane);
      for (int t = g; t < N; t += nrb) {
        const int iy = PR ? W::row(t) : t;
        row_apply_v<N, T, PR, false>(cmp + iy * RS, sc, rp, vv, vz + iy * N, nullptr, lane);
      }
    }
  }
  __syncthreads();
  if constexpr (PR) {
    if (g >= 0 && g < nrb)
      for (int a = g; a < st.n_act; a += nrb)
        group_fft_line_x<N, -1, RSW, false, true>(cmp + a, sc, lane, tw);
  } else {
    if (g >= 0)
      for (int a = g; a < st.n_act; a += ng) group_fft_line<N, -1>(cmp + a, RS, lane, tw, PostId());
  }
  __syncthreads();
  for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) plane[s] = cmp[st.iy[s] * RS + st.act[s]];
}

// --- C: forward z-FFT + gather + kinetic + Chebyshev epilogue --------------
// c and inv_e come from `coef` (device, {c, inv_e}) when it is given, so a
// captured graph of the filter replays with new spectral bounds.
template <int N, typename T, int MODE, int SB>
__global__ void __launch_bounds__((StickTpb<N, SB>::value)) k_stick_fwd(const cplx<T>* __restrict__ sbuf, StickArgs st,
                                                   const cplx<T>* __restrict__ psi,
                                                   const double* __restrict__ kin,
                                                   cplx<T>* out, const cplx<T>* aux0, T c, T inv_e,
                                                   const double* __restrict__ coef, T inv_n3,
                                                   double* __restrict__ part,
                                                   const cplx<T>* __restrict__ tw,
                                                   const cplx<T>* __restrict__ resw,
                                                   const double* __restrict__ ress) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[SB * LS];
  __shared__ double red[2][16];  // one slot per wavefront
  if (coef) {
    c = (T)coef[0];
    inv_e = (T)coef[1];
  }
  const int s0 = blockIdx.x * SB;
  const int b = blockIdx.y;
  const int ns = min(SB, st.n_st - s0);
  const cplx<T>* in = sbuf + (size_t)b * N * st.ld + s0;
  for (int t = threadIdx.x; t < N * SB; t += blockDim.x) {
    const int z = t / SB, sl = t - z * SB;
    if (sl < ns) lds[sl * LS + z] = in[(size_t)z * st.ld + sl];
  }
  __syncthreads();
  int lane;
  const int g = Groups<N>::id(&lane), ng = Groups<N>::count();
  if (g >= 0)
    for (int sl = g; sl < ns; sl += ng) group_fft_line<N, -1>(lds + sl * LS, 1, lane, tw, PostId());
  __syncthreads();
  const size_t boff = (size_t)b * st.n_pw;
  double so = 0.0, sp = 0.0;
  const int p_end = st.pw0[s0 + ns - 1] + st.cnt[s0 + ns - 1];
  for (int gi = st.pw0[s0] + threadIdx.x; gi < p_end; gi += blockDim.x) {
    const int si = st.pw_siz[gi];
    const cplx<T> v = lds[((si >> 8) - s0) * LS + (si & 255)];
    const cplx<T> p = psi[boff + gi];
    const T k = (T)kin[gi];
    cplx<T> h = mk(k * p.x + inv_n3 * v.x, k * p.y + inv_n3 * v.y);
    cplx<T> o;
    if (MODE == EPI_HPSI) {
      o = h;
    } else if (MODE == EPI_CHEB1) {
      o = mk((h.x - c * p.x) * inv_e, (h.y - c * p.y) * inv_e);
    } else {
      const cplx<T> q = aux0[boff + gi];
      o = mk((T)2 * (h.x - c * p.x) * inv_e - q.x, (T)2 * (h.y - c * p.y) * inv_e - q.y);
      if (MODE == EPI_CHEBK_RES) {
        // residual recurrence: + (2/e) T_k(x_b) W_b, the band's scale in ress
        const cplx<T> w = resw[boff + gi];
        const T sb = (T)ress[b];
        o = mk(o.x + sb * w.x, o.y + sb * w.y);
      }
    }
    out[boff + gi] = o;
    so += (double)o.x * o.x + (double)o.y * o.y;
    sp += (double)p.x * p.x + (double)p.y * p.y;
  }
  if (part) {
    for (int off = 32; off > 0; off >>= 1) {
      so += __shfl_down(so, off);
      sp += __shfl_down(sp, off);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = so;
      red[1][threadIdx.x >> 6] = sp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0, bb = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        a += red[0][w];
        bb += red[1][w];
      }
      const size_t slot = (size_t)b * gridDim.x + blockIdx.x;
      part[2 * slot] = a;
      part[2 * slot + 1] = bb;
    }
  }
}

// --- D: density planes ------------------------------------------------------
// Per (z-plane, chunk of bands): for each band with occ > 0, sticks -> LDS
// plane, inverse y-FFT, then per row: expand and inverse x-FFT whose final
// store accumulates occ * |psi(r)|^2 into the block's own partial plane
// part[chunk][z][N^2] (each element owned by one lane: no atomics).
template <int N, int TPB>
__global__ void __launch_bounds__(TPB) k_density_plane(const z64* __restrict__ sbuf, StickArgs st,
                                                       const double* __restrict__ occ, int n_b,
                                                       int chunk, int nrb, double* __restrict__ part,
                                                       const z64* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int LS = N + 1;
  const int RS = st.rs;
  z64* cmp = reinterpret_cast<z64*>(smem_raw);
  z64* rowbuf = cmp + N * RS;
  const int z = blockIdx.x, c = blockIdx.y;
  const int b0 = c * chunk, b1 = min(n_b, b0 + chunk);
  double* out = part + ((size_t)c * N + z) * N * N;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) out[i] = 0.0;
  int lane;
  const int g = Groups<N>::id(&lane), ng = Groups<N>::count();
  for (int b = b0; b < b1; ++b) {
    const double f = occ[b];
    if (f == 0.0) continue;
    const z64* plane = sbuf + ((size_t)b * N + z) * st.ld;
    __syncthreads();
    for (int i = threadIdx.x; i < N * RS; i += blockDim.x) cmp[i] = mk(0.0, 0.0);
    __syncthreads();
    for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) cmp[st.iy[s] * RS + st.act[s]] = plane[s];
    __syncthreads();
    if (g >= 0)
      for (int a = g; a < st.n_act; a += ng) group_fft_line<N, 1>(cmp + a, RS, lane, tw, PostId());
    __syncthreads();
    if (g >= 0 && g < nrb) {
      z64* rb = rowbuf + g * LS;
      for (int iy = g; iy < N; iy += nrb) {
        expand_row<N, double>(cmp + iy * RS, rb, st, lane);
        group_fft_line<N, 1>(rb, 1, lane, tw, AccRho{out + iy * N, f});
      }
    }
  }
}

// Density planes for a stick table inside the zero window (st.win), either
// precision with FP64 accumulation: per (z-plane, chunk of bands) and
// occupied band, sticks -> compact plane (stride Window<N>::RS), inverse
// y-lines through the groups' tiles (first nrb groups), then the rows in the
// bank-spreading order, each an inverse x-transform read straight from the
// compact row (row_density).
// complex64: at most 128 registers, so two 512-thread workgroups share a CU
// (with half the row buffers, which the 160 KB LDS then holds twice)
template <int N, typename T, int TPB>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
k_density_plane_w(const cplx<T>* __restrict__ sbuf,
                  StickArgs st,
                  const double* __restrict__ occ, int n_b,
                  int chunk, int nrb,
                  double* __restrict__ part,
                  const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  using W = Window<N>;
  constexpr int RS = W::RS;
  cplx<T>* cmp = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* rowbuf = cmp + N * RS;
  int* act_of = reinterpret_cast<int*>(rowbuf + nrb * RowScratch<N, T>::STRIDE);
  const int z = blockIdx.x, c = blockIdx.y;
  const int b0 = c * chunk, b1 = min(n_b, b0 + chunk);
  double* out = part + ((size_t)c * N + z) * N * N;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) out[i] = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) act_of[i] = st.act_of[i];
  __syncthreads();
  int lane;
  const int g = Groups<N>::id(&lane);
  const bool grp = g >= 0 && g < nrb;
  cplx<T>* sc = rowbuf + (grp ? g : 0) * RowScratch<N, T>::STRIDE;
  const RowPlanRegs<N, T> rp(act_of, tw, grp ? lane : 0);
  for (int b = b0; b < b1; ++b) {
    const double f = occ[b];
    if (f == 0.0) continue;
    const cplx<T>* plane = sbuf + ((size_t)b * N + z) * st.ld;
    __syncthreads();
    for (int i = threadIdx.x; i < N * RS; i += blockDim.x) cmp[i] = zero_c<T>();
    __syncthreads();
    for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) cmp[st.iy[s] * RS + st.act[s]] = plane[s];
    __syncthreads();
    if (grp)
      for (int a = g; a < st.n_act; a += nrb)
        group_fft_line_x<N, 1, RS, true, false>(cmp + a, sc, lane, tw);
    __syncthreads();
    if (grp)
      for (int t = g; t < N; t += nrb) {
        const int iy = W::row(t);
        row_density<N, T>(cmp + iy * RS, sc, rp, out + iy * N, f, lane);
      }
  }
}

__global__ void k_density_reduce(const double* __restrict__ part, int n_chunks, size_t n3,
                                 double inv_omega, int accumulate, double* __restrict__ rho) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  double s = 0;
  for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * n3 + i];
  rho[i] = accumulate ? rho[i] + s * inv_omega : s * inv_omega;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int build_sticks(chs_context* ctx) {
  Hamiltonian& H = ctx->impl->ham;
  const int n = H.n_grid, n_pw = H.n_pw;
  std::vector<int32_t> mil(3 * (size_t)n_pw);
  CHS_HIP(ctx, hipMemcpyAsync(mil.data(), H.miller, mil.size() * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += mil.size() * sizeof(int32_t);
  std::vector<int32_t> ix, iy, pw0, cnt, iz0;
  auto wrap = [n](int v) { return v >= 0 ? v : v + n; };
  for (int g = 0; g < n_pw; ++g) {
    const int h = mil[3 * g], k = mil[3 * g + 1], l = mil[3 * g + 2];
    if (2 * std::abs(h) >= n || 2 * std::abs(k) >= n || 2 * std::abs(l) >= n)
      return set_error(ctx, CHS_ERR_INVALID_ARG, "Miller index (%d,%d,%d) outside the %d^3 grid", h,
                       k, l, n);
    const bool same = !ix.empty() && ix.back() == wrap(h) && iy.back() == wrap(k);
    if (same) {
      const int prev_l = mil[3 * (g - 1) + 2];
      if (l != prev_l + 1)
        return set_error(ctx, CHS_ERR_INVALID_ARG,
                         "plane waves must be ordered lexicographically by (h, k, l)");
      cnt.back() += 1;
    } else {
      ix.push_back(wrap(h));
      iy.push_back(wrap(k));
      pw0.push_back(g);
      cnt.push_back(1);
      iz0.push_back(wrap(l));
    }
  }
  const int n_st = (int)ix.size();
  std::vector<int32_t> act_ix(ix);
  std::sort(act_ix.begin(), act_ix.end());
  act_ix.erase(std::unique(act_ix.begin(), act_ix.end()), act_ix.end());
  std::vector<int32_t> act_of(n, -1);
  for (size_t a = 0; a < act_ix.size(); ++a) act_of[act_ix[a]] = (int32_t)a;
  std::vector<int32_t> act(n_st);
  for (int s = 0; s < n_st; ++s) act[s] = act_of[ix[s]];
  const int n_act = (int)act_ix.size();
  // plane wave -> (stick, z index): sticks are consecutive plane-wave ranges
  std::vector<int32_t> pw_siz(n_pw);
  for (int s = 0; s < n_st; ++s)
    for (int j = 0; j < cnt[s]; ++j) {
      int iz = iz0[s] + j;
      if (iz >= n) iz -= n;
      pw_siz[pw0[s] + j] = (s << 8) | iz;
    }
  std::vector<int32_t> all;
  all.reserve(5 * (size_t)n_st + n_act + n + n_pw);
  for (auto* v : {&iy, &act, &pw0, &cnt, &iz0}) all.insert(all.end(), v->begin(), v->end());
  all.insert(all.end(), act_ix.begin(), act_ix.end());
  all.insert(all.end(), act_of.begin(), act_of.end());
  all.insert(all.end(), pw_siz.begin(), pw_siz.end());
  CHS_TRY(ensure(ctx, H.st_mem, all.size() * sizeof(int32_t)));
  CHS_HIP(ctx, hipMemcpyAsync(H.st_mem.p, all.data(), all.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->h2d += all.size() * sizeof(int32_t);
  int32_t* d = (int32_t*)H.st_mem.p;
  StickTable& st = H.st;
  st.n_st = n_st;
  st.n_act = n_act;
  st.n_grid = n;
  st.n_pw = n_pw;
  st.iy = d;
  st.act = d + n_st;
  st.pw0 = d + 2 * (size_t)n_st;
  st.cnt = d + 3 * (size_t)n_st;
  st.iz0 = d + 4 * (size_t)n_st;
  st.act_ix = d + 5 * (size_t)n_st;
  st.act_of = d + 5 * (size_t)n_st + n_act;
  st.pw_siz = d + 5 * (size_t)n_st + n_act + n;
  st.h_act_ix = act_ix;
  // the plane kernels' zero window (fft.hpp Window<N>): usable when no
  // active column and no stick row falls in it
  st.rs_win = 0;
  if (ctx->impl->plane_window && fft_length_supported(n)) {
    int lo = 0, hi = 0, rsw = 0;
    CHS_DISPATCH_N(n, {
      lo = Window<NN>::LO;
      hi = Window<NN>::HI;
      rsw = Window<NN>::RS;
    });
    bool fits = hi > lo && n_act < rsw;
    for (int x : act_ix) fits = fits && (x < lo || x >= hi);
    for (int y : iy) fits = fits && (y < lo || y >= hi);
    if (fits) st.rs_win = rsw;
  }
  return CHS_OK;
}

// Row stride of the stick buffer sbuf[b][z][stick]: whole 256-byte lines per
// z-row, so the SB-stick chunks of the z passes never straddle a line.
static inline int stick_ld(int n_st) { return (n_st + 31) / 32 * 32; }

static StickArgs stick_args(const StickTable& st) {
  StickArgs a;
  a.n_st = st.n_st;
  a.ld = stick_ld(st.n_st);
  a.n_act = st.n_act;
  a.n_pw = st.n_pw;
  a.win = st.rs_win > 0 ? 1 : 0;
  a.rs = st.n_act + 1;
  while (a.rs % 8 != 1) ++a.rs;  // spreads column accesses over LDS banks
  if (a.win) a.rs = st.rs_win;
  a.iy = st.iy;
  a.act = st.act;
  a.pw0 = st.pw0;
  a.cnt = st.cnt;
  a.iz0 = st.iz0;
  a.act_ix = st.act_ix;
  a.act_of = st.act_of;
  a.pw_siz = st.pw_siz;
  return a;
}

static int max_lds_per_block(chs_context* ctx) {
  static int cached = 0;
  if (!cached) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) !=
            hipSuccess ||
        v <= 0)
      v = 65536;
    cached = v;
  }
  return cached;
}

// Dynamic LDS of the plane kernels: compact plane + nrb row buffers.
struct PlaneLaunch {
  size_t lds = 0;
  int nrb = 0;
  int threads = 256;
};
static int plane_launch(chs_context* ctx, int n, const StickArgs& sa, size_t csz, int n2,
                        int nrb_div, int threads, int row_elems, PlaneLaunch* out,
                        size_t extra_bytes = 0) {
  const size_t cmp_bytes = (size_t)n * sa.rs * csz;
  const size_t row_bytes = (size_t)row_elems * csz;
  const int lim = max_lds_per_block(ctx);
  if (cmp_bytes + row_bytes + (size_t)n * sizeof(int32_t) + extra_bytes > (size_t)lim)
    return set_error(ctx, CHS_ERR_FFT, "plane of %d^3 grid does not fit in LDS (%zu > %d bytes)",
                     n, cmp_bytes + row_bytes + extra_bytes, lim);
  const int groups = (threads / 64) * (64 / n2);
  const size_t tab_bytes = (size_t)n * sizeof(int32_t) + extra_bytes;  // act_of (+ twiddles)
  int nrb = (int)((lim - cmp_bytes - tab_bytes) / row_bytes);
  nrb = std::max(1, std::min(nrb, groups) / std::max(1, nrb_div));
  out->nrb = nrb;
  out->threads = threads;
  out->lds = cmp_bytes + (size_t)nrb * row_bytes + tab_bytes;
  return CHS_OK;
}

// Tuned launch shape of the pipeline for (grid, precision).
static VlocTune get_tune(chs_context* ctx, int n, int precision) {
  auto it = ctx->impl->vloc_tune.find(std::make_pair(n, precision));
  return it == ctx->impl->vloc_tune.end() ? VlocTune{} : it->second;
}

template <int NN, typename T, int SB>
static int vloc_launch(chs_context* ctx, const StickArgs& sa, const PlaneLaunch& pl, int n_b,
                       const cplx<T>* psi, cplx<T>* sbuf, const T* veff, const cplx<T>* tw,
                       const double* kin, cplx<T>* out, const cplx<T>* aux0, int mode, double c,
                       double inv_e, const double* coef, double inv_n3, double* part,
                       hipEvent_t* ev, const cplx<T>* resw, const double* ress) {
  const int gz = ceil_div(sa.n_st, SB);
  if (ev) CHS_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
  hipLaunchKernelGGL((k_stick_inv<NN, T, SB>), dim3(gz, n_b), dim3(StickTpb<NN, SB>::value), 0,
                     ctx->stream, psi, sa,
                     sbuf, tw);
  if (ev) CHS_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
  const unsigned pgrid = (unsigned)(((size_t)NN * n_b + 7) / 8 * 8);
  // 1024-thread blocks for complex64 only (vloc_pipeline_t never asks for
  // them with complex128, whose codelets do not fit the 128 registers per
  // lane that 16 wavefronts per CU leave)
  // the window variant needs the compact plane laid out with its stride
  if (sa.win && sa.rs != Window<NN>::RS)
    return set_error(ctx, CHS_ERR_FFT, "stick table row stride %d != window stride %d", sa.rs,
                     Window<NN>::RS);
  bool launched = false;
#define CHS_PLANE(TPB, PR)                                                                        hipLaunchKernelGGL((k_plane<NN, T, TPB, PR>), dim3(pgrid), dim3(TPB), pl.lds, ctx->stream, sbuf,                      sa, veff, pl.nrb, n_b, tw)
  if constexpr (sizeof(T) == 4) {
    if (pl.threads == 1024) {
      if (sa.win)
        CHS_PLANE(1024, true);
      else
        CHS_PLANE(1024, false);
      launched = true;
    }
  }
  if (!launched) {
    if (pl.threads == 512) {
      if (sa.win)
        CHS_PLANE(512, true);
      else
        CHS_PLANE(512, false);
    } else {
      if (sa.win)
        CHS_PLANE(256, true);
      else
        CHS_PLANE(256, false);
    }
  }
#undef CHS_PLANE
  if (ev) CHS_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
#define CHS_FWD(M)                                                                            \
  hipLaunchKernelGGL((k_stick_fwd<NN, T, M, SB>), dim3(gz, n_b), dim3(StickTpb<NN, SB>::value), \
                     0, ctx->stream,                                                           \
                     sbuf, sa, psi, kin, out, aux0, (T)c, (T)inv_e, coef, (T)inv_n3, part, tw,  \
                     resw, ress)
  if (mode == EPI_HPSI)
    CHS_FWD(EPI_HPSI);
  else if (mode == EPI_CHEB1)
    CHS_FWD(EPI_CHEB1);
  else if (mode == EPI_CHEBK_RES)
    CHS_FWD(EPI_CHEBK_RES);
  else
    CHS_FWD(EPI_CHEBK);
#undef CHS_FWD
  if (ev) CHS_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

// Sticks per block that fit the static LDS of the z passes for (N, T).
static bool sb_fits(int n, size_t csz, int sb) { return (size_t)sb * (n + 1) * csz <= 60000; }

template <typename T>
static int vloc_pipeline_t(chs_context* ctx, int precision, const void* psi_v, int n_b, void* out_v,
                           int mode, const void* aux0_v, double c, double inv_e,
                           const double* coef, double* part, int* n_part, int reps, double* ms,
                           const void* resw_v = nullptr, const double* ress = nullptr) {
  Hamiltonian& H = ctx->impl->ham;
  const StickTable& st = H.st;
  const int n = H.n_grid;
  using C = cplx<T>;
  CHS_TRY(ensure(ctx, ctx->impl->stick_buf, (size_t)n_b * n * stick_ld(st.n_st) * sizeof(C)));
  C* sbuf = (C*)ctx->impl->stick_buf.p;
  const void* twv = nullptr;
  CHS_TRY(twiddles(ctx, n, precision, &twv));
  const C* tw = (const C*)twv;
  const T* veff = precision == CHS_C128 ? (const T*)H.veff : (const T*)H.veff_f.p;
  StickArgs sa = stick_args(st);
  const C* psi = (const C*)psi_v;
  C* out = (C*)out_v;
  const C* aux0 = (const C*)aux0_v;
  const C* resw = (const C*)resw_v;
  if (mode == EPI_CHEBK_RES && (!resw || !ress))
    return set_error(ctx, CHS_ERR_INVALID_ARG, "residual epilogue without its residual");
  const double inv_n3 = 1.0 / ((double)n * n * n);
  VlocTune tn = get_tune(ctx, n, precision);
  if (!sb_fits(n, sizeof(C), tn.sb)) tn.sb = 16;
  if (tn.threads != 256 && tn.threads != 512 && (tn.threads != 1024 || sizeof(T) != 4))
    tn.threads = 512;
  if (n_part) *n_part = (int)ceil_div(st.n_st, tn.sb) * n_b;
  hipEvent_t ev[4];
  if (ms)
    for (auto& e : ev) CHS_HIP(ctx, hipEventCreate(&e));
  double tot[3] = {0, 0, 0};
  for (int r = 0; r < std::max(1, reps); ++r) {
    CHS_DISPATCH_N(n, {
      PlaneLaunch pl;
      // the 512-thread complex64 plane kernel keeps its row twiddles in LDS
      const size_t lanetw = tn.threads == 512 && sizeof(C) == 8
                                ? (size_t)LaneTw<NN>::SIZE * sizeof(C) : 0;
      CHS_TRY(plane_launch(ctx, NN, sa, sizeof(C), Split<NN>::N2, tn.nrb_div, tn.threads,
                           RowScratch<NN, T>::STRIDE, &pl, lanetw));
      hipEvent_t* evp = ms ? ev : nullptr;
      if (tn.sb == 16)
        CHS_TRY((vloc_launch<NN, T, 16>(ctx, sa, pl, n_b, psi, sbuf, veff, tw, H.kin, out, aux0,
                                        mode, c, inv_e, coef, inv_n3, part, evp, resw, ress)));
      else if (tn.sb == 64)
        CHS_TRY((vloc_launch<NN, T, 64>(ctx, sa, pl, n_b, psi, sbuf, veff, tw, H.kin, out, aux0,
                                        mode, c, inv_e, coef, inv_n3, part, evp, resw, ress)));
      else
        CHS_TRY((vloc_launch<NN, T, 32>(ctx, sa, pl, n_b, psi, sbuf, veff, tw, H.kin, out, aux0,
                                        mode, c, inv_e, coef, inv_n3, part, evp, resw, ress)));
    });
    if (ms) {
      CHS_HIP(ctx, hipEventSynchronize(ev[3]));
      for (int k = 0; k < 3; ++k) {
        float t = 0;
        CHS_HIP(ctx, hipEventElapsedTime(&t, ev[k], ev[k + 1]));
        tot[k] += t;
      }
    }
  }
  if (ms) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 3; ++k) ms[k] = tot[k] / std::max(1, reps);
  }
  return CHS_OK;
}

int vloc_pipeline(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int mode,
                  const void* aux0, const double* d_coef, double c, double inv_e,
                  double* d_norm_part, int* n_part, const void* res_w, const double* res_s) {
  if (precision == CHS_C128)
    return vloc_pipeline_t<double>(ctx, precision, psi, n_b, out, mode, aux0, c, inv_e, d_coef,
                                   d_norm_part, n_part, 1, nullptr, res_w, res_s);
  return vloc_pipeline_t<float>(ctx, precision, psi, n_b, out, mode, aux0, c, inv_e, d_coef,
                                d_norm_part, n_part, 1, nullptr, res_w, res_s);
}

// The three kernels of one steady-state Chebyshev step `reps` times with HIP
// events around each launch on the context stream; ms[k] = mean duration of
// kernel k (A, B, C). out doubles as Y_{k-2}.
int vloc_timed(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int reps,
               double* ms) {
  if (precision == CHS_C128)
    return vloc_pipeline_t<double>(ctx, precision, psi, n_b, out, EPI_CHEBK, out, 0.5, 0.1,
                                   nullptr, nullptr, nullptr, reps, ms);
  return vloc_pipeline_t<float>(ctx, precision, psi, n_b, out, EPI_CHEBK, out, 0.5, 0.1, nullptr,
                                nullptr, nullptr, reps, ms);
}

// Launch-shape autotuner of the H|psi> pipeline for the bound grid: every
// (sticks per block, plane-kernel threads, row-buffer divisor) candidate
// that fits is timed on
// n_b bands at the given precision (HIP events, `reps` runs after one
// warm-up) and the fastest is kept for (grid, precision).
int vloc_autotune(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int reps,
                  VlocTune* best_out, double* best_ms) {
  Hamiltonian& H = ctx->impl->ham;
  const int n = H.n_grid;
  const size_t csz = precision == CHS_C128 ? 16 : 8;
  auto key = std::make_pair(n, precision);
  VlocTune best{};
  double best_t = 1e300;
  for (int sb : {16, 32, 64}) {
    if (!sb_fits(n, csz, sb)) continue;
    for (int threads : {256, 512, 1024})
      for (int div : {1, 2, 4}) {
        if (threads == 1024 && csz != 8) continue;
        const VlocTune cand{sb, div, threads};
        ctx->impl->vloc_tune[key] = cand;
        double ms[3];
        CHS_TRY(vloc_timed(ctx, precision, psi, n_b, out, 1, ms));
        CHS_TRY(vloc_timed(ctx, precision, psi, n_b, out, reps, ms));
        const double t = ms[0] + ms[1] + ms[2];
        if (t < best_t) {
          best_t = t;
          best = cand;
        }
      }
  }
  ctx->impl->vloc_tune[key] = best;
  tune_cache_store(ctx, "vloc " + std::to_string(n) + " " + std::to_string(precision) + " " +
                            std::to_string(best.sb) + " " + std::to_string(best.nrb_div) + " " +
                            std::to_string(best.threads));
  if (best_out) *best_out = best;
  if (best_ms) *best_ms = best_t;
  return CHS_OK;
}

int ensure_precision_ready(chs_context* ctx, int precision);

// rho = sum_b occ_b |psi_b(r)|^2 / omega from a band-major block in either
// precision (FP64 accumulation); with `accumulate` the sum is added to rho.
template <typename T>
static int density_t(chs_context* ctx, const cplx<T>* psi, int n_b, const double* occ_host,
                     double omega, double* d_rho, bool accumulate) {
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  auto* im = ctx->impl;
  const StickTable& st = H.st;
  const int n = H.n_grid;
  const size_t n3 = (size_t)n * n * n;
  const int precision = sizeof(T) == 8 ? CHS_C128 : CHS_C64;
  CHS_TRY(ensure_precision_ready(ctx, precision));
  int nb = 0;  // bands up to the last occupied one
  for (int b = 0; b < n_b; ++b)
    if (occ_host[b] != 0.0) nb = b + 1;
  const int chunk = 8;  // bands per block (more blocks in flight; partial planes are cheap)
  const int n_chunks = std::max(1, (nb + chunk - 1) / chunk);
  CHS_TRY(ensure(ctx, im->ws6, (size_t)n_b * sizeof(double) + (size_t)n_chunks * n3 * sizeof(double)));
  double* docc = (double*)im->ws6.p;
  double* part = docc + n_b;
  CHS_HIP(ctx, hipMemcpyAsync(docc, occ_host, n_b * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += n_b * sizeof(double);
  if (nb == 0) {
    if (!accumulate) CHS_HIP(ctx, hipMemsetAsync(d_rho, 0, n3 * sizeof(double), ctx->stream));
    return CHS_OK;
  }
  CHS_TRY(ensure(ctx, im->stick_buf, (size_t)nb * n * stick_ld(st.n_st) * sizeof(cplx<T>)));
  cplx<T>* sbuf = (cplx<T>*)im->stick_buf.p;
  const void* twv = nullptr;
  CHS_TRY(twiddles(ctx, n, precision, &twv));
  const cplx<T>* tw = (const cplx<T>*)twv;
  StickArgs sa = stick_args(st);
  constexpr int SB = 16;
  const int gz = ceil_div(st.n_st, SB);
  // launch shape of the density planes, timed once per (grid, precision) on
  // the first call: workgroup size (256 or 512 threads, low 16 bits) and the
  // row-buffer divisor (high bits; rows go to fewer groups, so more
  // workgroups fit a CU's LDS). Every shape gives bit-identical partial
  // planes (each output element sums its bands in the same order).
  const int key = 2 * n + (sizeof(T) == 4 ? 1 : 0);
  auto tpb_it = im->density_tpb.find(key);
  const bool tune = tpb_it == im->density_tpb.end();
  CHS_DISPATCH_N(n, {
    if (sa.win && sa.rs != Window<NN>::RS)
      return set_error(ctx, CHS_ERR_FFT, "stick table row stride %d != window stride %d", sa.rs,
                       Window<NN>::RS);
    hipLaunchKernelGGL((k_stick_inv<NN, T, SB>), dim3(gz, nb), dim3(StickTpb<NN, SB>::value), 0,
                       ctx->stream, psi,
                       sa, sbuf, tw);
    CHS_LAUNCH_CHECK(ctx);
    auto planes = [&](int shape) -> int {
      const int tpb = (shape & 0xffff) == 512 ? 512 : 256;
      const int div = std::min(8, std::max(1, shape >> 16));
      PlaneLaunch pl;
      if (sa.win) {
        CHS_TRY(plane_launch(ctx, NN, sa, sizeof(cplx<T>), Split<NN>::N2, div, tpb,
                             RowScratch<NN, T>::STRIDE, &pl));
        if (tpb == 512)
          hipLaunchKernelGGL((k_density_plane_w<NN, T, 512>), dim3(NN, n_chunks), dim3(512),
                             pl.lds, ctx->stream, sbuf, sa, docc, nb, chunk, pl.nrb, part, tw);
        else
          hipLaunchKernelGGL((k_density_plane_w<NN, T, 256>), dim3(NN, n_chunks), dim3(256),
                             pl.lds, ctx->stream, sbuf, sa, docc, nb, chunk, pl.nrb, part, tw);
      } else if constexpr (sizeof(T) == 8) {
        CHS_TRY(plane_launch(ctx, NN, sa, sizeof(z64), Split<NN>::N2, div, tpb, NN + 1, &pl));
        if (tpb == 512)
          hipLaunchKernelGGL((k_density_plane<NN, 512>), dim3(NN, n_chunks), dim3(512), pl.lds,
                             ctx->stream, sbuf, sa, docc, nb, chunk, pl.nrb, part, tw);
        else
          hipLaunchKernelGGL((k_density_plane<NN, 256>), dim3(NN, n_chunks), dim3(256), pl.lds,
                             ctx->stream, sbuf, sa, docc, nb, chunk, pl.nrb, part, tw);
      } else {
        return set_error(ctx, CHS_ERR_INVALID_ARG,
                         "complex64 density needs a stick table inside the zero window");
      }
      CHS_LAUNCH_CHECK(ctx);
      return CHS_OK;
    };
    if (tune) {
      int best = 256;
      float best_t = 1e30f;
      for (int shape : {256 | 1 << 16, 512 | 1 << 16, 512 | 2 << 16, 512 | 4 << 16}) {
        CHS_HIP(ctx, hipEventRecord(im->ev0, ctx->stream));
        CHS_TRY(planes(shape));
        CHS_HIP(ctx, hipEventRecord(im->ev1, ctx->stream));
        CHS_HIP(ctx, hipEventSynchronize(im->ev1));
        float t = 0;
        CHS_HIP(ctx, hipEventElapsedTime(&t, im->ev0, im->ev1));
        if (t < best_t) {
          best_t = t;
          best = shape;
        }
      }
      im->density_tpb[key] = best;
      tune_cache_store(ctx, "dens " + std::to_string(key) + " " + std::to_string(best));
    } else {
      CHS_TRY(planes(tpb_it->second));
    }
  });
  hipLaunchKernelGGL(k_density_reduce, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, part,
                     n_chunks, n3, 1.0 / omega, accumulate ? 1 : 0, d_rho);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int density_impl(chs_context* ctx, const z64* psi, int n_b, const double* occ_host, double omega,
                 double* d_rho, bool accumulate) {
  return density_t<double>(ctx, psi, n_b, occ_host, omega, d_rho, accumulate);
}

int density_c64(chs_context* ctx, const c32* psi, int n_b, const double* occ_host, double omega,
                double* d_rho, bool accumulate) {
  return density_t<float>(ctx, psi, n_b, occ_host, omega, d_rho, accumulate);
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

static int refresh_veff_f(chs_context* ctx) {
  Hamiltonian& H = ctx->impl->ham;
  const size_t n3 = (size_t)H.n_grid * H.n_grid * H.n_grid;
  CHS_TRY(ensure(ctx, H.veff_f, n3 * sizeof(float)));
  hipLaunchKernelGGL(k_d2f, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, H.veff,
                     (float*)H.veff_f.p, n3);
  CHS_LAUNCH_CHECK(ctx);
  H.veff_f_valid = true;
  return CHS_OK;
}

int ensure_precision_ready(chs_context* ctx, int precision) {
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  if (precision == CHS_C64 && !H.veff_f_valid) CHS_TRY(refresh_veff_f(ctx));
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_bind_hamiltonian(chs_context* ctx, int n_grid, int n_pw, const int32_t* d_miller,
                         const double* d_kin, const double* d_veff) {
  if (!ctx || n_pw <= 0 || !d_miller || !d_kin || !d_veff) return CHS_ERR_INVALID_ARG;
  kset_swap(ctx, -1);  // a selected k-point goes back to its set
  return bind_hamiltonian(ctx, n_grid, n_pw, d_miller, d_kin, d_veff);
}

}  // extern "C"

namespace chs {

int bind_hamiltonian(chs_context* ctx, int n_grid, int n_pw, const int32_t* d_miller,
                     const double* d_kin, const double* d_veff) {
  if (!fft_length_supported(n_grid))
    return set_error(ctx, CHS_ERR_FFT, "unsupported FFT length %d", n_grid);
  Hamiltonian& H = ctx->impl->ham;
  H.bound = false;
  H.n_grid = n_grid;
  H.n_pw = n_pw;
  H.miller = d_miller;
  H.kin = d_kin;
  H.veff = d_veff;
  H.veff_f_valid = false;
  H.n_proj = 0;
  H.has_s = false;
  CHS_TRY(build_sticks(ctx));
  H.bound = true;
  return CHS_OK;
}

}  // namespace chs

extern "C" {

int chs_set_veff(chs_context* ctx, const double* d_veff) {
  if (!ctx || !d_veff) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  H.veff = d_veff;
  H.veff_f_valid = false;
  return CHS_OK;
}

}  // extern "C"
