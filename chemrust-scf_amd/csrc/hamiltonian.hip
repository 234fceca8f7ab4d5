// Fused local-potential H|psi> for the SCF eigensolver hot path.
//
// Reference: apply_v_loc_hamiltonian (src/eigensolver/chebyshev.rs:279-365)
// runs init_kinetic, zero_buffer, scatter_pw_to_grid, a batched cuFFT
// inverse, veff_multiply, a batched cuFFT forward and gather_add_kinetic over
// n_b full N^3 grids. Here the same operator is three kernels that never
// materialise a full grid:
//   A  k_stick_inv : per (band, 32 sticks): scatter the sticks' plane waves
//                    into LDS, inverse z-FFT, write the stick buffer
//                    sbuf[b][z][stick] (coalesced over sticks);
//   B  k_plane     : per (band, z-plane): sticks -> LDS plane in compact
//                    form [iy][active ix], inverse y-FFT on the active
//                    columns, then row chunks: expand to full x rows,
//                    inverse x-FFT, x V_eff(r), forward x-FFT, compact;
//                    forward y-FFT, sticks back to sbuf (in place);
//   C  k_stick_fwd : per (band, 32 sticks): forward z-FFT in LDS, gather the
//                    plane waves, x 1/N^3, + 0.5|k+G|^2 psi, and the
//                    Chebyshev three-term update of chebyshev.rs:700-762
//                    fused in the same epilogue, with per-block norm
//                    partial sums for the divergence guard (:546-563).
#include <algorithm>
#include <cmath>

#include "fft.hpp"
#include "impl.hpp"
#include "scf_dev.hpp"

namespace chs {

constexpr int kSB = 32;  // sticks per block in the z passes

struct StickArgs {
  int n_st, n_act, n_pw;
  const int32_t* iy;
  const int32_t* act;
  const int32_t* pw0;
  const int32_t* cnt;
  const int32_t* iz0;
  const int32_t* act_ix;
};

template <typename T> __device__ __forceinline__ cplx<T> zero_c() { return mk((T)0, (T)0); }

// --- A: scatter + inverse z-FFT -------------------------------------------
template <int N, typename T>
__global__ void __launch_bounds__(256) k_stick_inv(const cplx<T>* __restrict__ psi, StickArgs st,
                                                   cplx<T>* __restrict__ sbuf,
                                                   const cplx<T>* __restrict__ tw) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[kSB * LS];
  const int s0 = blockIdx.x * kSB;
  const int b = blockIdx.y;
  const int ns = min(kSB, st.n_st - s0);
  for (int i = threadIdx.x; i < kSB * LS; i += blockDim.x) lds[i] = zero_c<T>();
  __syncthreads();
  const cplx<T>* pb = psi + (size_t)b * st.n_pw;
  for (int t = threadIdx.x; t < ns * N; t += blockDim.x) {
    const int sl = t / N, j = t - sl * N;
    const int s = s0 + sl;
    if (j < st.cnt[s]) {
      int iz = st.iz0[s] + j;
      if (iz >= N) iz -= N;
      lds[sl * LS + iz] = pb[st.pw0[s] + j];
    }
  }
  lds_fft_lines<N, 1, T>(lds, ns, [](int l) { return l * LS; }, 1, tw);
  cplx<T>* out = sbuf + (size_t)b * N * st.n_st + s0;
  for (int t = threadIdx.x; t < N * kSB; t += blockDim.x) {
    const int z = t / kSB, sl = t - z * kSB;
    if (sl < ns) out[(size_t)z * st.n_st + sl] = lds[sl * LS + z];
  }
}

// --- B: plane pass with V_eff ----------------------------------------------
template <int N, typename T>
__global__ void __launch_bounds__(256) k_plane(cplx<T>* __restrict__ sbuf, StickArgs st,
                                               const T* __restrict__ veff, int ch,
                                               const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* cmp = reinterpret_cast<cplx<T>*>(smem_raw);
  const int RS = st.n_act + 1;
  cplx<T>* stage = cmp + N * RS;
  constexpr int LS = N + 1;
  const int z = blockIdx.x, b = blockIdx.y;
  cplx<T>* plane = sbuf + ((size_t)b * N + z) * st.n_st;
  for (int i = threadIdx.x; i < N * RS; i += blockDim.x) cmp[i] = zero_c<T>();
  __syncthreads();
  for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) cmp[st.iy[s] * RS + st.act[s]] = plane[s];
  // inverse y-FFT on the active columns
  lds_fft_lines<N, 1, T>(cmp, st.n_act, [](int a) { return a; }, RS, tw);
  const T* vz = veff + (size_t)z * N * N;
  for (int r0 = 0; r0 < N; r0 += ch) {
    const int nr = min(ch, N - r0);
    for (int i = threadIdx.x; i < ch * LS; i += blockDim.x) stage[i] = zero_c<T>();
    __syncthreads();
    for (int t = threadIdx.x; t < nr * st.n_act; t += blockDim.x) {
      const int r = t / st.n_act, a = t - r * st.n_act;
      stage[r * LS + st.act_ix[a]] = cmp[(r0 + r) * RS + a];
    }
    lds_fft_lines<N, 1, T>(stage, nr, [](int l) { return l * LS; }, 1, tw);
    for (int t = threadIdx.x; t < nr * N; t += blockDim.x) {
      const int r = t / N, x = t - r * N;
      const T v = vz[(r0 + r) * N + x];
      cplx<T> c = stage[r * LS + x];
      stage[r * LS + x] = mk(c.x * v, c.y * v);
    }
    lds_fft_lines<N, -1, T>(stage, nr, [](int l) { return l * LS; }, 1, tw);
    for (int t = threadIdx.x; t < nr * st.n_act; t += blockDim.x) {
      const int r = t / st.n_act, a = t - r * st.n_act;
      cmp[(r0 + r) * RS + a] = stage[r * LS + st.act_ix[a]];
    }
    __syncthreads();
  }
  lds_fft_lines<N, -1, T>(cmp, st.n_act, [](int a) { return a; }, RS, tw);
  for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) plane[s] = cmp[st.iy[s] * RS + st.act[s]];
}

// --- C: forward z-FFT + gather + kinetic + Chebyshev epilogue --------------
template <int N, typename T, int MODE>
__global__ void __launch_bounds__(256) k_stick_fwd(const cplx<T>* __restrict__ sbuf, StickArgs st,
                                                   const cplx<T>* __restrict__ psi,
                                                   const double* __restrict__ kin,
                                                   cplx<T>* out, const cplx<T>* aux0, T c, T inv_e,
                                                   T inv_n3, double* __restrict__ part,
                                                   const cplx<T>* __restrict__ tw) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[kSB * LS];
  __shared__ double red[2][4];
  const int s0 = blockIdx.x * kSB;
  const int b = blockIdx.y;
  const int ns = min(kSB, st.n_st - s0);
  const cplx<T>* in = sbuf + (size_t)b * N * st.n_st + s0;
  for (int t = threadIdx.x; t < N * kSB; t += blockDim.x) {
    const int z = t / kSB, sl = t - z * kSB;
    if (sl < ns) lds[sl * LS + z] = in[(size_t)z * st.n_st + sl];
  }
  lds_fft_lines<N, -1, T>(lds, ns, [](int l) { return l * LS; }, 1, tw);
  const size_t boff = (size_t)b * st.n_pw;
  double so = 0.0, sp = 0.0;
  for (int t = threadIdx.x; t < ns * N; t += blockDim.x) {
    const int sl = t / N, j = t - sl * N;
    const int s = s0 + sl;
    if (j < st.cnt[s]) {
      int iz = st.iz0[s] + j;
      if (iz >= N) iz -= N;
      const int g = st.pw0[s] + j;
      const cplx<T> v = lds[sl * LS + iz];
      const cplx<T> p = psi[boff + g];
      const T k = (T)kin[g];
      cplx<T> h = mk(k * p.x + inv_n3 * v.x, k * p.y + inv_n3 * v.y);
      cplx<T> o;
      if (MODE == EPI_HPSI) {
        o = h;
      } else if (MODE == EPI_CHEB1) {
        o = mk((h.x - c * p.x) * inv_e, (h.y - c * p.y) * inv_e);
      } else {
        const cplx<T> q = aux0[boff + g];
        o = mk((T)2 * (h.x - c * p.x) * inv_e - q.x, (T)2 * (h.y - c * p.y) * inv_e - q.y);
      }
      out[boff + g] = o;
      so += (double)o.x * o.x + (double)o.y * o.y;
      sp += (double)p.x * p.x + (double)p.y * p.y;
    }
  }
  if (part) {
    // wave reduction, then across the 4 waves
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

// Sum the per-block norm partials (fixed order) and apply the reference's
// 10x growth guard for step k; the first failing step is kept in flag[0..4].
__global__ void k_norm_guard(const double* __restrict__ part, int n_part, int k,
                             double* __restrict__ flag) {
  __shared__ double r0[256], r1[256];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nc = sqrt(r0[0]), np = sqrt(r1[0]);
    const double growth = np > 0.0 ? nc / np : 1.0;
    if (flag[0] == 0.0 && growth > 10.0) {
      flag[0] = k;
      flag[1] = np;
      flag[2] = nc;
      flag[3] = growth;
    }
  }
}

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
  std::vector<int32_t> all;
  all.reserve(5 * (size_t)n_st + n_act);
  for (auto* v : {&iy, &act, &pw0, &cnt, &iz0}) all.insert(all.end(), v->begin(), v->end());
  all.insert(all.end(), act_ix.begin(), act_ix.end());
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
  st.h_act_ix = act_ix;
  return CHS_OK;
}

static StickArgs stick_args(const StickTable& st) {
  StickArgs a;
  a.n_st = st.n_st;
  a.n_act = st.n_act;
  a.n_pw = st.n_pw;
  a.iy = st.iy;
  a.act = st.act;
  a.pw0 = st.pw0;
  a.cnt = st.cnt;
  a.iz0 = st.iz0;
  a.act_ix = st.act_ix;
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

template <typename T>
static int vloc_pipeline_t(chs_context* ctx, int precision, const void* psi_v, int n_b, void* out_v,
                           int mode, const void* aux0_v, double c, double inv_e, double* part,
                           int* n_part) {
  Hamiltonian& H = ctx->impl->ham;
  const StickTable& st = H.st;
  const int n = H.n_grid;
  using C = cplx<T>;
  const size_t sbytes = (size_t)n_b * n * st.n_st * sizeof(C);
  CHS_TRY(ensure(ctx, ctx->impl->stick_buf, sbytes));
  C* sbuf = (C*)ctx->impl->stick_buf.p;
  const void* twv = nullptr;
  CHS_TRY(twiddles(ctx, n, precision, &twv));
  const C* tw = (const C*)twv;
  const T* veff;
  if (precision == CHS_C128) {
    veff = (const T*)H.veff;
  } else {
    veff = (const T*)H.veff_f.p;
  }
  StickArgs sa = stick_args(st);
  const C* psi = (const C*)psi_v;
  C* out = (C*)out_v;
  const C* aux0 = (const C*)aux0_v;
  const double inv_n3 = 1.0 / ((double)n * n * n);
  const int gz = ceil_div(st.n_st, kSB);
  if (n_part) *n_part = gz * n_b;
  const size_t cmp_bytes = (size_t)n * (st.n_act + 1) * sizeof(C);
  const size_t row_bytes = (size_t)(n + 1) * sizeof(C);
  const int lim = max_lds_per_block(ctx);
  if (cmp_bytes + row_bytes > (size_t)lim)
    return set_error(ctx, CHS_ERR_FFT, "plane of %d^3 grid does not fit in LDS (%zu > %d bytes)",
                     n, cmp_bytes + row_bytes, lim);
  int ch = (int)((lim - cmp_bytes) / row_bytes);
  ch = std::min(ch, 32);
  ch = std::min(ch, n);
  const size_t plane_lds = cmp_bytes + (size_t)ch * row_bytes;
  CHS_DISPATCH_N(n, {
    hipLaunchKernelGGL((k_stick_inv<NN, T>), dim3(gz, n_b), dim3(256), 0, ctx->stream, psi, sa,
                       sbuf, tw);
    CHS_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL((k_plane<NN, T>), dim3(NN, n_b), dim3(256), plane_lds, ctx->stream, sbuf,
                       sa, veff, ch, tw);
    CHS_LAUNCH_CHECK(ctx);
    if (mode == EPI_HPSI)
      hipLaunchKernelGGL((k_stick_fwd<NN, T, EPI_HPSI>), dim3(gz, n_b), dim3(256), 0, ctx->stream,
                         sbuf, sa, psi, H.kin, out, aux0, (T)c, (T)inv_e, (T)inv_n3, part, tw);
    else if (mode == EPI_CHEB1)
      hipLaunchKernelGGL((k_stick_fwd<NN, T, EPI_CHEB1>), dim3(gz, n_b), dim3(256), 0,
                         ctx->stream, sbuf, sa, psi, H.kin, out, aux0, (T)c, (T)inv_e, (T)inv_n3,
                         part, tw);
    else
      hipLaunchKernelGGL((k_stick_fwd<NN, T, EPI_CHEBK>), dim3(gz, n_b), dim3(256), 0,
                         ctx->stream, sbuf, sa, psi, H.kin, out, aux0, (T)c, (T)inv_e, (T)inv_n3,
                         part, tw);
    CHS_LAUNCH_CHECK(ctx);
  });
  return CHS_OK;
}

// The three kernels of one Chebyshev step (steady-state epilogue) `reps`
// times with HIP events around each launch on the context stream; ms[k] is
// the mean duration of kernel k (A, B, C).
template <typename T>
static int vloc_timed_t(chs_context* ctx, int precision, const void* psi_v, int n_b, void* out_v,
                        int reps, double* ms) {
  Hamiltonian& H = ctx->impl->ham;
  const StickTable& st = H.st;
  const int n = H.n_grid;
  using C = cplx<T>;
  CHS_TRY(ensure(ctx, ctx->impl->stick_buf, (size_t)n_b * n * st.n_st * sizeof(C)));
  C* sbuf = (C*)ctx->impl->stick_buf.p;
  const void* twv = nullptr;
  CHS_TRY(twiddles(ctx, n, precision, &twv));
  const C* tw = (const C*)twv;
  const T* veff = precision == CHS_C128 ? (const T*)H.veff : (const T*)H.veff_f.p;
  StickArgs sa = stick_args(st);
  const C* psi = (const C*)psi_v;
  C* out = (C*)out_v;
  const double inv_n3 = 1.0 / ((double)n * n * n);
  const int gz = ceil_div(st.n_st, kSB);
  const size_t cmp_bytes = (size_t)n * (st.n_act + 1) * sizeof(C);
  const size_t row_bytes = (size_t)(n + 1) * sizeof(C);
  const int lim = max_lds_per_block(ctx);
  if (cmp_bytes + row_bytes > (size_t)lim) return set_error(ctx, CHS_ERR_FFT, "plane too large");
  const int ch = std::min(std::min(32, n), (int)((lim - cmp_bytes) / row_bytes));
  const size_t plane_lds = cmp_bytes + (size_t)ch * row_bytes;
  hipEvent_t ev[4];
  for (auto& e : ev) CHS_HIP(ctx, hipEventCreate(&e));
  double tot[3] = {0, 0, 0};
  for (int r = 0; r < reps; ++r) {
    CHS_DISPATCH_N(n, {
      CHS_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
      hipLaunchKernelGGL((k_stick_inv<NN, T>), dim3(gz, n_b), dim3(256), 0, ctx->stream, psi, sa,
                         sbuf, tw);
      CHS_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
      hipLaunchKernelGGL((k_plane<NN, T>), dim3(NN, n_b), dim3(256), plane_lds, ctx->stream,
                         sbuf, sa, veff, ch, tw);
      CHS_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
      hipLaunchKernelGGL((k_stick_fwd<NN, T, EPI_CHEBK>), dim3(gz, n_b), dim3(256), 0,
                         ctx->stream, sbuf, sa, psi, H.kin, out, out, (T)0.5, (T)0.1, (T)inv_n3,
                         nullptr, tw);
      CHS_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
    });
    CHS_LAUNCH_CHECK(ctx);
    CHS_HIP(ctx, hipEventSynchronize(ev[3]));
    for (int k = 0; k < 3; ++k) {
      float t = 0;
      CHS_HIP(ctx, hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      tot[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 3; ++k) ms[k] = tot[k] / reps;
  return CHS_OK;
}

int vloc_timed(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int reps,
               double* ms) {
  if (precision == CHS_C128) return vloc_timed_t<double>(ctx, precision, psi, n_b, out, reps, ms);
  return vloc_timed_t<float>(ctx, precision, psi, n_b, out, reps, ms);
}

int vloc_pipeline(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int mode,
                  const void* aux0, const void* /*aux1*/, double c, double inv_e,
                  double* d_norm_part, int* n_part) {
  if (precision == CHS_C128)
    return vloc_pipeline_t<double>(ctx, precision, psi, n_b, out, mode, aux0, c, inv_e,
                                   d_norm_part, n_part);
  return vloc_pipeline_t<float>(ctx, precision, psi, n_b, out, mode, aux0, c, inv_e, d_norm_part,
                                n_part);
}

// --- D: density planes ------------------------------------------------------
// Per (z-plane, chunk of bands): for each band with occ > 0, sticks -> LDS
// plane, inverse y-FFT, then row chunks: expand, inverse x-FFT and
// accumulate occ * |psi(r)|^2 in registers (thread t owns plane elements
// t + 256 i). The chunk's partial plane goes to part[chunk][z][N^2].
template <int N>
__global__ void __launch_bounds__(256) k_density_plane(const z64* __restrict__ sbuf, StickArgs st,
                                                       const double* __restrict__ occ, int n_b,
                                                       int chunk, int ch, double* __restrict__ part,
                                                       const z64* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  z64* cmp = reinterpret_cast<z64*>(smem_raw);
  const int RS = st.n_act + 1;
  z64* stage = cmp + N * RS;
  constexpr int LS = N + 1;
  constexpr int PER = (N * N + 255) / 256;
  const int z = blockIdx.x, c = blockIdx.y;
  const int b0 = c * chunk, b1 = min(n_b, b0 + chunk);
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.0;
  for (int b = b0; b < b1; ++b) {
    const double f = occ[b];
    if (f == 0.0) continue;
    const z64* plane = sbuf + ((size_t)b * N + z) * st.n_st;
    __syncthreads();
    for (int i = threadIdx.x; i < N * RS; i += blockDim.x) cmp[i] = mk(0.0, 0.0);
    __syncthreads();
    for (int s = threadIdx.x; s < st.n_st; s += blockDim.x) cmp[st.iy[s] * RS + st.act[s]] = plane[s];
    lds_fft_lines<N, 1, double>(cmp, st.n_act, [](int a) { return a; }, RS, tw);
    for (int r0 = 0; r0 < N; r0 += ch) {
      const int nr = min(ch, N - r0);
      for (int i = threadIdx.x; i < ch * LS; i += blockDim.x) stage[i] = mk(0.0, 0.0);
      __syncthreads();
      for (int t = threadIdx.x; t < nr * st.n_act; t += blockDim.x) {
        const int r = t / st.n_act, a = t - r * st.n_act;
        stage[r * LS + st.act_ix[a]] = cmp[(r0 + r) * RS + a];
      }
      lds_fft_lines<N, 1, double>(stage, nr, [](int l) { return l * LS; }, 1, tw);
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = threadIdx.x + 256 * i;
        const int row = e / N;
        if (e < N * N && row >= r0 && row < r0 + nr) {
          const z64 v = stage[(row - r0) * LS + (e - row * N)];
          acc[i] += f * (v.x * v.x + v.y * v.y);
        }
      }
      __syncthreads();
    }
  }
  double* out = part + ((size_t)c * N + z) * N * N;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + 256 * i;
    if (e < N * N) out[e] = acc[i];
  }
}

__global__ void k_density_reduce(const double* __restrict__ part, int n_chunks, size_t n3,
                                 double inv_omega, double* __restrict__ rho) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  double s = 0;
  for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * n3 + i];
  rho[i] = s * inv_omega;
}

int density_impl(chs_context* ctx, const z64* psi, int n_b, const double* occ_host, double omega,
                 double* d_rho) {
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  auto* im = ctx->impl;
  const StickTable& st = H.st;
  const int n = H.n_grid;
  const size_t n3 = (size_t)n * n * n;
  // only bands up to the last occupied one
  int nb = 0;
  for (int b = 0; b < n_b; ++b)
    if (occ_host[b] != 0.0) nb = b + 1;
  const int chunk = 16;
  const int n_chunks = std::max(1, (nb + chunk - 1) / chunk);
  CHS_TRY(ensure(ctx, im->ws6, (size_t)n_b * sizeof(double) + (size_t)n_chunks * n3 * sizeof(double)));
  double* docc = (double*)im->ws6.p;
  double* part = docc + n_b;
  CHS_HIP(ctx, hipMemcpyAsync(docc, occ_host, n_b * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += n_b * sizeof(double);
  if (nb == 0) {
    CHS_HIP(ctx, hipMemsetAsync(d_rho, 0, n3 * sizeof(double), ctx->stream));
    return CHS_OK;
  }
  CHS_TRY(ensure(ctx, im->stick_buf, (size_t)nb * n * st.n_st * sizeof(z64)));
  z64* sbuf = (z64*)im->stick_buf.p;
  const void* twv = nullptr;
  CHS_TRY(twiddles(ctx, n, CHS_C128, &twv));
  const z64* tw = (const z64*)twv;
  StickArgs sa = stick_args(st);
  const size_t cmp_bytes = (size_t)n * (st.n_act + 1) * sizeof(z64);
  const size_t row_bytes = (size_t)(n + 1) * sizeof(z64);
  const int lim = max_lds_per_block(ctx);
  if (cmp_bytes + row_bytes > (size_t)lim)
    return set_error(ctx, CHS_ERR_FFT, "density plane of %d^3 grid does not fit in LDS", n);
  int ch = std::min(32, std::min(n, (int)((lim - cmp_bytes) / row_bytes)));
  const size_t lds = cmp_bytes + (size_t)ch * row_bytes;
  const int gz = ceil_div(st.n_st, kSB);
  CHS_DISPATCH_N(n, {
    hipLaunchKernelGGL((k_stick_inv<NN, double>), dim3(gz, nb), dim3(256), 0, ctx->stream, psi, sa,
                       sbuf, tw);
    CHS_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL((k_density_plane<NN>), dim3(NN, n_chunks), dim3(256), lds, ctx->stream,
                       sbuf, sa, docc, nb, chunk, ch, part, tw);
    CHS_LAUNCH_CHECK(ctx);
  });
  hipLaunchKernelGGL(k_density_reduce, dim3(ceil_div(n3, 256)), dim3(256), 0, ctx->stream, part,
                     n_chunks, n3, 1.0 / omega, d_rho);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
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

int chs_set_veff(chs_context* ctx, const double* d_veff) {
  if (!ctx || !d_veff) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  H.veff = d_veff;
  H.veff_f_valid = false;
  return CHS_OK;
}

}  // extern "C"
