// Learned band-CG preconditioner: a 4 -> 64 -> 64 -> 64 -> 1 SiLU network
// with bfloat16 weights and FP32 accumulation that maps per-(band, plane
// wave) features to a preconditioner scale, trained on the device.
//
// There is no reference counterpart (the reference's band CG uses the TPA
// preconditioner only, src/eigensolver/uspp_preconditioner.rs:148-180);
// the network's output o modulates that same polynomial:
//   f = [kin_g / E_b, log E_b, V_mean, V_std],  s = t(kin_g / E_b) e^o / E_b,
// with E_b the band's kinetic energy and V the bound local potential, so an
// untrained (zero-output) network is the per-band TPA preconditioner.
// Training pairs come per band (residual r and correction e over the band's
// plane waves). The loss is the band-mean misalignment of the
// preconditioned residual with the correction,
//   L = (1/n_bands) sum_b [1 - (sum_g s_g a_g)^2 / (sum_g s_g^2 c_g sum_g |e_g|^2)],
//   a = -Re(conj(r) e), c = |r|^2,
// which is scale-free per band as band CG's line search is (one
// preconditioned step along s r leaves an error of relative size L_b);
// minimised by Adam on FP32 master weights, the forward pass always with
// the bfloat16-rounded weights.
//
// Kernels: k_mlp_pass<0> (forward over band-aligned 32-sample tiles in LDS,
// per-tile sums), k_mlp_band_sums, k_mlp_pass<1> (forward + backward with
// the band coefficients, per-block gradient partials), k_mlp_reduce
// (fixed-order sums), k_mlp_adam, k_mlp_table (one thread per (band, plane
// wave) inference), and the training-pair builder.
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas1.hpp"
#include "impl.hpp"
#include "linalg.hpp"
#include "ops.hpp"

struct chs_mlp {
  chs::DevBuf theta, m, v, grad, wbf, part, tsum, bcoef;
  int t = 0;
};

namespace chs {

constexpr int kH = 64, kIn = 4;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH,
              kNP = oB4 + 1;  // 8705
static_assert(kNP == CHS_MLP_PARAMS, "parameter count");
constexpr int kTS = 32;      // samples per training tile
constexpr int kTSP = kTS + 4;  // row stride of the [feature][sample] activations
constexpr int kWP = kH + 1;    // padded row stride of W2, W3 in LDS
constexpr int kMlpBlocks = 512;

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float sg = 1.0f / (1.0f + expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}
__device__ __forceinline__ float tpa_t(float x) {
  const float x2 = x * x, num = 27.0f + 18.0f * x + 12.0f * x2 + 8.0f * x2 * x;
  return num / (num + 16.0f * x2 * x2);
}

// bfloat16 weights -> FP32 in LDS: W1 [64][4], W2 / W3 [64][65] (row = output)
struct MlpLds {
  float w1[kH * kIn], b1[kH], w2[kH * kWP], b2[kH], w3[kH * kWP], b3[kH], w4[kH], b4;
};
__device__ void load_weights(MlpLds& L, const unsigned short* __restrict__ wbf) {
  for (int t = threadIdx.x; t < kNP; t += blockDim.x) {
    const float v = bf2f(wbf[t]);
    if (t < oB1) L.w1[t] = v;
    else if (t < oW2) L.b1[t - oB1] = v;
    else if (t < oB2) { const int q = t - oW2; L.w2[(q / kH) * kWP + q % kH] = v; }
    else if (t < oW3) L.b2[t - oB2] = v;
    else if (t < oB3) { const int q = t - oW3; L.w3[(q / kH) * kWP + q % kH] = v; }
    else if (t < oW4) L.b3[t - oB3] = v;
    else if (t < oB4) L.w4[t - oW4] = v;
    else L.b4 = v;
  }
}

__global__ void k_mlp_to_bf16(const float* __restrict__ theta, unsigned short* __restrict__ wbf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kNP) wbf[t] = f2bf(theta[t]);
}

// Sample j of a batch: i = (off + j * stride) mod n.
__device__ __forceinline__ long long batch_index(long long j, long long off, long long stride,
                                                 long long n) {
  return (long long)(((unsigned long long)off + (unsigned long long)j * (unsigned long long)stride) %
                     (unsigned long long)n);
}

// Forward + backward of nb samples over 32-sample tiles; gradient partial of
// this block to part[blockIdx.x][kNP], loss partial to lpart[blockIdx.x].
// 128 threads: thread (mb, sb) = (t >> 3, t & 7) owns a 4 x 4 output block
// of every 64 x 32 layer product; for the weight gradients it owns the
// 4 x 8 block (j, i) = (4 (t >> 3) + q, 8 (t & 7) + p).
// Batch: nbb whole bands; batch band j is band (off + j stride) mod NB of
// the pair arrays (band length Lb); tile tt = (j, q0) covers samples q0 ..
// q0 + 31 of band j. PASS 0 writes per-tile sums {sum s a, sum s^2 c,
// sum |e|^2} to tsum; PASS 1 back-propagates dL/do = g1 (g2 s c - a) s with
// the band coefficients bcoef[j] = {g1, g2}.
template <int PASS>
__global__ void __launch_bounds__(128) k_mlp_pass(const unsigned short* __restrict__ wbf,
                                                  const float* __restrict__ feat,
                                                  const float* __restrict__ rr,
                                                  const float* __restrict__ ee,
                                                  const float* __restrict__ ww, long long NB,
                                                  int Lb, int nbb, long long off,
                                                  long long stride, double* __restrict__ tsum,
                                                  const double* __restrict__ bcoef,
                                                  float* __restrict__ part) {
  __shared__ MlpLds L;
  __shared__ __attribute__((aligned(16))) float Z1[kH * kTSP], Z2[kH * kTSP], Z3[kH * kTSP];
  __shared__ __attribute__((aligned(16))) float FT[kIn * kTSP];
  __shared__ float dO[kTS];
    load_weights(L, wbf);
  const int t = threadIdx.x, mb = t >> 3, sb = t & 7;
  float g3[4][8], g2[4][8], g1[2];
  for (int q = 0; q < 4; ++q)
    for (int p = 0; p < 8; ++p) g3[q][p] = g2[q][p] = 0.0f;
  g1[0] = g1[1] = 0.0f;
  float gb1 = 0, gb2 = 0, gb3 = 0, gw4 = 0, gb4 = 0;
  const int tpb = (Lb + kTS - 1) / kTS;
  const long long tiles = (long long)nbb * tpb;
  // out[m][s] = act_in?(sum_k A(k, m) B[k][s]) for the thread's 4 x 4 block
  auto mm_fwd = [&](const float* W, int ldw, int K, const float* B, bool silu_b, float acc[4][4]) {
    for (int k = 0; k < K; ++k) {
      const float4 bv = *reinterpret_cast<const float4*>(B + k * kTSP + 4 * sb);
      float b4[4] = {bv.x, bv.y, bv.z, bv.w};
      if (silu_b)
        for (int p = 0; p < 4; ++p) b4[p] = silu(b4[p]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float a = W[(4 * mb + q) * ldw + k];
#pragma unroll
        for (int p = 0; p < 4; ++p) acc[q][p] = fmaf(a, b4[p], acc[q][p]);
      }
    }
  };
  for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int jb = (int)(tile / tpb), q0 = (int)(tile % tpb) * kTS;
    const long long base = batch_index(jb, off, stride, NB) * Lb;  // first pair of the band
    __syncthreads();
    for (int e = t; e < kIn * kTS; e += blockDim.x) {
      const int s = e / kIn, c = e % kIn;
      FT[c * kTSP + s] = q0 + s < Lb ? feat[(base + q0 + s) * kIn + c] : 0.0f;
    }
    __syncthreads();
    // forward: Z1 = W1 F + b1, Z2 = W2 silu(Z1) + b2, Z3 = W3 silu(Z2) + b3
    {
      float acc[4][4];
      for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) acc[q][p] = L.b1[4 * mb + q];
      mm_fwd(L.w1, kIn, kIn, FT, false, acc);
      for (int q = 0; q < 4; ++q)
        *reinterpret_cast<float4*>(Z1 + (4 * mb + q) * kTSP + 4 * sb) =
            make_float4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]);
    }
    __syncthreads();
    {
      float acc[4][4];
      for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) acc[q][p] = L.b2[4 * mb + q];
      mm_fwd(L.w2, kWP, kH, Z1, true, acc);
      for (int q = 0; q < 4; ++q)
        *reinterpret_cast<float4*>(Z2 + (4 * mb + q) * kTSP + 4 * sb) =
            make_float4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]);
    }
    __syncthreads();
    {
      float acc[4][4];
      for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) acc[q][p] = L.b3[4 * mb + q];
      mm_fwd(L.w3, kWP, kH, Z2, true, acc);
      for (int q = 0; q < 4; ++q)
        *reinterpret_cast<float4*>(Z3 + (4 * mb + q) * kTSP + 4 * sb) =
            make_float4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]);
    }
    __syncthreads();
    // output s, the band sums (PASS 0) or dL/do (PASS 1) per sample
    double sa = 0.0, sc = 0.0, se = 0.0;
    if (t < kTS) {
      float d = 0.0f;
      const long long i = base + q0 + t;
      if (q0 + t < Lb && ww[i] > 0.0f) {
        float o = L.b4;
        for (int k = 0; k < kH; ++k) o = fmaf(L.w4[k], silu(Z3[k * kTSP + t]), o);
        const float x = FT[0 * kTSP + t], eb = expf(FT[1 * kTSP + t]);
        const float s = tpa_t(x) * expf(o) / eb;
        const float r0 = rr[2 * i], r1 = rr[2 * i + 1], e0 = ee[2 * i], e1 = ee[2 * i + 1];
        const float a = -(r0 * e0 + r1 * e1), c = r0 * r0 + r1 * r1;
        if (PASS == 0) {
          sa = (double)s * a;
          sc = (double)s * s * c;
          se = (double)e0 * e0 + (double)e1 * e1;
        } else {
          const float g1 = (float)bcoef[2 * jb], g2 = (float)bcoef[2 * jb + 1];
          d = g1 * (g2 * s * c - a) * s;
        }
      }
      dO[t] = d;
    }
    if (PASS == 0) {
      for (int o2 = 32; o2 > 0; o2 >>= 1) {
        sa += __shfl_down(sa, o2);
        sc += __shfl_down(sc, o2);
        se += __shfl_down(se, o2);
      }
      if (t == 0) {  // samples live in threads 0 .. 31 (wave 0)
        tsum[3 * tile] = sa;
        tsum[3 * tile + 1] = sc;
        tsum[3 * tile + 2] = se;
      }
      continue;
    }
    __syncthreads();
    // output layer gradients; dZ3 = dO w4 dsilu(Z3) in place
    if (t < kH) {
      float a = 0.0f;
      for (int s = 0; s < kTS; ++s) a = fmaf(dO[s], silu(Z3[t * kTSP + s]), a);
      gw4 += a;
      if (t == 0)
        for (int s = 0; s < kTS; ++s) gb4 += dO[s];
    }
    __syncthreads();
    for (int e = t; e < kH * kTS; e += blockDim.x) {
      const int k = e / kTS, s = e % kTS;
      Z3[k * kTSP + s] = dO[s] * L.w4[k] * dsilu(Z3[k * kTSP + s]);
    }
    __syncthreads();
    // weight gradient G[j][i] += sum_s D[j][s] act(Bin[i][s]) for the 4 x 8 block
    auto wgrad = [&](const float* D, const float* Bin, bool silu_b, int nin, float g[4][8]) {
      for (int s = 0; s < kTS; s += 4) {
        float dv[4][4], bv[8][4];
        for (int q = 0; q < 4; ++q) {
          const float4 v = *reinterpret_cast<const float4*>(D + (4 * mb + q) * kTSP + s);
          dv[q][0] = v.x; dv[q][1] = v.y; dv[q][2] = v.z; dv[q][3] = v.w;
        }
        for (int p = 0; p < 8; ++p) {
          const int i = 8 * sb + p;
          if (i < nin) {
            const float4 v = *reinterpret_cast<const float4*>(Bin + i * kTSP + s);
            bv[p][0] = v.x; bv[p][1] = v.y; bv[p][2] = v.z; bv[p][3] = v.w;
            if (silu_b)
              for (int c = 0; c < 4; ++c) bv[p][c] = silu(bv[p][c]);
          } else {
            for (int c = 0; c < 4; ++c) bv[p][c] = 0.0f;
          }
        }
        for (int q = 0; q < 4; ++q)
          for (int p = 0; p < 8; ++p)
            for (int c = 0; c < 4; ++c) g[q][p] = fmaf(dv[q][c], bv[p][c], g[q][p]);
      }
    };
    // in-place dZ_prev[i][s] = dsilu(Z_prev[i][s]) * sum_j W[j][i] D[j][s]
    auto back = [&](const float* W, const float* D, float* Zp) {
      float acc[4][4];
      for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) acc[q][p] = 0.0f;
      for (int j = 0; j < kH; ++j) {
        const float4 bv = *reinterpret_cast<const float4*>(D + j * kTSP + 4 * sb);
        const float b4[4] = {bv.x, bv.y, bv.z, bv.w};
        for (int q = 0; q < 4; ++q) {
          const float a = W[j * kWP + 4 * mb + q];
          for (int p = 0; p < 4; ++p) acc[q][p] = fmaf(a, b4[p], acc[q][p]);
        }
      }
      for (int q = 0; q < 4; ++q) {
        float* z = Zp + (4 * mb + q) * kTSP + 4 * sb;
        const float4 zv = *reinterpret_cast<const float4*>(z);
        *reinterpret_cast<float4*>(z) =
            make_float4(acc[q][0] * dsilu(zv.x), acc[q][1] * dsilu(zv.y), acc[q][2] * dsilu(zv.z),
                        acc[q][3] * dsilu(zv.w));
      }
    };
    wgrad(Z3, Z2, true, kH, g3);
    if (t < kH)
      for (int s = 0; s < kTS; ++s) gb3 += Z3[t * kTSP + s];
    __syncthreads();
    back(L.w3, Z3, Z2);  // Z2 <- dZ2
    __syncthreads();
    wgrad(Z2, Z1, true, kH, g2);
    if (t < kH)
      for (int s = 0; s < kTS; ++s) gb2 += Z2[t * kTSP + s];
    __syncthreads();
    back(L.w2, Z2, Z1);  // Z1 <- dZ1
    __syncthreads();
    // W1 gradient (64 x 4 entries, 2 per thread) and b1
    for (int c = 0; c < 2; ++c) {
      const int idx = 2 * t + c, j = idx / kIn, i = idx % kIn;
      float a = 0.0f;
      for (int s = 0; s < kTS; ++s) a = fmaf(Z1[j * kTSP + s], FT[i * kTSP + s], a);
      g1[c] += a;
    }
    if (t < kH)
      for (int s = 0; s < kTS; ++s) gb1 += Z1[t * kTSP + s];
  }
  if (PASS == 0) return;
  // block partial in the flat parameter layout
  float* pb = part + (size_t)blockIdx.x * kNP;
  for (int q = 0; q < 4; ++q)
    for (int p = 0; p < 8; ++p) {
      const int j = 4 * mb + q, i = 8 * sb + p;
      pb[oW3 + j * kH + i] = g3[q][p];
      pb[oW2 + j * kH + i] = g2[q][p];
    }
  pb[oW1 + 2 * t] = g1[0];
  pb[oW1 + 2 * t + 1] = g1[1];
  if (t < kH) {
    pb[oB1 + t] = gb1;
    pb[oB2 + t] = gb2;
    pb[oB3 + t] = gb3;
    pb[oW4 + t] = gw4;
  }
  if (t == 0) pb[oB4] = gb4;
}

// Per batch band: A, B, E from its tiles (fixed order); L_b = 1 - A^2/(B E)
// for used bands (B, E > 0), 0 otherwise; bcoef = {2A/(B E) / nbb, A/B};
// lossv[j] = L_b / nbb.
__global__ void k_mlp_band_sums(const double* __restrict__ tsum, int nbb, int tpb,
                                double* __restrict__ bcoef, double* __restrict__ lossv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nbb) return;
  double A = 0, B = 0, E = 0;
  for (int q = 0; q < tpb; ++q) {
    const double* p = tsum + 3 * ((size_t)j * tpb + q);
    A += p[0];
    B += p[1];
    E += p[2];
  }
  const bool used = B > 0.0 && E > 0.0;
  bcoef[2 * j] = used ? 2.0 * A / (B * E) / nbb : 0.0;
  bcoef[2 * j + 1] = used ? A / B : 0.0;
  lossv[j] = used ? (1.0 - A * A / (B * E)) / nbb : 0.0;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int nblk,
                             const double* __restrict__ lossv, int nl, float* __restrict__ grad,
                             double* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNP) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * kNP + p];
    grad[p] = (float)s;
  }
  if (p == 0) {
    double s = 0.0;
    for (int b = 0; b < nl; ++b) s += lossv[b];
    *loss = s;
  }
}

__global__ void k_mlp_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float bc1, float bc2,
                           unsigned short* __restrict__ wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  const float gp = g[p];
  const float mp = 0.9f * m[p] + 0.1f * gp;
  const float vp = 0.999f * v[p] + 0.001f * gp * gp;
  m[p] = mp;
  v[p] = vp;
  const float th = theta[p] - lr * (mp / bc1) / (sqrtf(vp / bc2) + 1e-8f);
  theta[p] = th;
  wbf[p] = f2bf(th);
}

// E_b = sum_g kin_g |psi_bg|^2 / sum_g |psi_bg|^2, one block per band.
__global__ void __launch_bounds__(256) k_band_ekin(const z64* __restrict__ psi,
                                                   const double* __restrict__ kin, int n_pw,
                                                   double* __restrict__ ekin) {
  __shared__ double sh[8];
  const z64* pb = psi + (size_t)blockIdx.x * n_pw;
  double a = 0.0, b = 0.0;
  for (int g = threadIdx.x; g < n_pw; g += blockDim.x) {
    const double w = norm2(pb[g]);
    a += kin[g] * w;
    b += w;
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) ekin[blockIdx.x] = sb > 0.0 ? sa / sb : 0.0;
}

// mean and standard deviation of the bound local potential (one block)
__global__ void __launch_bounds__(256) k_grid_stats(const double* __restrict__ v, size_t n3,
                                                    double* __restrict__ out) {
  __shared__ double sh[8];
  double a = 0.0, b = 0.0;
  for (size_t i = threadIdx.x; i < n3; i += blockDim.x) {
    a += v[i];
    b += v[i] * v[i];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    const double mean = sa / (double)n3;
    out[0] = mean;
    out[1] = sqrt(fmax(sb / (double)n3 - mean * mean, 0.0));
  }
}

// One thread per (band, plane wave): scale table for band CG.
__global__ void __launch_bounds__(256) k_mlp_table(const unsigned short* __restrict__ wbf,
                                                   const double* __restrict__ kin,
                                                   const double* __restrict__ ekin,
                                                   const double* __restrict__ vstat, int n_pw,
                                                   float* __restrict__ table) {
  __shared__ MlpLds L;
  load_weights(L, wbf);
  __syncthreads();
  const int b = blockIdx.y;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  const float eb = (float)ekin[b];
  const float f[4] = {(float)(kin[g] / ekin[b]), logf(eb), (float)vstat[0], (float)vstat[1]};
  float h1[kH], h2[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    float a = L.b1[j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a = fmaf(L.w1[j * kIn + i], f[i], a);
    h1[j] = silu(a);
  }
#pragma unroll 4
  for (int j = 0; j < kH; ++j) {
    float a = L.b2[j];
#pragma unroll
    for (int i = 0; i < kH; ++i) a = fmaf(L.w2[j * kWP + i], h1[i], a);
    h2[j] = silu(a);
  }
  float o = L.b4;
#pragma unroll 4
  for (int j = 0; j < kH; ++j) {
    float a = L.b3[j];
#pragma unroll
    for (int i = 0; i < kH; ++i) a = fmaf(L.w3[j * kWP + i], h2[i], a);
    o = fmaf(L.w4[j], silu(a), o);
  }
  table[(size_t)b * n_pw + g] = tpa_t(f[0]) * expf(o) / eb;
}

// ---- training pairs -------------------------------------------------------
// Per band b of a block psi (not yet converged) and the converged block
// psi_ref: psi^ = psi / sqrt(<psi|S|psi>), r = (H - eps S) psi^ with
// eps = <psi^|H|psi^>, e = phase-aligned psi_ref_b - psi^ with its
// psi^ component removed. dots[b] slots: 0 <psi|S psi>, 1 <psi|H psi>,
// 2 <psi_ref|S psi>, 3 <S psi^ | e>, 4 |e|^2.
__global__ void __launch_bounds__(256) k_pair_dots(const z64* __restrict__ x, const z64* __restrict__ y,
                                                   int n_pw, z64* __restrict__ out, int slot) {
  __shared__ double sh[8];
  const z64* xb = x + (size_t)blockIdx.x * n_pw;
  const z64* yb = y + (size_t)blockIdx.x * n_pw;
  double re = 0, im = 0;
  for (int g = threadIdx.x; g < n_pw; g += blockDim.x) {
    const z64 c = cmulc(xb[g], yb[g]);
    re += c.x;
    im += c.y;
  }
  const double a = block_sum(re, sh);
  const double b = block_sum(im, sh);
  if (threadIdx.x == 0) out[(size_t)blockIdx.x * 8 + slot] = mk(a, b);
}

// r = (hp - eps sp) / sqrt(n) ; e = t - psi^ with t (in e) the normalised
// projection of psi^ on the reference eigenspace of the band
__global__ void k_pair_build(const z64* __restrict__ psi, const z64* __restrict__ hp,
                             const z64* __restrict__ sp, int n_pw, const z64* __restrict__ dots,
                             z64* __restrict__ r, z64* __restrict__ e) {
  const int b = blockIdx.y, g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  const z64* d = dots + (size_t)b * 8;
  const double nn = d[0].x, inv = 1.0 / sqrt(nn), eps = d[1].x / nn;
  const size_t i = (size_t)b * n_pw + g;
  const z64 p = psi[i], h = hp[i], s = sp[i];
  r[i] = mk((h.x - eps * s.x) * inv, (h.y - eps * s.y) * inv);
  e[i] = e[i] - mk(p.x * inv, p.y * inv);
}

// e -= psi^ <S psi^|e>, then the f32 pair arrays and features
__global__ void k_pair_emit(const z64* __restrict__ psi, const z64* __restrict__ r,
                            const z64* __restrict__ e, const double* __restrict__ kin,
                            const double* __restrict__ ekin, const double* __restrict__ vstat,
                            const double* __restrict__ wband, int n_pw,
                            const z64* __restrict__ dots, float* __restrict__ feat,
                            float* __restrict__ rf, float* __restrict__ ef, float* __restrict__ wf) {
  const int b = blockIdx.y, g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  const z64* d = dots + (size_t)b * 8;
  // psi^ <S psi^|e> = psi <S psi|e> / <psi|S psi>
  const double inv2 = 1.0 / d[0].x;
  const z64 proj = d[3];
  const size_t i = (size_t)b * n_pw + g;
  const z64 p = psi[i];
  const z64 ev = e[i] - mk(inv2 * (p.x * proj.x - p.y * proj.y), inv2 * (p.x * proj.y + p.y * proj.x));
  const z64 rv = r[i];
  rf[2 * i] = (float)rv.x;
  rf[2 * i + 1] = (float)rv.y;
  ef[2 * i] = (float)ev.x;
  ef[2 * i + 1] = (float)ev.y;
  wf[i] = (float)wband[b];
  feat[4 * i] = (float)(kin[g] / ekin[b]);
  feat[4 * i + 1] = (float)log(ekin[b]);
  feat[4 * i + 2] = (float)vstat[0];
  feat[4 * i + 3] = (float)vstat[1];
}

__global__ void k_sub_scaled(const z64* __restrict__ e, const z64* __restrict__ psi, int n_pw,
                             const z64* __restrict__ dots, z64* __restrict__ out) {
  const int b = blockIdx.y, g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pw) return;
  const z64* d = dots + (size_t)b * 8;
  const double inv2 = 1.0 / d[0].x;
  const z64 proj = d[3];
  const size_t i = (size_t)b * n_pw + g;
  const z64 p = psi[i];
  out[i] = e[i] - mk(inv2 * (p.x * proj.x - p.y * proj.y), inv2 * (p.x * proj.y + p.y * proj.x));
}

static int mlp_stats(chs_context* ctx, const z64* psi, int n_b, double** ekin, double** vstat) {
  Hamiltonian& H = ctx->impl->ham;
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->small, 64 * sizeof(double)));
  CHS_TRY(ensure(ctx, im->mlp_ws, (size_t)n_b * sizeof(double) * 2 + 256));
  *ekin = (double*)im->mlp_ws.p;
  *vstat = (double*)im->small.p + 40;
  hipLaunchKernelGGL(k_band_ekin, dim3(n_b), dim3(256), 0, ctx->stream, psi, H.kin, H.n_pw, *ekin);
  const size_t n3 = (size_t)H.n_grid * H.n_grid * H.n_grid;
  hipLaunchKernelGGL(k_grid_stats, dim3(1), dim3(256), 0, ctx->stream, H.veff, n3, *vstat);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

static int mlp_batch(chs_context* ctx, chs_mlp* m, const float* feat, const float* r,
                     const float* e, const float* w, long long NB, int Lb, int nbb, long long off,
                     long long stride, double* d_loss) {
  const int tpb = (Lb + kTS - 1) / kTS;
  const long long tiles = (long long)nbb * tpb;
  CHS_TRY(ensure(ctx, m->part, (size_t)kMlpBlocks * kNP * sizeof(float)));
  CHS_TRY(ensure(ctx, m->tsum, (size_t)tiles * 3 * sizeof(double)));
  CHS_TRY(ensure(ctx, m->bcoef, (size_t)nbb * 3 * sizeof(double)));
  double* bcoef = (double*)m->bcoef.p;
  double* lossv = bcoef + 2 * (size_t)nbb;
  const int nblk = (int)std::min<long long>(kMlpBlocks, std::max<long long>(1, tiles));
  const unsigned short* wbf = (const unsigned short*)m->wbf.p;
  hipLaunchKernelGGL(k_mlp_pass<0>, dim3(nblk), dim3(128), 0, ctx->stream, wbf, feat, r, e, w, NB,
                     Lb, nbb, off, stride, (double*)m->tsum.p, (const double*)nullptr,
                     (float*)nullptr);
  hipLaunchKernelGGL(k_mlp_band_sums, dim3(ceil_div(nbb, 256)), dim3(256), 0, ctx->stream,
                     (const double*)m->tsum.p, nbb, tpb, bcoef, lossv);
  hipLaunchKernelGGL(k_mlp_pass<1>, dim3(nblk), dim3(128), 0, ctx->stream, wbf, feat, r, e, w, NB,
                     Lb, nbb, off, stride, (double*)nullptr, (const double*)bcoef,
                     (float*)m->part.p);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kNP, 256)), dim3(256), 0, ctx->stream,
                     (const float*)m->part.p, nblk, (const double*)lossv, nbb, (float*)m->grad.p,
                     d_loss);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

static long long coprime_stride(long long n) {
  long long s = 1000003;
  auto gcd = [](long long a, long long b) {
    while (b) {
      const long long t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  while (gcd(s, n) != 1) s += 2;
  return s % n == 0 ? 1 : s % n;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_mlp_create(chs_context* ctx, uint64_t seed, chs_mlp** out) {
  if (!ctx || !out) return CHS_ERR_INVALID_ARG;
  *out = nullptr;
  // Glorot-uniform W1..W3 from the 48-bit LCG; biases, w4, b4 zero
  std::vector<float> th(kNP, 0.0f);
  uint64_t s = seed & 0xFFFFFFFFFFFFull;
  auto uni = [&s]() {
    s = (s * 25214903917ull + 11ull) & 0xFFFFFFFFFFFFull;
    return (double)s / 281474976710656.0;
  };
  const struct { int off, fo, fi; } layers[3] = {{oW1, kH, kIn}, {oW2, kH, kH}, {oW3, kH, kH}};
  for (const auto& l : layers) {
    const double lim = std::sqrt(6.0 / (l.fo + l.fi));
    for (int k = 0; k < l.fo * l.fi; ++k) th[l.off + k] = (float)((2.0 * uni() - 1.0) * lim);
  }
  auto* m = new chs_mlp;
  int rc = chs_mlp_set_params(ctx, m, th.data());
  if (rc != CHS_OK) {
    chs_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return CHS_OK;
}

int chs_mlp_destroy(chs_mlp* m) {
  if (!m) return CHS_ERR_INVALID_ARG;
  for (DevBuf* b : {&m->theta, &m->m, &m->v, &m->grad, &m->wbf, &m->part, &m->tsum, &m->bcoef})
    release(*b);
  delete m;
  return CHS_OK;
}

int chs_mlp_set_params(chs_context* ctx, chs_mlp* m, const float* theta) {
  if (!ctx || !m || !theta) return CHS_ERR_INVALID_ARG;
  for (DevBuf* b : {&m->theta, &m->m, &m->v, &m->grad})
    CHS_TRY(ensure(ctx, *b, kNP * sizeof(float)));
  CHS_TRY(ensure(ctx, m->wbf, kNP * sizeof(unsigned short)));
  CHS_HIP(ctx, hipMemcpyAsync(m->theta.p, theta, kNP * sizeof(float), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += kNP * sizeof(float);
  CHS_HIP(ctx, hipMemsetAsync(m->m.p, 0, kNP * sizeof(float), ctx->stream));
  CHS_HIP(ctx, hipMemsetAsync(m->v.p, 0, kNP * sizeof(float), ctx->stream));
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3(ceil_div(kNP, 256)), dim3(256), 0, ctx->stream,
                     (const float*)m->theta.p, (unsigned short*)m->wbf.p);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->h2d += kNP * sizeof(float);
  m->t = 0;
  return CHS_OK;
}

int chs_mlp_get_params(chs_context* ctx, chs_mlp* m, float* theta) {
  if (!ctx || !m || !theta || !m->theta.p) return CHS_ERR_INVALID_ARG;
  CHS_HIP(ctx, hipMemcpyAsync(theta, m->theta.p, kNP * sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += kNP * sizeof(float);
  return CHS_OK;
}

int chs_mlp_loss_grad(chs_context* ctx, chs_mlp* m, const float* d_feat, const float* d_r,
                      const float* d_e, const float* d_w, int64_t n, int band_len, double* loss,
                      float* grad) {
  if (!ctx || !m || !d_feat || !d_r || !d_e || !d_w || n <= 0 || band_len <= 0 ||
      n % band_len != 0)
    return CHS_ERR_INVALID_ARG;
  const long long NB = n / band_len;
  if (NB > (1 << 30)) return CHS_ERR_INVALID_ARG;
  CHS_TRY(ensure(ctx, ctx->impl->small, 64 * sizeof(double)));
  double* dl = (double*)ctx->impl->small.p + 32;
  CHS_TRY(mlp_batch(ctx, m, d_feat, d_r, d_e, d_w, NB, band_len, (int)NB, 0, 1, dl));
  double h = 0.0;
  CHS_HIP(ctx, hipMemcpyAsync(&h, dl, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->d2h += sizeof(double);
  if (grad) {
    CHS_HIP(ctx, hipMemcpyAsync(grad, m->grad.p, kNP * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
    ctx->d2h += kNP * sizeof(float);
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = h;
  return CHS_OK;
}

int chs_mlp_train(chs_context* ctx, chs_mlp* m, const float* d_feat, const float* d_r,
                  const float* d_e, const float* d_w, int64_t n, int band_len, int steps,
                  int batch_bands, float lr, double* loss_trace) {
  if (!ctx || !m || !d_feat || !d_r || !d_e || !d_w || n <= 0 || band_len <= 0 ||
      n % band_len != 0 || steps < 0 || batch_bands <= 0 || !(lr > 0))
    return CHS_ERR_INVALID_ARG;
  const long long NB = n / band_len;
  const int nbb = (int)std::min<long long>(batch_bands, NB);
  const long long stride = coprime_stride(NB);
  CHS_TRY(ensure(ctx, ctx->impl->mlp_ws, (size_t)std::max(steps, 1) * sizeof(double) + 256));
  double* dtrace = (double*)ctx->impl->mlp_ws.p;
  for (int s = 0; s < steps; ++s) {
    const long long off = (long long)(((unsigned long long)s * (unsigned long long)nbb) %
                                      (unsigned long long)NB) * stride % NB;
    CHS_TRY(mlp_batch(ctx, m, d_feat, d_r, d_e, d_w, NB, band_len, nbb, off, stride, dtrace + s));
    m->t += 1;
    const float bc1 = 1.0f - std::pow(0.9f, (float)m->t), bc2 = 1.0f - std::pow(0.999f, (float)m->t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(ceil_div(kNP, 256)), dim3(256), 0, ctx->stream,
                       (float*)m->theta.p, (float*)m->m.p, (float*)m->v.p, (const float*)m->grad.p,
                       lr, bc1, bc2, (unsigned short*)m->wbf.p);
    CHS_LAUNCH_CHECK(ctx);
  }
  if (loss_trace && steps > 0) {
    CHS_HIP(ctx, hipMemcpyAsync(loss_trace, dtrace, steps * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    ctx->d2h += steps * sizeof(double);
  }
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

int chs_mlp_scale_table(chs_context* ctx, chs_mlp* m, const double* d_psi, int n_b,
                        float* d_table) {
  if (!ctx || !m || !d_psi || n_b <= 0 || !d_table) return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  double *ekin, *vstat;
  CHS_TRY(mlp_stats(ctx, (const z64*)d_psi, n_b, &ekin, &vstat));
  hipLaunchKernelGGL(k_mlp_table, dim3(ceil_div(H.n_pw, 256), n_b), dim3(256), 0, ctx->stream,
                     (const unsigned short*)m->wbf.p, H.kin, ekin, vstat, H.n_pw, d_table);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int chs_mlp_pairs(chs_context* ctx, const double* d_psi, const double* d_psi_ref,
                  const double* eps_ref, int n_b, double gap_min, float* d_feat, float* d_r,
                  float* d_e, float* d_w, int* n_used) {
  if (!ctx || !d_psi || !d_psi_ref || !eps_ref || n_b <= 0 || !d_feat || !d_r || !d_e || !d_w)
    return CHS_ERR_INVALID_ARG;
  Hamiltonian& H = ctx->impl->ham;
  if (!H.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  auto* im = ctx->impl;
  const int n_pw = H.n_pw;
  const size_t ne = (size_t)n_b * n_pw, nn2 = (size_t)n_b * n_b;
  CHS_TRY(ensure(ctx, im->ws4, ne * sizeof(z64)));
  CHS_TRY(ensure(ctx, im->ws5, ne * sizeof(z64)));
  CHS_TRY(ensure(ctx, im->ws3, 2 * ne * sizeof(z64)));
  z64* HP = (z64*)im->ws4.p;
  z64* SP = (z64*)im->ws5.p;
  z64* R = (z64*)im->ws3.p;
  z64* E = R + ne;
  const z64* psi = (const z64*)d_psi;
  const z64* ref = (const z64*)d_psi_ref;
  CHS_TRY(apply_h_impl(ctx, CHS_C128, psi, n_b, HP, SP));
  // sized before mlp_stats takes pointers into it
  const size_t dots_off = (((size_t)n_b * 2 * sizeof(double) + 255) / 256) * 256;
  const size_t ov_off = dots_off + (((size_t)n_b * 8 * sizeof(z64) + 255) / 256) * 256;
  CHS_TRY(ensure(ctx, im->mlp_ws, ov_off + nn2 * sizeof(z64) + 256));
  double *ekin, *vstat;
  CHS_TRY(mlp_stats(ctx, psi, n_b, &ekin, &vstat));
  z64* dots = (z64*)((char*)im->mlp_ws.p + dots_off);
  z64* OV = (z64*)((char*)im->mlp_ws.p + ov_off);
  double* wband = ekin + n_b;
  const dim3 eg(ceil_div(n_pw, 256), n_b);
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  hipLaunchKernelGGL(k_pair_dots, dim3(n_b), dim3(256), 0, ctx->stream, psi, SP, n_pw, dots, 0);
  hipLaunchKernelGGL(k_pair_dots, dim3(n_b), dim3(256), 0, ctx->stream, psi, HP, n_pw, dots, 1);
  CHS_LAUNCH_CHECK(ctx);
  // OV[j][b] = <psi*_j | S | psi_b>
  CHS_TRY(zgemm(ctx, 'C', 'N', n_b, n_b, n_pw, one, ref, n_pw, SP, n_pw, zero, OV, n_b));
  std::vector<z64> hov(nn2), hd((size_t)n_b * 8);
  CHS_HIP(ctx, hipMemcpyAsync(hov.data(), OV, nn2 * sizeof(z64), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipMemcpyAsync(hd.data(), dots, hd.size() * sizeof(z64), hipMemcpyDeviceToHost,
                              ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += (nn2 + hd.size()) * sizeof(z64);
  // reference eigenspaces: runs of eigenvalues closer than gap_min; the run
  // touching the top of the block may be incomplete and is not used
  std::vector<int> g0(n_b), g1(n_b);
  for (int b = 0; b < n_b;) {
    int e = b + 1;
    while (e < n_b && eps_ref[e] - eps_ref[e - 1] <= gap_min) ++e;
    for (int j = b; j < e; ++j) {
      g0[j] = b;
      g1[j] = e;
    }
    b = e;
  }
  // t_b = P* psi^_b / |P* psi^_b|: masked, normalised overlap columns
  std::vector<double> wb(n_b, 0.0);
  std::vector<char> ok(n_b, 0);
  for (int b = 0; b < n_b; ++b) {
    double p2 = 0.0;
    for (int j = g0[b]; j < g1[b]; ++j) p2 += norm2(hov[(size_t)j + (size_t)b * n_b]);
    const double nn = hd[(size_t)b * 8].x;
    ok[b] = g1[b] < n_b && nn > 0.0 && p2 / nn > 0.81;
    const double sc = p2 > 0.0 ? 1.0 / std::sqrt(p2) : 0.0;
    for (int j = 0; j < n_b; ++j) {
      z64& o = hov[(size_t)j + (size_t)b * n_b];
      o = (j >= g0[b] && j < g1[b]) ? mk(o.x * sc, o.y * sc) : mk(0.0, 0.0);
    }
  }
  CHS_HIP(ctx, hipMemcpyAsync(OV, hov.data(), nn2 * sizeof(z64), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += nn2 * sizeof(z64);
  CHS_TRY(zgemm(ctx, 'N', 'N', n_pw, n_b, n_b, one, ref, n_pw, OV, n_b, zero, E, n_pw));
  hipLaunchKernelGGL(k_pair_build, eg, dim3(256), 0, ctx->stream, psi, HP, SP, n_pw, dots, R, E);
  hipLaunchKernelGGL(k_pair_dots, dim3(n_b), dim3(256), 0, ctx->stream, SP, E, n_pw, dots, 3);
  // |e|^2 after the projection, for the band weights
  hipLaunchKernelGGL(k_sub_scaled, eg, dim3(256), 0, ctx->stream, E, psi, n_pw, dots, HP);
  hipLaunchKernelGGL(k_pair_dots, dim3(n_b), dim3(256), 0, ctx->stream, HP, HP, n_pw, dots, 4);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipMemcpyAsync(hd.data(), dots, hd.size() * sizeof(z64), hipMemcpyDeviceToHost,
                              ctx->stream));
  ctx->d2h += hd.size() * sizeof(z64);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int used = 0;
  for (int b = 0; b < n_b; ++b) {
    const double e2 = hd[(size_t)b * 8 + 4].x;
    if (ok[b] && e2 > 1e-24) {
      wb[b] = 1.0 / e2;
      ++used;
    }
  }
  for (double& w : wb) w = used ? w / used : 0.0;
  CHS_HIP(ctx, hipMemcpyAsync(wband, wb.data(), n_b * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  ctx->h2d += n_b * sizeof(double);
  hipLaunchKernelGGL(k_pair_emit, eg, dim3(256), 0, ctx->stream, psi, R, E, H.kin, ekin, vstat,
                     wband, n_pw, dots, d_feat, d_r, d_e, d_w);
  CHS_LAUNCH_CHECK(ctx);
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (n_used) *n_used = used;
  return CHS_OK;
}

}  // extern "C"
