// Batched in-place 3D FFT of full n^3 grids (BatchedFftPlan3d,
// reference src/device/fft.rs:116-228), as three LDS line passes.
#include "fft.hpp"
#include "impl.hpp"

namespace chs {

constexpr int kLines = 16;

// One pass over axis `axis` (0 = x, 1 = y, 2 = z; N = that axis' length) of
// `batch` nx x ny x nz grids (x fastest). A block transforms kLines lines:
// for x, kLines consecutive rows; for y and z, kLines consecutive ix at fixed
// iz resp. iy, so global loads coalesce.
template <int N, int S, typename T>
__global__ void __launch_bounds__(256) k_fft_axis(cplx<T>* __restrict__ g, int axis, int nx,
                                                  int ny, int nz,
                                                  const cplx<T>* __restrict__ tw) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[kLines * LS];
  const size_t nxy = (size_t)nx * ny, n3 = nxy * nz;
  cplx<T>* grid = g + (size_t)blockIdx.y * n3;
  size_t base, si, se;
  int nl = kLines;
  if (axis == 0) {
    const size_t row0 = (size_t)blockIdx.x * kLines, rows = (size_t)ny * nz;
    base = row0 * N;
    si = N;
    se = 1;
    if (row0 + nl > rows) nl = (int)(rows - row0);
  } else {
    const int gpp = (nx + kLines - 1) / kLines;  // line groups per plane
    const int o = blockIdx.x / gpp;              // iz (axis 1) or iy (axis 2)
    const int ix0 = (blockIdx.x % gpp) * kLines;
    base = (axis == 1 ? (size_t)o * nxy : (size_t)o * nx) + ix0;
    si = 1;
    se = axis == 1 ? (size_t)nx : nxy;
    if (ix0 + nl > nx) nl = nx - ix0;
  }
  const int tot = kLines * N;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int i, j;
    if (si == 1) { i = t % kLines; j = t / kLines; } else { j = t % N; i = t / N; }
    if (i < nl) lds[i * LS + j] = grid[base + i * si + j * se];
  }
  lds_fft_lines<N, S, T>(lds, nl, [](int l) { return l * LS; }, 1, tw);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int i, j;
    if (si == 1) { i = t % kLines; j = t / kLines; } else { j = t % N; i = t / N; }
    if (i < nl) grid[base + i * si + j * se] = lds[i * LS + j];
  }
}

template <typename T>
static int fft3d_impl(chs_context* ctx, int precision, T* grid, const int dims[3], int batch,
                      int sign) {
  using C = cplx<T>;
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int gpp = (nx + kLines - 1) / kLines;
  for (int axis = 0; axis < 3; ++axis) {
    const int len = dims[axis];
    const void* tw = nullptr;
    CHS_TRY(twiddles(ctx, len, precision, &tw));
    const dim3 grd(axis == 0 ? ceil_div((size_t)ny * nz, kLines)
                             : (unsigned)((axis == 1 ? nz : ny) * gpp),
                   batch);
    CHS_DISPATCH_N(len, {
      if (sign > 0)
        hipLaunchKernelGGL((k_fft_axis<NN, 1, T>), grd, dim3(256), 0, ctx->stream, (C*)grid,
                           axis, nx, ny, nz, (const C*)tw);
      else
        hipLaunchKernelGGL((k_fft_axis<NN, -1, T>), grd, dim3(256), 0, ctx->stream, (C*)grid,
                           axis, nx, ny, nz, (const C*)tw);
      CHS_LAUNCH_CHECK(ctx);
    });
  }
  return CHS_OK;
}

int fft3d_xyz(chs_context* ctx, int precision, void* grid, const int dims[3], int batch,
              int sign) {
  if (precision == CHS_C128) return fft3d_impl(ctx, precision, (double*)grid, dims, batch, sign);
  return fft3d_impl(ctx, precision, (float*)grid, dims, batch, sign);
}

int fft3d(chs_context* ctx, int precision, void* grid, int n, int batch, int sign) {
  const int dims[3] = {n, n, n};
  return fft3d_xyz(ctx, precision, grid, dims, batch, sign);
}

// Hermitian half [nz][ny][nx/2 + 1] (x fastest) -> full complex grid
// [nz][ny][nx]: entries past the half are conj(a(-x, -y, -z)).
__global__ void k_hermitian_expand(const z64* __restrict__ half, int nx, int ny, int nz,
                                   z64* __restrict__ full) {
  const size_t n3 = (size_t)nx * ny * nz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const int b = blockIdx.y;
  const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((size_t)nx * ny));
  const int nh = nx / 2 + 1;
  const z64* hb = half + (size_t)b * nh * ny * nz;
  z64 v;
  if (ix < nh) {
    v = hb[ix + (size_t)nh * (iy + (size_t)ny * iz)];
  } else {
    const int jx = nx - ix, jy = (ny - iy) % ny, jz = (nz - iz) % nz;
    v = conj(hb[jx + (size_t)nh * (jy + (size_t)ny * jz)]);
  }
  full[(size_t)b * n3 + i] = v;
}

__global__ void k_real_of(const z64* __restrict__ a, double* __restrict__ r, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = a[i].x;
}

__global__ void k_real_to_complex(const double* __restrict__ r, z64* __restrict__ c, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = mk(r[i], 0.0);
}

// Wave-grid point (ix, iy, iz) takes the fine-grid coefficient of its
// frequency: f = i for i <= n/2, i - n above (an even wave grid's Nyquist
// plane takes the fine grid's +n/2 coefficient), wrapped onto the fine grid.
__device__ __forceinline__ int wave_to_fine(int i, int n, int m) {
  const int f = i <= n / 2 ? i : i - n;
  return f >= 0 ? f : m + f;
}

__global__ void k_wave_gather(const z64* __restrict__ fine, int fx, int fy, int fz, int wx, int wy,
                              int wz, z64* __restrict__ wave) {
  const size_t n = (size_t)wx * wy * wz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ix = (int)(i % wx), iy = (int)((i / wx) % wy), iz = (int)(i / ((size_t)wx * wy));
  const int jx = wave_to_fine(ix, wx, fx), jy = wave_to_fine(iy, wy, fy),
            jz = wave_to_fine(iz, wz, fz);
  wave[i] = fine[(size_t)jx + (size_t)fx * (jy + (size_t)fy * jz)];
}

// out = Re(a) / n_fine
__global__ void k_real_scaled(const z64* __restrict__ a, double s, double* __restrict__ r,
                              size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = a[i].x * s;
}

}  // namespace chs

extern "C" {

int chs_fft3d_c2c(chs_context* ctx, int precision, void* d_grid, int nx, int ny, int nz,
                  int batch, int sign) {
  if (!ctx || !d_grid || batch <= 0 || batch > 65535 || (sign != 1 && sign != -1) ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  const int dims[3] = {nx, ny, nz};
  for (int d : dims)
    if (!chs::fft_length_supported(d))
      return chs::set_error(ctx, CHS_ERR_FFT, "unsupported FFT length %d", d);
  return chs::fft3d_xyz(ctx, precision, d_grid, dims, batch, sign);
}

int chs_fft3d_z2d(chs_context* ctx, const double* d_half, int nx, int ny, int nz, int batch,
                  double* d_real) {
  if (!ctx || !d_half || !d_real || batch <= 0 || batch > 65535) return CHS_ERR_INVALID_ARG;
  const int dims[3] = {nx, ny, nz};
  for (int d : dims)
    if (!chs::fft_length_supported(d))
      return chs::set_error(ctx, CHS_ERR_FFT, "unsupported FFT length %d", d);
  using namespace chs;
  const size_t n3 = (size_t)nx * ny * nz;
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->fft_ws, n3 * batch * sizeof(z64)));
  z64* full = (z64*)im->fft_ws.p;
  hipLaunchKernelGGL(k_hermitian_expand, dim3(ceil_div(n3, 256), batch), dim3(256), 0,
                     ctx->stream, (const z64*)d_half, nx, ny, nz, full);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d_xyz(ctx, CHS_C128, full, dims, batch, 1));
  hipLaunchKernelGGL(k_real_of, dim3(ceil_div(n3 * batch, 256)), dim3(256), 0, ctx->stream, full,
                     d_real, n3 * batch);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

int chs_downsample_to_wave_grid(chs_context* ctx, const double* d_fine, const int fine_dims[3],
                                const int wave_dims[3], double* d_wave) {
  if (!ctx || !d_fine || !d_wave || !fine_dims || !wave_dims) return CHS_ERR_INVALID_ARG;
  using namespace chs;
  for (int a = 0; a < 3; ++a) {
    if (!fft_length_supported(fine_dims[a]) || !fft_length_supported(wave_dims[a]))
      return set_error(ctx, CHS_ERR_FFT, "unsupported FFT length");
    if (wave_dims[a] > fine_dims[a])
      return set_error(ctx, CHS_ERR_INVALID_ARG, "wave grid larger than the fine grid");
  }
  const size_t nf = (size_t)fine_dims[0] * fine_dims[1] * fine_dims[2];
  const size_t nw = (size_t)wave_dims[0] * wave_dims[1] * wave_dims[2];
  auto* im = ctx->impl;
  CHS_TRY(ensure(ctx, im->fft_ws, (nf + nw) * sizeof(z64)));
  z64* g = (z64*)im->fft_ws.p;
  z64* w = g + nf;
  hipLaunchKernelGGL(k_real_to_complex, dim3(ceil_div(nf, 256)), dim3(256), 0, ctx->stream, d_fine,
                     g, nf);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d_xyz(ctx, CHS_C128, g, fine_dims, 1, -1));
  hipLaunchKernelGGL(k_wave_gather, dim3(ceil_div(nw, 256)), dim3(256), 0, ctx->stream,
                     (const z64*)g, fine_dims[0], fine_dims[1], fine_dims[2], wave_dims[0],
                     wave_dims[1], wave_dims[2], w);
  CHS_LAUNCH_CHECK(ctx);
  CHS_TRY(fft3d_xyz(ctx, CHS_C128, w, wave_dims, 1, +1));
  hipLaunchKernelGGL(k_real_scaled, dim3(ceil_div(nw, 256)), dim3(256), 0, ctx->stream,
                     (const z64*)w, 1.0 / (double)nf, d_wave, nw);
  CHS_LAUNCH_CHECK(ctx);
  return CHS_OK;
}

}  // extern "C"
