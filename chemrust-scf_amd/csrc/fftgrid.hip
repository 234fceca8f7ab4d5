// Batched in-place 3D FFT of full n^3 grids (BatchedFftPlan3d,
// reference src/device/fft.rs:116-228), as three LDS line passes.
#include "fft.hpp"
#include "impl.hpp"

namespace chs {

constexpr int kLines = 16;

// One pass over axis `axis` (0 = x, 1 = y, 2 = z) of `batch` grids. A block
// transforms kLines lines: for x, kLines consecutive rows; for y and z,
// kLines consecutive ix at fixed (iz) resp. (iy), so global loads coalesce.
template <int N, int S, typename T>
__global__ void __launch_bounds__(256) k_fft_axis(cplx<T>* __restrict__ g, int axis,
                                                  const cplx<T>* __restrict__ tw) {
  constexpr int LS = N + 1;
  __shared__ cplx<T> lds[kLines * LS];
  const size_t n2 = (size_t)N * N, n3 = n2 * N;
  const int groups_per_plane = (N + kLines - 1) / kLines;
  const int b = blockIdx.y;
  cplx<T>* grid = g + (size_t)b * n3;
  size_t base, si, se;
  int nl = kLines;
  if (axis == 0) {
    const size_t row0 = (size_t)blockIdx.x * kLines;
    base = row0 * N;
    si = N;
    se = 1;
    if (row0 + nl > n2) nl = (int)(n2 - row0);
  } else {
    const int o = blockIdx.x / groups_per_plane;  // iz (axis 1) or iy (axis 2)
    const int ix0 = (blockIdx.x % groups_per_plane) * kLines;
    base = (axis == 1 ? (size_t)o * n2 : (size_t)o * N) + ix0;
    si = 1;
    se = axis == 1 ? N : n2;
    if (ix0 + nl > N) nl = N - ix0;
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
static int fft3d_impl(chs_context* ctx, int precision, T* grid, int n, int batch, int sign) {
  const void* tw = nullptr;
  CHS_TRY(twiddles(ctx, n, precision, &tw));
  using C = cplx<T>;
  CHS_DISPATCH_N(n, {
    const int groups = (NN + kLines - 1) / kLines;
    dim3 gx(ceil_div((size_t)NN * NN, kLines), batch), gyz(NN * groups, batch);
    for (int axis = 0; axis < 3; ++axis) {
      dim3 grd = axis == 0 ? gx : gyz;
      if (sign > 0)
        hipLaunchKernelGGL((k_fft_axis<NN, 1, T>), grd, dim3(256), 0, ctx->stream, (C*)grid,
                           axis, (const C*)tw);
      else
        hipLaunchKernelGGL((k_fft_axis<NN, -1, T>), grd, dim3(256), 0, ctx->stream, (C*)grid,
                           axis, (const C*)tw);
      CHS_LAUNCH_CHECK(ctx);
    }
  });
  return CHS_OK;
}

int fft3d(chs_context* ctx, int precision, void* grid, int n, int batch, int sign) {
  if (precision == CHS_C128) return fft3d_impl(ctx, precision, (double*)grid, n, batch, sign);
  return fft3d_impl(ctx, precision, (float*)grid, n, batch, sign);
}

}  // namespace chs

extern "C" int chs_fft3d_c2c(chs_context* ctx, int precision, void* d_grid, int n, int batch,
                             int sign) {
  if (!ctx || !d_grid || batch <= 0 || (sign != 1 && sign != -1) ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  if (!chs::fft_length_supported(n))
    return chs::set_error(ctx, CHS_ERR_FFT, "unsupported FFT length %d", n);
  return chs::fft3d(ctx, precision, d_grid, n, batch, sign);
}
