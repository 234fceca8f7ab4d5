// Shared device/host definitions for libchs_gpu (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/chs.h"

namespace chs {

// ---------------------------------------------------------------------------
// Complex arithmetic on interleaved pairs.
// ---------------------------------------------------------------------------
template <typename T> struct cplx {
  T x, y;
};
// complex64 on its natural 8-byte alignment, so that an element moves with
// one 64-bit LDS access (ds_read_b64 / ds_write_b64) rather than a pair of
// 32-bit ones
template <> struct alignas(8) cplx<float> {
  float x, y;
};
using z64 = cplx<double>;
using c32 = cplx<float>;

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) {
  return mk(s * a.x, s * a.y);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) {
  return mk(a.x, -a.y);
}
// conj(a) * b
template <typename T> __host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename T> __host__ __device__ __forceinline__ T norm2(cplx<T> a) {
  return a.x * a.x + a.y * a.y;
}

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
struct Status {
  int code = CHS_OK;
  std::string msg;
};

}  // namespace chs

// Context: one HIP stream, transfer counters, cached workspaces.
struct chs_context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t h2d = 0, d2h = 0;
  std::string last_error;
  chs_diag diag{};
  struct Impl;
  Impl* impl = nullptr;
};

namespace chs {

int set_error(chs_context* ctx, int code, const char* fmt, ...);

#define CHS_HIP(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      return ::chs::set_error((ctx), _e == hipErrorOutOfMemory ? CHS_ERR_OUT_OF_MEMORY  \
                                                               : CHS_ERR_DEVICE,        \
                              "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),    \
                              __FILE__, __LINE__);                                      \
  } while (0)

#define CHS_TRY(expr)            \
  do {                           \
    int _rc = (expr);            \
    if (_rc != CHS_OK) return _rc; \
  } while (0)

#define CHS_LAUNCH_CHECK(ctx) CHS_HIP(ctx, hipGetLastError())

// Device buffer owned by the context (reused, grown on demand).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
int ensure(chs_context* ctx, DevBuf& b, size_t bytes);
void release(DevBuf& b);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace chs
