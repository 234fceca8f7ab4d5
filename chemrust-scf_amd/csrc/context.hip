// Context, memory, transfer accounting and status reporting.
#include <cstdlib>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <dlfcn.h>

#include <fstream>
#include <mutex>
#include <sstream>

#include "impl.hpp"

namespace chs {

// ---- on-disk tuning cache ------------------------------------------------
// Header "chs-tune 1 <library hash> <device>": entries count only for the
// library build (FNV-1a of the loaded .so) and device (architecture and CU
// count) that wrote them; another header means the file is rewritten.
static std::mutex g_tune_mu;

static std::string library_hash() {
  static std::string h;
  static std::once_flag once;
  std::call_once(once, [] {
    Dl_info info{};
    uint64_t f = 1469598103934665603ull;
    if (dladdr((void*)&chs_context_create, &info) && info.dli_fname) {
      std::ifstream in(info.dli_fname, std::ios::binary);
      char buf[1 << 16];
      while (in.read(buf, sizeof(buf)) || in.gcount() > 0) {
        for (std::streamsize i = 0; i < in.gcount(); ++i) {
          f ^= (unsigned char)buf[i];
          f *= 1099511628211ull;
        }
      }
    }
    char hex[17];
    snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)f);
    h = hex;
  });
  return h;
}

static std::string tune_header(chs_context* ctx) {
  hipDeviceProp_t prop{};
  std::string dev = "unknown";
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess)
    dev = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
  for (char& c : dev)
    if (c == ' ') c = '_';
  return "chs-tune 1 " + library_hash() + " " + dev;
}

void tune_cache_load(chs_context* ctx) {
  auto* im = ctx->impl;
  if (im->tune_path.empty()) return;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  std::ifstream in(im->tune_path);
  std::string line;
  if (!in || !std::getline(in, line) || line != tune_header(ctx)) return;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string kind;
    ss >> kind;
    if (kind == "vloc") {
      int n, prec, sb, div, thr;
      if (ss >> n >> prec >> sb >> div >> thr) im->vloc_tune[{n, prec}] = VlocTune{sb, div, thr};
    } else if (kind == "gemm") {
      GemmKey k{};
      int choice;
      if (ss >> k.tsize >> k.oa >> k.ob >> k.m >> k.n >> k.k >> choice) im->gemm_tune[k] = choice;
    } else if (kind == "dens") {
      int key, shape;
      if (ss >> key >> shape) im->density_tpb[key] = shape;
    }
  }
}

void tune_cache_store(chs_context* ctx, const std::string& line) {
  auto* im = ctx->impl;
  if (im->tune_path.empty()) return;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const std::string head = tune_header(ctx);
  std::string first;
  {
    std::ifstream in(im->tune_path);
    if (in) std::getline(in, first);
  }
  if (first != head) {
    std::ofstream out(im->tune_path, std::ios::trunc);
    out << head << "\n";
  }
  std::ofstream out(im->tune_path, std::ios::app);
  out << line << "\n";
}

int set_error(chs_context* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->last_error = buf;
  return code;
}

int ensure(chs_context* ctx, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return CHS_OK;
  if (b.p) {
    CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CHS_HIP(ctx, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  CHS_HIP(ctx, hipMalloc(&b.p, want));
  b.bytes = want;
  ctx->impl->ws_gen += 1;
  return CHS_OK;
}

void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int twiddles(chs_context* ctx, int n, int precision, const void** out) {
  auto key = std::make_pair(n, precision);
  auto it = ctx->impl->twiddles.find(key);
  if (it != ctx->impl->twiddles.end()) {
    *out = it->second;
    return CHS_OK;
  }
  void* d = nullptr;
  if (precision == CHS_C128) {
    std::vector<double> h(2 * n);
    for (int t = 0; t < n; ++t) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / n;
      h[2 * t] = (double)cosl(a);
      h[2 * t + 1] = (double)sinl(a);
    }
    CHS_HIP(ctx, hipMalloc(&d, h.size() * sizeof(double)));
    CHS_HIP(ctx, hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->h2d += h.size() * sizeof(double);
  } else {
    std::vector<float> h(2 * n);
    for (int t = 0; t < n; ++t) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / n;
      h[2 * t] = (float)cosl(a);
      h[2 * t + 1] = (float)sinl(a);
    }
    CHS_HIP(ctx, hipMalloc(&d, h.size() * sizeof(float)));
    CHS_HIP(ctx, hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->h2d += h.size() * sizeof(float);
  }
  ctx->impl->twiddles[key] = d;
  *out = d;
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_context_create(int device, chs_context** out) {
  if (!out) return CHS_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CHS_ERR_DEVICE;
  if (device < 0 || device >= n) return CHS_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return CHS_ERR_DEVICE;
  auto* ctx = new chs_context;
  ctx->device = device;
  ctx->impl = new chs_context::Impl;
  if (const char* ng = std::getenv("CHS_NO_GRAPHS")) ctx->impl->use_graphs = !(ng[0] && ng[0] != '0');
  if (const char* gs = std::getenv("CHS_GEMM_SIMT")) ctx->impl->use_mfma = !(gs[0] && gs[0] != '0');
  if (const char* gt = std::getenv("CHS_GEMM_TUNE")) ctx->impl->gemm_autotune = !(gt[0] == '0');
  if (const char* gc = std::getenv("CHS_GEMM_CHOICE")) ctx->impl->gemm_choice = std::atoi(gc);
  if (const char* ht = std::getenv("CHS_HPSI_TUNE")) ctx->impl->hpsi_autotune = !(ht[0] == '0');
  if (const char* pw = std::getenv("CHS_PLANE_WINDOW")) ctx->impl->plane_window = !(pw[0] == '0');
  if (const char* g3 = std::getenv("CHS_GEMM_FP16X3")) ctx->impl->use_h3 = g3[0] && g3[0] != '0';
  if (const char* tc = std::getenv("CHS_TUNE_CACHE")) ctx->impl->tune_path = tc;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->impl->ev0) != hipSuccess ||
      hipEventCreate(&ctx->impl->ev1) != hipSuccess) {
    delete ctx->impl;
    delete ctx;
    return CHS_ERR_DEVICE;
  }
  tune_cache_load(ctx);
  *out = ctx;
  return CHS_OK;
}

int chs_context_destroy(chs_context* ctx) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  auto* im = ctx->impl;
  for (chs_scf* s : im->scf_children) scf_orphan(s);
  im->scf_children.clear();
  for (auto& kv : im->twiddles) (void)hipFree(kv.second);
  DevBuf* bufs[] = {&im->stick_buf, &im->part, &im->small, &im->ws1, &im->ws2, &im->ws3,
                    &im->ws4, &im->ws5, &im->ws6, &im->rr32, &im->res64, &im->res32, &im->res_br, &im->res_tab, &im->lanczos_v0, &im->gemm_ws, &im->gemv_ws, &im->h3_bws, &im->lin_ws, &im->scf_ws,
                    &im->proj_ws, &im->chol_ws, &im->eig_ws, &im->mlp_ws, &im->fft_ws, &im->kset_ws,
                    &im->gemm_tune_ws, &im->gemm_tune_c};
  for (DevBuf* b : bufs) release(*b);
  release_hamiltonian(im->ham);
  for (Hamiltonian& h : im->kset) release_hamiltonian(h);
  if (im->kset_graph) (void)hipGraphExecDestroy(im->kset_graph);
  if (im->eigh_graph.exec) (void)hipGraphExecDestroy(im->eigh_graph.exec);
  if (im->jacobi.exec) (void)hipGraphExecDestroy(im->jacobi.exec);
  release(im->jacobi.sched);
  release(im->jacobi.rot);
  (void)hipEventDestroy(im->ev0);
  (void)hipEventDestroy(im->ev1);
  (void)hipStreamDestroy(ctx->stream);
  delete im;
  delete ctx;
  return CHS_OK;
}

int chs_synchronize(chs_context* ctx) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CHS_OK;
}

const char* chs_last_error(chs_context* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int chs_diag_get(chs_context* ctx, chs_diag* out) {
  if (!ctx || !out) return CHS_ERR_INVALID_ARG;
  *out = ctx->diag;
  return CHS_OK;
}

int chs_transfer_bytes(chs_context* ctx, uint64_t* h2d, uint64_t* d2h) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  if (h2d) *h2d = ctx->h2d;
  if (d2h) *d2h = ctx->d2h;
  return CHS_OK;
}

int chs_tuning_entries(chs_context* ctx, int* n_hpsi, int* n_gemm, int* n_density) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  if (n_hpsi) *n_hpsi = (int)ctx->impl->vloc_tune.size();
  if (n_gemm) *n_gemm = (int)ctx->impl->gemm_tune.size();
  if (n_density) *n_density = (int)ctx->impl->density_tpb.size();
  return CHS_OK;
}

int chs_transfer_reset(chs_context* ctx) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  ctx->h2d = ctx->d2h = 0;
  return CHS_OK;
}

int chs_alloc(chs_context* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return CHS_ERR_INVALID_ARG;
  *dptr = nullptr;
  if (bytes == 0) bytes = 1;
  CHS_HIP(ctx, hipMalloc(dptr, bytes));
  return CHS_OK;
}

int chs_free(chs_context* ctx, void* dptr) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  if (!dptr) return CHS_OK;
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CHS_HIP(ctx, hipFree(dptr));
  return CHS_OK;
}

int chs_h2d(chs_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return CHS_ERR_INVALID_ARG;
  if (!bytes) return CHS_OK;
  CHS_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->h2d += bytes;
  return CHS_OK;
}

int chs_d2h(chs_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return CHS_ERR_INVALID_ARG;
  if (!bytes) return CHS_OK;
  CHS_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += bytes;
  return CHS_OK;
}

int chs_d2d(chs_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return CHS_ERR_INVALID_ARG;
  if (!bytes) return CHS_OK;
  CHS_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return CHS_OK;
}

int chs_memset(chs_context* ctx, void* dptr, int value, size_t bytes) {
  if (!ctx || (!dptr && bytes)) return CHS_ERR_INVALID_ARG;
  if (!bytes) return CHS_OK;
  CHS_HIP(ctx, hipMemsetAsync(dptr, value, bytes, ctx->stream));
  return CHS_OK;
}

int chs_spectral_bounds(const double* eig, int n_eig, double lambda_max, double* out4) {
  // src/eigensolver/chebyshev.rs:214-254
  if (!out4 || (n_eig > 0 && !eig)) return CHS_ERR_INVALID_ARG;
  double eps_cut;
  if (n_eig <= 0) {
    eps_cut = lambda_max / 3.0;
  } else {
    const double e0 = eig[0], el = eig[n_eig - 1];
    const double raw = el + 0.2 * (el - e0);
    eps_cut = raw < lambda_max * 0.95 ? raw : lambda_max * 0.95;
  }
  out4[0] = lambda_max;
  out4[1] = eps_cut;
  out4[2] = 0.5 * (lambda_max + eps_cut);
  out4[3] = 0.5 * (lambda_max - eps_cut);
  return CHS_OK;
}

}  // extern "C"
