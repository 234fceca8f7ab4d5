// k-point sets: K plane-wave bindings of one cell (uneven plane-wave counts,
// own kinetic energies and projectors per k, one shared V_eff) whose
// Chebyshev filters run as ONE captured HIP graph: every k-point's degree-m
// recurrence (three H|psi> kernels, the projector GEMMs and the growth
// guard per step) is recorded once and replayed with a single
// hipGraphLaunch; the filter windows live in device memory so a replay
// takes each call's new bounds.
//
// Reference: the k-point loop of ScfIteration::diagonalize
// (src/scf.rs:169-265) calls chebyshev_filter (src/eigensolver/
// chebyshev.rs:626-816) once per k-point with per-call plan and buffer
// set-up; here all k-points share one set of workspaces and launches.
#include <vector>

#include "impl.hpp"
#include "ops.hpp"

namespace chs {

// Make k-point k (or nothing, k < 0) the bound Hamiltonian.
void kset_swap(chs_context* ctx, int k) {
  auto* im = ctx->impl;
  if (im->kset_cur >= 0) {
    std::swap(im->ham, im->kset[im->kset_cur]);
    im->kset_cur = -1;
  }
  if (k >= 0) {
    std::swap(im->ham, im->kset[k]);
    im->kset_cur = k;
  }
}

static void kset_drop_graph(chs_context* ctx) {
  auto* im = ctx->impl;
  if (im->kset_graph) (void)hipGraphExecDestroy(im->kset_graph);
  im->kset_graph = nullptr;
  im->kset_key.clear();
}

void kset_release(chs_context* ctx) {
  auto* im = ctx->impl;
  kset_swap(ctx, -1);
  kset_drop_graph(ctx);
  for (Hamiltonian& h : im->kset) release_hamiltonian(h);
  im->kset.clear();
}

// Enqueue every k-point's filter; each result ends in its caller buffer.
static int kset_enqueue(chs_context* ctx, int precision, void* const* d_psi, int n_b, int degree) {
  auto* im = ctx->impl;
  const int K = (int)im->kset.size();
  double* coef = (double*)im->kset_ws.p;
  double* flags = coef + 2 * K;
  const size_t cs = precision == CHS_C128 ? 16 : 8;
  for (int k = 0; k < K; ++k) {
    kset_swap(ctx, k);
    Hamiltonian& H = im->ham;
    const size_t bytes = (size_t)n_b * H.n_pw * cs;
    int rc = ensure(ctx, H.fscr, bytes);
    void* res = nullptr;
    if (rc == CHS_OK)
      rc = filter_steps(ctx, precision, d_psi[k], H.fscr.p, n_b, degree, coef + 2 * k,
                        flags + 8 * k, &res);
    if (rc == CHS_OK && res != d_psi[k]) {
      const hipError_t e =
          hipMemcpyAsync(d_psi[k], res, bytes, hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) rc = set_error(ctx, CHS_ERR_DEVICE, "k-set copy: %s", hipGetErrorString(e));
    }
    kset_swap(ctx, -1);
    CHS_TRY(rc);
  }
  return CHS_OK;
}

}  // namespace chs

using namespace chs;

extern "C" {

int chs_kset_bind(chs_context* ctx, int n_k, int n_grid, const int* n_pw,
                  const int32_t* const* d_miller, const double* const* d_kin,
                  const double* d_veff, int n_proj, const double* const* d_beta,
                  const double* d_d, const double* d_q) {
  if (!ctx || n_k <= 0 || !n_pw || !d_miller || !d_kin || !d_veff ||
      (n_proj > 0 && (!d_beta || !d_d || !d_q)))
    return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  kset_swap(ctx, -1);
  kset_drop_graph(ctx);
  for (Hamiltonian& h : im->kset) release_hamiltonian(h);
  im->kset.assign(n_k, Hamiltonian{});
  for (int k = 0; k < n_k; ++k) {
    kset_swap(ctx, k);
    int rc = bind_hamiltonian(ctx, n_grid, n_pw[k], d_miller[k], d_kin[k], d_veff);
    if (rc == CHS_OK && n_proj > 0) rc = chs_bind_projectors(ctx, n_proj, d_beta[k], d_d, d_q);
    kset_swap(ctx, -1);
    if (rc != CHS_OK) {
      for (Hamiltonian& h : im->kset) release_hamiltonian(h);
      im->kset.clear();
      return rc;
    }
  }
  return CHS_OK;
}

int chs_kset_set_veff(chs_context* ctx, const double* d_veff) {
  if (!ctx || !d_veff) return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  if (im->kset.empty()) return set_error(ctx, CHS_ERR_INVALID_ARG, "no k-point set bound");
  kset_swap(ctx, -1);
  for (Hamiltonian& h : im->kset) {
    if (h.veff != d_veff) kset_drop_graph(ctx);
    h.veff = d_veff;
    h.veff_f_valid = false;
  }
  return CHS_OK;
}

int chs_kset_select(chs_context* ctx, int k) {
  if (!ctx) return CHS_ERR_INVALID_ARG;
  if (k >= (int)ctx->impl->kset.size()) return set_error(ctx, CHS_ERR_INVALID_ARG, "k out of range");
  kset_swap(ctx, k);
  return CHS_OK;
}

int chs_kset_filter(chs_context* ctx, int precision, void* const* d_psi, int n_b, int degree,
                    const double* bounds4) {
  if (!ctx || !d_psi || n_b <= 0 || degree <= 0 || !bounds4 ||
      (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  const int K = (int)im->kset.size();
  if (K == 0) return set_error(ctx, CHS_ERR_INVALID_ARG, "no k-point set bound");
  for (int k = 0; k < K; ++k)
    if (!(bounds4[4 * k + 3] > 0.0) || !d_psi[k])
      return set_error(ctx, CHS_ERR_INVALID_ARG, "k-point %d: bad window or buffer", k);
  kset_swap(ctx, -1);
  // precision-dependent copies (FP32 V_eff, projectors) outside the graph
  for (int k = 0; k < K; ++k) {
    kset_swap(ctx, k);
    const int rc = ensure_precision_ready(ctx, precision);
    kset_swap(ctx, -1);
    CHS_TRY(rc);
  }
  CHS_TRY(ensure(ctx, im->kset_ws, (size_t)K * 10 * sizeof(double)));
  std::vector<double> coef(2 * (size_t)K);
  for (int k = 0; k < K; ++k) {
    coef[2 * k] = bounds4[4 * k + 2];
    coef[2 * k + 1] = 1.0 / bounds4[4 * k + 3];
  }
  double* d_coef = (double*)im->kset_ws.p;
  double* d_flags = d_coef + 2 * K;
  CHS_HIP(ctx, hipMemcpyAsync(d_coef, coef.data(), coef.size() * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  CHS_HIP(ctx, hipMemsetAsync(d_flags, 0, 8 * (size_t)K * sizeof(double), ctx->stream));
  ctx->h2d += coef.size() * sizeof(double);
  std::vector<uintptr_t> key = {(uintptr_t)K, (uintptr_t)n_b, (uintptr_t)degree,
                                (uintptr_t)precision, (uintptr_t)im->tune_gen};
  for (int k = 0; k < K; ++k) key.push_back((uintptr_t)d_psi[k]);
  if (im->use_graphs && im->kset_graph && key == im->kset_key &&
      im->kset_key_gen == im->ws_gen) {
    CHS_HIP(ctx, hipGraphLaunch(im->kset_graph, ctx->stream));
  } else {
    // eager run (allocates what the graph will name), then capture for the
    // next calls; the capture itself does not execute
    CHS_TRY(kset_enqueue(ctx, precision, d_psi, n_b, degree));
    if (im->use_graphs) {
      kset_drop_graph(ctx);
      CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const uint64_t gen = im->ws_gen;
      hipGraphExec_t exec = nullptr;
      CHS_TRY(capture_graph(
          ctx,
          [&](chs_context* c) {
            const int rc = kset_enqueue(c, precision, d_psi, n_b, degree);
            if (rc == CHS_OK && im->ws_gen != gen)
              return set_error(c, CHS_ERR_DEVICE, "k-set capture allocated workspace");
            return rc;
          },
          &exec));
      im->kset_graph = exec;
      im->kset_key = key;
      im->kset_key_gen = im->ws_gen;
    }
  }
  // divergence report of the first k-point that diverged
  std::vector<double> hf(8 * (size_t)K);
  CHS_HIP(ctx, hipMemcpyAsync(hf.data(), d_flags, hf.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  CHS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->d2h += hf.size() * sizeof(double);
  for (int k = 0; k < K; ++k) {
    const double* f = hf.data() + 8 * k;
    if (f[0] != 0.0) {
      ctx->diag.iteration = (int)f[0];
      ctx->diag.norm_previous = f[1];
      ctx->diag.norm_current = f[2];
      ctx->diag.growth_factor = f[3];
      return set_error(ctx, CHS_ERR_CHEBYSHEV_DIVERGED,
                       "k-point %d: Chebyshev filter diverged at iteration %d (x%.3f)", k,
                       (int)f[0], f[3]);
    }
  }
  return CHS_OK;
}

int chs_autotune_hpsi(chs_context* ctx, int precision, int n_b, int* sticks_per_block,
                      int* row_div, int* plane_threads, double* ms) {
  if (!ctx || n_b <= 0 || (precision != CHS_C128 && precision != CHS_C64))
    return CHS_ERR_INVALID_ARG;
  auto* im = ctx->impl;
  if (!im->ham.bound) return set_error(ctx, CHS_ERR_INVALID_ARG, "no Hamiltonian bound");
  CHS_TRY(ensure_precision_ready(ctx, precision));
  const size_t bytes = (size_t)n_b * im->ham.n_pw * (precision == CHS_C128 ? 16 : 8);
  CHS_TRY(ensure(ctx, im->ws1, bytes));
  CHS_TRY(ensure(ctx, im->ws2, bytes));
  CHS_HIP(ctx, hipMemsetAsync(im->ws1.p, 0, bytes, ctx->stream));
  CHS_HIP(ctx, hipMemsetAsync(im->ws2.p, 0, bytes, ctx->stream));
  VlocTune best;
  double t = 0;
  CHS_TRY(vloc_autotune(ctx, precision, im->ws1.p, n_b, im->ws2.p, 5, &best, &t));
  im->tune_gen += 1;
  if (sticks_per_block) *sticks_per_block = best.sb;
  if (row_div) *row_div = best.nrb_div;
  if (plane_threads) *plane_threads = best.threads;
  if (ms) *ms = t;
  return CHS_OK;
}

}  // extern "C"
