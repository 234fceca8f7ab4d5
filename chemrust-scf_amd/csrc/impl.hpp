// Context internals shared by the translation units of libchs_gpu.
#pragma once
#include <map>
#include <string>
#include <vector>
#include <utility>

#include "common.hpp"

namespace chs {

// Column ("stick") decomposition of the plane-wave sphere: one stick per
// (ix, iy) grid column that holds at least one plane wave. Plane waves are
// stored lexicographically in (h, k, l), so the plane waves of a stick are a
// contiguous index range with consecutive l.
struct StickTable {
  int n_st = 0;    // sticks
  int n_act = 0;   // distinct ix among sticks ("active" x columns)
  int n_grid = 0;
  int n_pw = 0;
  // device arrays (int32), length n_st unless noted
  int32_t* iy = nullptr;     // FFT y index of the stick
  int32_t* act = nullptr;    // index of the stick's ix in the active list
  int32_t* pw0 = nullptr;    // first plane wave of the stick
  int32_t* cnt = nullptr;    // plane waves in the stick
  int32_t* iz0 = nullptr;    // FFT z index of the first plane wave
  int32_t* act_ix = nullptr; // n_act: ix of each active column
  int32_t* act_of = nullptr; // n_grid: active index of each x column, or -1
  int32_t* pw_siz = nullptr; // n_pw: (stick << 8) | FFT z index of each plane wave
  std::vector<int32_t> h_act_ix;
  // compact-plane row stride of the plane kernels that skip the zero window
  // (Window<N> in fft.hpp), or 0 when some stick or active column lies in it
  int rs_win = 0;
};

// Six fp16 planes [m][k] of a fixed complex operand for the fp16x3 GEMM
// (ar, ai, ar + ai, each as hi and lo of the value scaled by 2^ea); rows
// padded to m_pad, row pitch ld (both with zero padding).
struct H3Planes {
  DevBuf buf;
  int m = 0, k = 0, m_pad = 0, ld = 0, ea = 0;
  size_t plane = 0;  // halfs per plane
};

struct Hamiltonian {
  bool bound = false;
  int n_grid = 0, n_pw = 0;
  const int32_t* miller = nullptr;
  const double* kin = nullptr;
  const double* veff = nullptr;
  DevBuf veff_f;      // float copy of V_eff for the c64 path
  bool veff_f_valid = false;
  StickTable st;
  DevBuf st_mem;
  // projectors
  int n_proj = 0;
  const double* beta = nullptr;  // c128, P x n_pw
  const double* dion = nullptr;  // P
  const double* qion = nullptr;  // P
  bool has_s = false;
  DevBuf beta_f;                 // c64 copy
  DevBuf sinv_m;                 // c128 P x P (Q^-1 + beta^H beta)^-1
  DevBuf sinv_m_f;               // c64 copy
  DevBuf bmat;                   // c128 P x P beta^H beta
  DevBuf bmat_f;                 // c64 copy
  // projector coefficients carried through a Chebyshev filter (nl_ws, five
  // P x n_b slots): slots of W_{k-1}, W_k and G = beta^H O; 3 and 4 scratch
  DevBuf nl_ws;
  int nl_wp = 0, nl_w = 1, nl_g = 2;
  DevBuf fscr;                   // second Chebyshev buffer of a k-point set member
  // fp16x3 split planes of beta^H and beta (complex64 filter), and the
  // per-column exponents of the two products' right operands (2 n_b ints)
  H3Planes h3_cn, h3_nn;
  DevBuf h3_exp;
};

inline void release_hamiltonian(Hamiltonian& h) {
  for (DevBuf* b : {&h.veff_f, &h.st_mem, &h.beta_f, &h.sinv_m, &h.sinv_m_f, &h.bmat, &h.bmat_f,
                    &h.nl_ws, &h.fscr, &h.h3_cn.buf, &h.h3_nn.buf, &h.h3_exp})
    release(*b);
  h = Hamiltonian{};
}

// Forget the caller arrays the bound Hamiltonian points at (plane waves,
// V_eff, projectors); its own buffers stay allocated for the next binding.
inline void unbind_hamiltonian_fields(Hamiltonian& h) {
  h.bound = false;
  h.miller = nullptr;
  h.kin = nullptr;
  h.veff = nullptr;
  h.veff_f_valid = false;
  h.n_proj = 0;
  h.beta = h.dion = h.qion = nullptr;
  h.has_s = false;
}

// Captured launch sequence of the Householder tridiagonal reduction, valid
// for one (n, matrix, workspace) triple.
struct EighGraph {
  int n = 0;
  const void* gA = nullptr;
  const void* gws = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Round-robin pair schedule and rotation buffer of the Jacobi eigensolver,
// with one captured sweep for the last (matrix, vectors) pair.
struct JacobiPlan {
  int n = 0;
  int rounds = 0, pairs = 0;
  DevBuf sched;  // int2 [rounds][pairs]
  DevBuf rot;    // double4 [pairs]
  hipGraphExec_t exec = nullptr;
  const void* gA = nullptr;
  const void* gV = nullptr;
};

// GEMM shape key of the kernel-variant cache.
struct GemmKey {
  int tsize, oa, ob, m, n, k;
  bool operator<(const GemmKey& o) const {
    const int a[6] = {tsize, oa, ob, m, n, k}, b[6] = {o.tsize, o.oa, o.ob, o.m, o.n, o.k};
    for (int i = 0; i < 6; ++i)
      if (a[i] != b[i]) return a[i] < b[i];
    return false;
  }
};

// Launch shape of the H|psi> pipeline: sticks per block of the z passes, the
// divisor of the plane kernel's row-buffer count and its workgroup size.
struct VlocTune {
  int sb = 32;
  int nrb_div = 1;
  int threads = 256;  // plane kernel workgroup size (256 or 512)
};

}  // namespace chs

struct chs_scf;  // a caller-supplied SCF system (scf_driver.hip)

struct chs_context::Impl {
  std::map<std::pair<int, int>, void*> twiddles;  // (N, precision) -> table
  std::map<std::pair<int, int>, chs::VlocTune> vloc_tune;  // (N, precision) -> launch shape
  std::map<int, int> density_tpb;  // 2 N + (complex64) -> density-plane launch shape
  bool hpsi_autotune = true;  // chs_scf_run tunes new grids (CHS_HPSI_TUNE=0: off)
  // plane kernels that skip the zero window of each line when the stick
  // table allows (CHS_PLANE_WINDOW=0: the general kernels throughout)
  bool plane_window = true;
  chs::Hamiltonian ham;
  // workspaces
  chs::DevBuf stick_buf, part, small, ws1, ws2, ws3, ws4, ws5, ws6;
  chs::DevBuf rr32;  // complex64 H|Y> and beta^H Y of Rayleigh-Ritz in the complex64 phase
  // residual filter: FP64 residual, its complex64 copy, its projector
  // coefficients, per-band tables (+ guard flags and window)
  chs::DevBuf res64, res32, res_br, res_tab;
  // Lanczos start vector of the last (seed, n_pw): generated on the host once
  chs::DevBuf lanczos_v0;
  uint64_t lanczos_seed = 0;
  int lanczos_n = 0;
  chs::DevBuf gemm_ws, gemv_ws, h3_bws, lin_ws, scf_ws, proj_ws, chol_ws, eig_ws, mlp_ws, fft_ws;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Replay launch-bound loops (Jacobi sweeps) as captured HIP graphs. Off
  // when CHS_NO_GRAPHS is set to a non-zero value: some kernel tracers
  // cannot follow graph capture, and direct launches run the same kernels.
  bool use_graphs = true;
  // Complex GEMMs on the matrix cores (MFMA); CHS_GEMM_SIMT=1 selects the
  // vector-FMA kernel of the same tiling instead.
  bool use_mfma = true;
  // complex64 filter projector products on the fp16 matrix cores (fp16x3);
  // opt-in: its split operands carry 1.5x the bytes of FP32 and the kernel is
  // slower than the FP32 form
  bool use_h3 = false;
  // per-shape GEMM kernel variant, timed on first use (CHS_GEMM_TUNE=0: off)
  bool gemm_autotune = true;
  int gemm_choice = -1;  // CHS_GEMM_CHOICE: one fixed kernel choice for every shape (>= 0)
  std::map<chs::GemmKey, int> gemm_tune;
  chs::DevBuf gemm_tune_ws, gemm_tune_c;
  // counts workspace (re)allocations: a captured graph is valid only while
  // the buffers it names are the ones it was captured with
  uint64_t ws_gen = 0;
  uint64_t tune_gen = 0;
  // k-point set: member k lives in kset[k] except while swapped into `ham`
  // (kset_cur == k; kset[k] then holds the displaced binding)
  std::vector<chs::Hamiltonian> kset;
  int kset_cur = -1;
  chs::DevBuf kset_ws;          // per-k filter windows {c, inv_e} and flags
  hipGraphExec_t kset_graph = nullptr;
  chs::EighGraph eigh_graph;  // captured tridiagonal reduction of the last (n, A, workspace)
  chs::JacobiPlan jacobi;     // Jacobi eigensolver schedule (CHS_EIGH_JACOBI)
  std::vector<uintptr_t> kset_key;
  uint64_t kset_key_gen = 0;
  // SCF systems created on this context: destroying the context frees their
  // device arrays and detaches them, so a later chs_scf_context_destroy only
  // frees the host side
  std::vector<chs_scf*> scf_children;
  // on-disk tuning cache (CHS_TUNE_CACHE=<path>): launch shapes and GEMM
  // variants keyed by shape, precision, library build and device
  std::string tune_path;
};

namespace chs {

// Length-N root table e^{+2 pi i t/N} (t < N) in the given precision.
int twiddles(chs_context* ctx, int n, int precision, const void** out);
// Batched in-place full-grid 3D FFT (fftgrid.hip): n^3 grids, or
// dims[0] x dims[1] x dims[2] (x fastest).
int fft3d(chs_context* ctx, int precision, void* grid, int n, int batch, int sign);
int fft3d_xyz(chs_context* ctx, int precision, void* grid, const int dims[3], int batch,
              int sign);
// Stick tables for the bound plane-wave set.
int build_sticks(chs_context* ctx);
// Bind a plane-wave set + V_eff as the current Hamiltonian (no k-set swap).
int bind_hamiltonian(chs_context* ctx, int n_grid, int n_pw, const int32_t* d_miller,
                     const double* d_kin, const double* d_veff);
// Make k-point k of the bound set the current Hamiltonian (k < 0: none).
void kset_swap(chs_context* ctx, int k);
// Tuning cache: load the entries of this build and device (context set-up)
// and append one decision ("vloc n prec sb div threads", "gemm tsize oa ob m
// n k choice", "dens key tpb").
void tune_cache_load(chs_context* ctx);
void tune_cache_store(chs_context* ctx, const std::string& line);

// Drop the k-point set (its bindings, graph and per-k buffers).
void kset_release(chs_context* ctx);
// Free an SCF system's device arrays and detach it from its (dying) context.
void scf_orphan(chs_scf* s);
// Unbind the current Hamiltonian (a selected k-point first goes back to its
// set): the operator calls then report that nothing is bound.
inline void unbind_hamiltonian(chs_context* ctx) {
  kset_swap(ctx, -1);
  unbind_hamiltonian_fields(ctx->impl->ham);
}

// Local-potential H|psi> pipeline pieces (hamiltonian.hip).
// d_coef: device {c, inv_e} read by the kernels instead of c, inv_e (so a
// captured graph replays with new bounds); nullptr to use the values.
// EPI_CHEBK_RES also adds res_s[b] * res_w (band b's residual, filter
// precision; res_s: per-band scales, device f64).
int vloc_pipeline(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int mode,
                  const void* aux0, const double* d_coef, double c, double inv_e,
                  double* d_norm_part, int* n_part, const void* res_w = nullptr,
                  const double* res_s = nullptr);
int vloc_autotune(chs_context* ctx, int precision, const void* psi, int n_b, void* out, int reps,
                  VlocTune* best, double* best_ms);

// Records the launches `enqueue(ctx)` makes on the context stream into an
// executable graph. Every exit leaves the stream out of capture mode and frees
// the intermediate graph; *exec is written only when instantiation succeeded
// (a failed instantiate leaves no handle that a later call could launch or
// destroy).
template <typename F>
int capture_graph(chs_context* ctx, F&& enqueue, hipGraphExec_t* exec) {
  CHS_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue(ctx);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != CHS_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != CHS_OK) return rc;
    return set_error(ctx, CHS_ERR_DEVICE, "stream capture failed: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t e = nullptr;
  const hipError_t ie = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess)
    return set_error(ctx, CHS_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ie));
  *exec = e;
  return CHS_OK;
}

// Modes of the stick epilogue.
enum EpiMode {
  EPI_HPSI = 0,    // out = T psi + V psi
  EPI_CHEB1 = 1,   // out = (H psi - c psi) * inv_e
  EPI_CHEBK = 2,   // out = 2 (H psi - c psi) * inv_e - aux0
  EPI_CHEBK_RES = 3,  // EPI_CHEBK + res_s[b] res_w (residual-based filter)
};

}  // namespace chs
