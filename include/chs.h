/* chs.h -- C ABI of libchs_gpu.so, the MI355X (gfx950) implementation of the
 * chemrust-scf SCF eigensolver hot path.
 *
 * Every entry point replaces a piece of the reference's device layer or
 * eigensolver surface; the reference file:line each one stands in for is
 * given beside it (paths relative to the reference repository root).
 *
 * Conventions (identical to the reference):
 *  - complex numbers are interleaved (re, im): double pairs for CHS_C128,
 *    float pairs for CHS_C64 (src/device/mod.rs:24-43);
 *  - a wavefunction block is band-major ("ColumnDistributed",
 *    src/layout.rs:48-53): psi[b * n_pw + g], i.e. an (n_pw x n_b)
 *    column-major matrix with leading dimension n_pw;
 *  - FFT grids are ix-fastest: idx = ix + ngx * (iy + ngy * iz)
 *    (src/scf.rs:385-399); the inverse transform is e^{+iG.r}, the forward
 *    e^{-iG.r}, both unnormalised (src/eigensolver/chebyshev.rs:331,346);
 *  - eigenvalues ascend (ZHEGVD, src/device/solver.rs:47-95).
 *
 * Every function returns an int32 status: CHS_OK (0) or a negative code, one
 * per variant of the reference's Error enum (src/types.rs:225-248) plus
 * argument and allocation errors. Nothing is thrown or aborts across the
 * ABI. chs_last_error() gives a message; chs_diag() the structured fields
 * (ChebyshevDiverged, solver info, failing Cholesky column).
 *
 * Threading: one context owns one HIP stream; calls on a context are
 * stream-ordered and must not be made from two threads at once. Separate
 * contexts are independent (src/scf.rs:173-174, CONTEXT.md:160).
 * Device buffers passed in are owned by the caller; the library never frees
 * them. Workspaces are owned by the context and reused across calls. */
#ifndef CHS_H
#define CHS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (src/types.rs:225-248) ---------------------------- */
enum {
  CHS_OK = 0,
  CHS_ERR_NOT_IMPLEMENTED = -1,      /* Error::NotImplemented */
  CHS_ERR_CHEBYSHEV_DIVERGED = -2,   /* Error::ChebyshevDiverged{..} */
  CHS_ERR_RAYLEIGH_RITZ_FAILED = -3, /* Error::RayleighRitzFailed{info} */
  CHS_ERR_SOLVER = -4,               /* Error::Solver */
  CHS_ERR_BLAS = -5,                 /* Error::Blas */
  CHS_ERR_FFT = -6,                  /* Error::Fft */
  CHS_ERR_DEVICE = -7,               /* Error::Cuda (HIP runtime error) */
  CHS_ERR_KERNEL_BUILD = -8,         /* Error::Nvrtc (no runtime compile here: unused) */
  CHS_ERR_INVALID_ARG = -9,
  CHS_ERR_OUT_OF_MEMORY = -10,
  CHS_ERR_NOT_CONVERGED = -11        /* SCF hit max_iter */
};

/* precision of a compute call */
enum { CHS_C128 = 0, CHS_C64 = 1 };

typedef struct chs_diag {
  int32_t iteration;          /* ChebyshevDiverged.iteration */
  double norm_previous;       /* ChebyshevDiverged.norm_previous */
  double norm_current;        /* ChebyshevDiverged.norm_current */
  double growth_factor;       /* ChebyshevDiverged.growth_factor */
  int32_t solver_info;        /* RayleighRitzFailed.info / Cholesky column+1 */
} chs_diag;

typedef struct chs_context chs_context;

/* ---- context (src/scf.rs:173-177: CudaContext::new(0), default stream,
 * BLAS/solver handles, CudaKernelSet::new) ------------------------------ */
int chs_context_create(int device, chs_context** out);
int chs_context_destroy(chs_context* ctx);
int chs_synchronize(chs_context* ctx);
const char* chs_last_error(chs_context* ctx);
int chs_diag_get(chs_context* ctx, chs_diag* out);
/* PcieAccount (src/device/pcie.rs:21-70): bytes moved host<->device by this
 * context since creation or the last reset. */
int chs_transfer_bytes(chs_context* ctx, uint64_t* h2d, uint64_t* d2h);
int chs_transfer_reset(chs_context* ctx);
/* Tuning decisions the context holds: H|psi> launch shapes (per grid and
 * precision), GEMM kernel variants (per shape) and density-plane workgroup
 * sizes. With CHS_TUNE_CACHE=<path> set when the context is created, the
 * decisions of earlier processes for this library build and device are read
 * from that file at creation and every new one is appended to it (no
 * counterpart in the reference: its plans are rebuilt per call, D8). */
int chs_tuning_entries(chs_context* ctx, int* n_hpsi, int* n_gemm, int* n_density);

/* ---- memory (Gpu<T>::from_host / sync_to_host, src/device/mod.rs:216-265;
 * Gpu::from_host_with / sync_to_host_with, src/device/pcie.rs:44-70) ---- */
int chs_alloc(chs_context* ctx, size_t bytes, void** dptr);
int chs_free(chs_context* ctx, void* dptr);
int chs_h2d(chs_context* ctx, void* dst, const void* src, size_t bytes);
int chs_d2h(chs_context* ctx, void* dst, const void* src, size_t bytes);
int chs_d2d(chs_context* ctx, void* dst, const void* src, size_t bytes);
int chs_memset(chs_context* ctx, void* dptr, int value, size_t bytes);

/* ---- setup ------------------------------------------------------------ */
/* pw_coords_to_fft_indices (src/scf.rs:387-401): Miller triples [h,k,l]
 * (device int32, 3*n_pw) to ix + ngx*(iy + ngy*iz) with negative indices
 * wrapped by +n. Bit-exact. */
int chs_pw_coords_to_fft_indices(chs_context* ctx, const int32_t* d_miller, int n_pw, int ngx,
                                 int ngy, int ngz, int32_t* d_idx);
/* Plane-wave sphere 0.5|k+G|^2 <= ecut of a cubic cell of edge L on an
 * n^3 grid (the reference takes it from GVectorGrid/KptWaveBlock of its
 * chemrust-hamiltonian-core dependency; this library defines the order:
 * lexicographic (h, k, l), h slowest). With d_miller == NULL only *n_pw is
 * returned; otherwise capacity max_pw must hold. Outputs: Miller triples,
 * FFT indices, 0.5|k+G|^2. */
int chs_gsphere(chs_context* ctx, double L, double ecut, const double kfrac[3], int n_grid,
                int max_pw, int* n_pw, int32_t* d_miller, int32_t* d_idx, double* d_kin);

/* ---- 3D FFT (BatchedFftPlan3d, src/device/fft.rs:116-228) -----------
 * plan_batched_c2c(nx, ny, nz, batch) + c2c_forward / c2c_inverse
 * (:181-228): batched, in place, unnormalised, nx x ny x nz grids (x
 * fastest) stored back to back; sign = +1 inverse (e^{+iG.r}), -1 forward.
 * Each length one of 8, 12, 16, 18, 24, 32, 36, 48, 64, 72, 96, 108, 128,
 * 144 (else CHS_ERR_FFT); batch <= 65535. */
int chs_fft3d_c2c(chs_context* ctx, int precision, void* d_grid, int nx, int ny, int nz,
                  int batch, int sign);
/* plan_batched_z2d + z2d (:127-174): complex-to-real inverse transform of
 * Hermitian-symmetric grids, unnormalised (e^{+iG.r}). The non-redundant
 * half is the fastest dimension: input (device c128) [nz][ny][nx/2 + 1]
 * per batch, output (device f64) [nz][ny][nx] per batch. (The reference
 * halves the dimension cuFFT stores fastest, which on its ix-fastest grids
 * is x; its comment calls it nz, discrepancy D2.) Uses a context workspace
 * of batch nx ny nz complex values. */
int chs_fft3d_z2d(chs_context* ctx, const double* d_half, int nx, int ny, int nz, int batch,
                  double* d_real);
/* downsample_array_to_wave_grid (src/scf.rs:413-454): a real array on the
 * fine grid (device f64, fine_dims x fastest) to the wave grid (device f64)
 * by G-space truncation: forward FFT on the fine grid, each wave-grid point
 * takes the fine coefficient of its frequency (f = i for i <= n/2, i - n
 * above; an even wave grid's Nyquist plane takes the +n/2 coefficient),
 * inverse FFT on the wave grid, real part / (fine point count). The
 * identity when the grids are equal. Uses a context workspace of
 * (fine + wave) complex values. */
int chs_downsample_to_wave_grid(chs_context* ctx, const double* d_fine, const int fine_dims[3],
                                const int wave_dims[3], double* d_wave);

/* ---- Hamiltonian ------------------------------------------------------ */
/* Bind the plane-wave set, grid and local potential used by the operator
 * calls below (the arguments of apply_full_hamiltonian,
 * src/eigensolver/chebyshev.rs:371-403): per-PW kinetic energies
 * (device f64, n_pw), Miller triples (device int32, 3*n_pw), V_eff on the
 * n^3 grid (device f64). Builds the column ("stick") tables of the fused
 * FFT kernels. The device arrays must stay valid while bound. */
int chs_bind_hamiltonian(chs_context* ctx, int n_grid, int n_pw, const int32_t* d_miller,
                         const double* d_kin, const double* d_veff);
/* Replace V_eff (device f64, n^3) in the bound Hamiltonian. */
int chs_set_veff(chs_context* ctx, const double* d_veff);
/* Nonlocal projectors (VnlBatchData::precompute, src/eigensolver/
 * vnl_data.rs:25-89): beta (device c128, projector-major P x n_pw, the
 * per-ion (n_exp x n_pw) blocks concatenated), diagonal D and Q (device f64,
 * P). P = 0 removes them. With q != 0 the overlap S = 1 + beta Q beta^H is
 * active and S^{-1} is prepared by the Woodbury identity. */
int chs_bind_projectors(chs_context* ctx, int n_proj, const double* d_beta, const double* d_d,
                        const double* d_q);
int chs_set_dion(chs_context* ctx, const double* d_d);
/* H|psi> (and S|psi> when d_spsi != NULL) for n_b bands
 * (apply_full_hamiltonian, src/eigensolver/chebyshev.rs:371-403; the
 * kinetic term per plane wave). psi/hpsi/spsi: device, band-major. */
int chs_apply_h(chs_context* ctx, int precision, const void* d_psi, int n_b, void* d_hpsi,
                void* d_spsi);

/* ---- eigensolver ------------------------------------------------------ */
/* compute_spectral_bounds (src/eigensolver/chebyshev.rs:214-254) with an
 * explicit lambda_max; out4 = {lambda_max, eps_cut, center, half_width}. */
int chs_spectral_bounds(const double* eig, int n_eig, double lambda_max, double* out4);
/* lambda_max of the reference's estimate: 0.5 gmax^2 (grid corner of a
 * cubic cell of edge L) + (max V_eff - min V_eff) (chebyshev.rs:225-237,
 * scf.rs:189-194). */
int chs_grid_lambda_max(chs_context* ctx, double L, double* out);
/* k-step Lanczos upper bound of H (or S^{-1}H), FP64:
 * max eig(T_k) + ||f_k|| (Zhou & Li, eq. 2.5). The reference has only the
 * spec (notes/pr-reviews/phase-2/deferred.md:131-160); it replaces the
 * lambda_max of compute_spectral_bounds. Start vector: complex N(0,1) per
 * plane wave from the LCG (src/eigensolver/subspace_diag.rs:434-447) with
 * `seed`. ritz_min (may be NULL) gets min eig(T_k). */
int chs_lanczos_bound(chs_context* ctx, int k, uint64_t seed, double* upper, double* ritz_min);
/* chebyshev_filter (src/eigensolver/chebyshev.rs:626-816): degree-`degree`
 * Chebyshev filter of the operator (H, or S^{-1}H with projectors) on the
 * window bounds4 = {lambda_max, eps_cut, center, half_width}; the
 * reference's 10x norm-growth guard returns CHS_ERR_CHEBYSHEV_DIVERGED
 * with the fields in chs_diag. d_psi (in) and d_out (out) band-major, in
 * the given precision; d_out may equal d_psi. */
int chs_chebyshev_filter(chs_context* ctx, int precision, const void* d_psi, int n_b, int degree,
                         const double* bounds4, void* d_out);
/* Residual-based Chebyshev filter (R-ChFSI: Algorithm 3 of the reference's
 * reference_paper/2025-rchfsi-inexact-mv-paper, main.tex:539-610, with the
 * approximate inverse taken as the exact Woodbury S^{-1}): the degree-m
 * filter of chs_chebyshev_filter on the Rayleigh-Ritz block d_x (c128,
 * n_b x n_pw) with eigenvalues eps (host, n_b), carried for the residuals
 * R_k = Y_k - X T_k((eps - c)/e) from the FP64 residual S^{-1}H X - X eps.
 * The recurrence runs in `precision` (CHS_C64: complex64 products whose
 * rounding scales with the residual); the residual and Y_m = R_m +
 * X T_m are FP64. Equal to the plain filter in exact arithmetic. d_out
 * (c128) may equal d_x. */
int chs_chebyshev_filter_residual(chs_context* ctx, int precision, const double* d_x, int n_b,
                                  const double* eps, int degree, const double* bounds4,
                                  double* d_out);
/* rayleigh_ritz (src/eigensolver/rayleigh_ritz.rs:48-245) with the
 * semantics of diagonalize_and_rotate (src/eigensolver/subspace_diag.rs:
 * 229-391): H_sub = Y^H H Y, S_sub = Y^H S Y (Hermitised), Cholesky-QR
 * S_sub = L L^H, B = L^{-1} H_sub L^{-H}, Hermitian eigensolve, Y <- Y
 * L^{-H} U in place (FP64 throughout). d_hy (may be NULL) receives H Y of
 * the rotated block. Eigenvalues to host array eps (n_b, ascending). */
int chs_rayleigh_ritz(chs_context* ctx, double* d_y, int n_b, double* eps, double* d_hy);
/* Hermitian eigensolver of a device n x n column-major c128 matrix
 * (SolverHandle::zhegvd, src/device/solver.rs:47-95, itype 1, lower, with
 * vectors): generalized when d_b != NULL (B is overwritten by its Cholesky
 * factor), standard otherwise. d_a is overwritten by the eigenvectors;
 * eigenvalues to host array w. */
int chs_zhegvd(chs_context* ctx, int n, double* d_a, double* d_b, double* w);
/* Cholesky-QR orthonormalisation (notes/pr-reviews/phase-2/deferred.md:
 * 27-48): S = Y^H S Y = L L^H, Y <- Y L^{-H} in place (FP64). */
int chs_cholqr(chs_context* ctx, double* d_y, int n_b);

/* band_cg_minimize (src/eigensolver/band_cg.rs:157-437) for every band of
 * the c128 block d_psi (n_b x n_pw, band-major, in place) at once: band i
 * is minimised against the converged set {(psi_j, S psi_j), j < i} of the
 * block as handed in (the Rayleigh-Ritz vectors), with the reference's
 * closed-form line search (line_search.rs:57-198), CG helpers
 * (cg_helpers.rs:49-216) and stopping rules. precond: 0 = TPA
 * (uspp_preconditioner.rs:148-180), 1 = TPA + USPP Woodbury term
 * (:56-131), 2 = a per-(band, plane wave) scale table d_mlp_scale (f32,
 * n_b x n_pw) in place of the TPA diagonal. precond + 256: the lower-band
 * S-orthogonalisation removes the complex overlap <psi_j|S g> instead of its
 * real part (cg_helpers.rs:188 keeps the real part, which lets complex bands
 * drift into lower ones; this option has no counterpart in the reference).
 * Outputs per band (host): eps, n_steps, converged. */
int chs_band_cg(chs_context* ctx, double* d_psi, int n_b, int max_steps, double tol, int precond,
                const float* d_mlp_scale, double* eps, int* n_steps, int* converged);

/* ---- primitives (BlasHandle, src/device/blas.rs:24-241) --------------
 * Column-major, op: 'N', 'T' or 'C'. */
int chs_zgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, const double alpha[2],
              const double* d_a, int lda, const double* d_b, int ldb, const double beta[2],
              double* d_c, int ldc);
/* complex64 counterpart (FP32 storage and accumulation): the projector
 * products of the complex64 filter */
int chs_cgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, const float alpha[2],
              const float* d_a, int lda, const float* d_b, int ldb, const float beta[2],
              float* d_c, int ldc);
int chs_zdotc(chs_context* ctx, int n, const double* d_x, const double* d_y, double out[2]);
/* chs_zdotc that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Its partial sums
 * live in a fixed device array of 514 doubles (4,112 bytes) compiled into
 * the library, one copy per GPU, held until the process ends; the partials
 * do not grow with n, so every n >= 0 is accepted. Calls on one GPU share
 * that array (each launch rewrites all of it before reading it): calls that
 * overlap on one GPU, from contexts on different streams, race. */
int chs_zdotc_noalloc(chs_context* ctx, int n, const double* d_x, const double* d_y,
                      double out[2]);
int chs_zaxpy(chs_context* ctx, int n, const double alpha[2], const double* d_x, double* d_y);
int chs_zscal(chs_context* ctx, int n, const double alpha[2], double* d_x);

/* ---- measurement (no reference counterpart: the reference was never
 * timed) ----------------------------------------------------------------
 * Runs the three kernels of one steady-state Chebyshev step of the bound
 * Hamiltonian on n_b synthetic bands `reps` times, with HIP events on the
 * context stream around each launch. ms3: mean milliseconds of kernel
 * A (scatter + inverse z-FFT), B (plane FFTs x V_eff), C (forward z-FFT +
 * gather + update); bytes3: algorithmic HBM bytes per launch of each;
 * dims4: {n_grid, n_pw, sticks, active x columns}. */
int chs_profile_hpsi(chs_context* ctx, int precision, int n_b, int reps, double* ms3,
                     double* bytes3, int32_t* dims4);
/* Mean duration (HIP events, context stream) of one Chebyshev step's two
 * n_pw-long projector products on the bound projectors with synthetic n_b-band
 * operands: ms2[0] for G = beta^H O, ms2[1] for O + beta C, in the form the
 * filter uses at `precision` (complex64: the FP32 matrix-core form, or the
 * fp16x3 form when CHS_GEMM_FP16X3=1). */
int chs_profile_projector_products(chs_context* ctx, int precision, int n_b, int reps,
                                   double* ms2);
/* Mean milliseconds (HIP events on the context stream) of `reps` complex
 * GEMMs C = op(A) op(B) with op(A) m x k, op(B) k x n, in CHS_C128 or
 * CHS_C64, on synthetic operands in context workspaces. */
int chs_profile_gemm(chs_context* ctx, int precision, char opa, char opb, int m, int n, int k,
                     int reps, double* ms);

/* ---- SCF -------------------------------------------------------------- */
/* Density from band-major c128 wavefunctions with occupations (host f64,
 * n_b): rho(r) = sum_i f_i |psi_i(r)|^2 / Omega (+ augmentation with
 * projectors). construct_density is todo!() in the reference
 * (src/scf.rs:272-280); spec PROJECT_ROOT_PLAN.md:568-611. */
int chs_density(chs_context* ctx, const double* d_psi, int n_b, const double* occ, double omega,
                double* d_rho);

typedef struct chs_scf_params {
  int32_t n_rep;           /* diamond-cubic supercell repetitions per axis */
  double a0;               /* conventional lattice constant (bohr) */
  double zval, rc_loc;     /* V_loc(r) = -zval erf(r/rc_loc)/r */
  double ecut;             /* plane-wave cutoff (Ha) */
  int32_t n_grid, n_bands;
  double kfrac[3];
  int32_t n_proj_per_atom; /* 0 or 4 (l = 0, 1) */
  double beta_rc, q_rc, q0, d_min, d_max;
  uint64_t seed_d, seed_psi, seed_lanczos;
  int32_t cheb_degree, first_passes, lanczos_k, pulay_depth;
  double mix_alpha, kerker_q, etol, smear_kt;
  int32_t max_iter;
  int32_t precision;       /* CHS_C128, or CHS_C64 filter with FP64 RR */
  int32_t fp64_final;      /* with CHS_C64: at least this many FP64 SCF steps at the end */
  double fp32_until;       /* with CHS_C64: |dE|/atom (Ha) below which the filter goes back
                              to FP64 (<= 0: max(2 etol, 3e-7)); it also goes back when |dE|
                              stalls (> 0.9 x the previous) below 10 fp32_until. The
                              convergence test (|dE|/atom < etol twice in a row) then runs
                              on FP64 steps only */
  int32_t fp64_refine;     /* with CHS_C64, the filter of those final FP64 iterations:
                              0 = the complex128 filter, 1 = the residual-based filter
                              (chs_chebyshev_filter_residual) in complex64 with the FP64
                              residual (k-point sets: one such filter per k-point) */
} chs_scf_params;

typedef struct chs_scf_result {
  int32_t n_iter, converged, n_pw, n_proj;  /* n_pw: of the first k-point */
  double energy, mu, seconds;
  double t_filter, t_rr, t_density, t_veff, t_lanczos; /* seconds per phase */
  int32_t n_iter_fp32;     /* SCF iterations whose filter ran in complex64 */
  double drho;             /* integral of |rho_out - rho_in| of the last iteration (electrons);
                              the density criterion of src/scf.rs:357 */
} chs_scf_result;

/* Full SCF (run_scf, src/scf.rs:358-376, with the todo!() transitions
 * build_v_eff / construct_density / check / mix filled in) of the synthetic
 * cell of p at the Gamma point (p->kfrac): atom positions (host f64,
 * 3 * n_atoms), energies (host, max_iter), eps (host, n_bands), rho (host,
 * n^3, may be NULL). Convergence: |dE|/atom < etol on two consecutive
 * iterations; with CHS_C64 both differences are between FP64 iterations. */
int chs_scf_run(chs_context* ctx, const chs_scf_params* p, const double* pos, chs_scf_result* res,
                double* energies, double* eps, double* rho);
/* The same over n_k k-points (the k loop of ScfIteration::diagonalize,
 * src/scf.rs:169-265): kfrac (host, 3 n_k, fractions of the reciprocal
 * lattice vectors; p->kfrac is not read), weights kweight (host, n_k, sum 1).
 * Every k-point has its own plane-wave sphere, start block (seed_psi),
 * projectors and Lanczos bound; V_eff and screened D are shared; one Fermi
 * level over all states; rho = sum_k w_k rho_k. With n_k > 1 the filters of
 * all k-points run as one captured graph launch per pass (chs_kset_filter).
 * Outputs: energies and drho (host, max_iter each, may be NULL), eps (host,
 * n_k * n_bands, k-major), rho (host n^3, may be NULL). */
int chs_scf_run_k(chs_context* ctx, const chs_scf_params* p, const double* pos, int n_k,
                  const double* kfrac, const double* kweight, chs_scf_result* res,
                  double* energies, double* drho, double* eps, double* rho);

/* ---- caller-supplied system (scf_context_create / scf_run /
 * scf_context_destroy, PROJECT_ROOT_PLAN.md:880-1029): the host hands in
 * the cell, grid, start density and wavefunctions, the plane-wave sets and
 * the projectors instead of a synthetic-cell recipe. Host arrays are copied
 * to the device at create and may be freed afterwards. ------------------ */
typedef struct chs_scf_system {
  int32_t grid[3];          /* wave (= density) grid [nx, ny, nz]; the fused FFT kernels are
                               cubic: nx = ny = nz, else CHS_ERR_NOT_IMPLEMENTED */
  double real_lattice[9];   /* a1, a2, a3 as rows (bohr); a cubic cell a I (the reciprocal
                               lattice is (2 pi / a) I), else CHS_ERR_NOT_IMPLEMENTED */
  int32_t n_atoms;
  const double* positions;  /* 3 n_atoms Cartesian (bohr) */
  double n_electrons;
  const double* v_ion;      /* n^3 local ionic potential on the grid (Ha, ix fastest) */
  int32_t n_bands;
  int32_t n_k;
  const double* k_frac;     /* 3 n_k, fractions of the reciprocal lattice vectors */
  const double* k_weight;   /* n_k, sum 1 */
  const int32_t* n_pw;      /* n_k */
  const int32_t* const* miller; /* per k: 3 n_pw[k] Miller triples [h, k, l] */
  int32_t proj_per_atom;    /* 0, 1 (l = 0) or 4 (l = 0, 1); projectors atom-major */
  const double* const* beta; /* per k: (proj_per_atom n_atoms) x n_pw[k] c128 (VnlBatchData,
                               src/eigensolver/vnl_data.rs:25-89); NULL without projectors */
  const double* d0;         /* proj_per_atom n_atoms unscreened diagonal D (Ha) */
  double q0, q_rc;          /* augmentation Q_ij = q0 delta_ij, a normalised Gaussian of width
                               q_rc on the projector's atom; q0 = 0: S = 1 */
  const double* rho0;       /* n^3 start density, or NULL: uniform n_electrons / Omega */
  const double* const* psi0; /* per k: n_bands x n_pw[k] c128 band-major, or NULL: the
                               seeded start of seed_psi */
} chs_scf_system;

typedef struct chs_scf chs_scf;
/* Uploads the system. solver: the solver fields of chs_scf_params
 * (cheb_degree, first_passes, lanczos_k, pulay_depth, mix_alpha, kerker_q,
 * etol, smear_kt, max_iter, precision, fp64_final, fp32_until, seed_psi,
 * seed_lanczos); its cell fields are not read. */
int chs_scf_context_create(chs_context* ctx, const chs_scf_system* sys,
                           const chs_scf_params* solver, chs_scf** out);
/* Runs the SCF from the context's current state (the start density and
 * wavefunctions, or the state the last run left). Returns CHS_OK when
 * converged, CHS_ERR_NOT_CONVERGED otherwise (as scf_run's -1); either way
 * the final density and wavefunctions become the state of the next run.
 * Outputs as chs_scf_run_k. */
int chs_scf_context_run(chs_scf* s, chs_scf_result* res, double* energies, double* drho,
                        double* eps, double* rho);
/* Wavefunctions of k-point k (host, n_bands x n_pw[k] c128 band-major). */
int chs_scf_get_psi(chs_scf* s, int k, double* psi);
/* Frees the system. Destroying its chs_context first frees the device arrays
 * and detaches the system: run and get_psi then return CHS_ERR_INVALID_ARG and
 * destroy frees the host side only. */
int chs_scf_context_destroy(chs_scf* s);

/* The SCF's seeded start block: psi_{b,g} = (x + i y) / (1 + 2 kin_g) with
 * x, y ~ N(0,1) from the reference's 48-bit LCG and Box-Muller
 * (src/eigensolver/subspace_diag.rs:434-447), drawn band-major then
 * plane-wave order; d_kin device f64 (n_pw), d_psi device c128 (n_b x n_pw).
 * Generated on the host (by up to 16 threads with exact LCG jump-ahead) and
 * uploaded. */
int chs_initial_psi(chs_context* ctx, const double* d_kin, int n_pw, int n_b, uint64_t seed,
                    double* d_psi);

/* ---- system set-up on the device (the construction run_scf performs
 * before its loop, src/scf.rs:358-376 with the synthetic cell of the
 * tests, tests/phase0_gate1_consistency.rs:75-112,187-197) --------------
 * Plane waves of p (lexicographic (h, k, l)), V_ion on the n^3 grid and,
 * with n_proj_per_atom > 0, the projectors beta (P x n_pw c128, atom-major),
 * unscreened D0 (seeded uniform [d_min, d_max]) and Q (= q0). Call with all
 * device pointers NULL to get n_pw and n_proj, then again with buffers. */
int chs_build_system(chs_context* ctx, const chs_scf_params* p, const double* pos, int* n_pw,
                     int* n_proj, int32_t* d_miller, double* d_kin, double* d_vion,
                     double* d_beta, double* d_d0, double* d_q);

/* ---- learned band-CG preconditioner (no reference counterpart: the
 * reference's band CG has the TPA preconditioner only,
 * uspp_preconditioner.rs:148-180) ---------------------------------------
 * Network 4 -> 64 -> 64 -> 64 -> 1 (SiLU), bfloat16 weights, FP32
 * accumulation; features f = [kin_g / E_b, log E_b, V_mean, V_std] of each
 * (band b, plane wave g) with E_b the band's kinetic energy and V the bound
 * local potential; scale s = t(kin_g / E_b) exp(o) / E_b (t: the TPA
 * polynomial), so a zero-output network is per-band TPA. Parameters (FP32
 * master, CHS_MLP_PARAMS floats): W1[64][4], b1, W2[64][64], b2, W3[64][64],
 * b3, w4[64], b4. Training pairs, band by band (band_len consecutive pairs
 * per band): features (f32, n x 4), residual r and correction e
 * (interleaved complex f32, n), mask w (f32, n; bands with w <= 0 unused).
 * Loss (band mean, scale-free per band): with a = -Re(conj(r) e),
 * c = |r|^2, L = mean_b [1 - (sum s a)^2 / (sum s^2 c sum |e|^2)];
 * minimised with Adam. */
#define CHS_MLP_PARAMS 8705
typedef struct chs_mlp chs_mlp;
int chs_mlp_create(chs_context* ctx, uint64_t seed, chs_mlp** out);
int chs_mlp_destroy(chs_mlp* m);
int chs_mlp_set_params(chs_context* ctx, chs_mlp* m, const float* theta); /* resets Adam */
int chs_mlp_get_params(chs_context* ctx, chs_mlp* m, float* theta);
/* L and dL/dtheta (host, CHS_MLP_PARAMS floats, may be NULL) over all bands
 * (n a multiple of band_len) */
int chs_mlp_loss_grad(chs_context* ctx, chs_mlp* m, const float* d_feat, const float* d_r,
                      const float* d_e, const float* d_w, int64_t n, int band_len, double* loss,
                      float* grad);
/* `steps` Adam steps (learning rate lr) on consecutive minibatches of
 * `batch_bands` whole bands of a fixed permutation; loss_trace (host,
 * steps, may be NULL) receives each minibatch's loss before its update. */
int chs_mlp_train(chs_context* ctx, chs_mlp* m, const float* d_feat, const float* d_r,
                  const float* d_e, const float* d_w, int64_t n, int band_len, int steps,
                  int batch_bands, float lr, double* loss_trace);
/* Scale table (f32, n_b x n_pw) of a c128 block on the bound Hamiltonian,
 * for chs_band_cg's precond 2. */
int chs_mlp_scale_table(chs_context* ctx, chs_mlp* m, const double* d_psi, int n_b,
                        float* d_table);
/* Training pairs of an unconverged c128 block d_psi against the converged
 * block d_psi_ref (eigenvalues eps_ref, host) on the bound Hamiltonian:
 * r = (H - eps S) psi^ with psi^ = psi / |psi|_S, e = t - psi^ with its
 * psi^ component removed, where t is the normalised projection of psi^ on
 * the reference eigenspace of band b (reference eigenvalues closer than
 * gap_min form one eigenspace). Bands whose projection norm is below 0.9,
 * or whose eigenspace reaches the top of the block, get w = 0; the others
 * w = 1 / (n_used |e_b|^2). Outputs n_b x n_pw pairs, band-major (band
 * length n_pw). */
int chs_mlp_pairs(chs_context* ctx, const double* d_psi, const double* d_psi_ref,
                  const double* eps_ref, int n_b, double gap_min, float* d_feat, float* d_r,
                  float* d_e, float* d_w, int* n_used);

/* ---- k-point sets (the k loop of ScfIteration::diagonalize,
 * src/scf.rs:169-265) ---------------------------------------------------
 * n_k plane-wave sets of one cell (n_pw[k] plane waves, Miller indices and
 * kinetic energies 0.5 |k + G|^2 per k), one shared V_eff and, with
 * n_proj > 0, per-k projectors d_beta[k] (n_proj x n_pw[k] c128) with
 * shared D and Q. */
int chs_kset_bind(chs_context* ctx, int n_k, int n_grid, const int* n_pw,
                  const int32_t* const* d_miller, const double* const* d_kin,
                  const double* d_veff, int n_proj, const double* const* d_beta,
                  const double* d_d, const double* d_q);
int chs_kset_set_veff(chs_context* ctx, const double* d_veff);
/* Make k-point k the bound Hamiltonian for the other entry points
 * (Rayleigh-Ritz, H|psi>, band CG); binding a new Hamiltonian returns it. */
int chs_kset_select(chs_context* ctx, int k);
/* Degree-m Chebyshev filter (chebyshev_filter, src/eigensolver/
 * chebyshev.rs:626-816) of every k-point's block d_psi[k] (n_b x n_pw[k],
 * in place) with windows bounds4[4k .. 4k+3] ({lambda_max, eps_cut, c, e}),
 * all k-points and steps recorded as one HIP graph on the first call for
 * these buffers and replayed by a single launch afterwards. */
int chs_kset_filter(chs_context* ctx, int precision, void* const* d_psi, int n_b, int degree,
                    const double* bounds4);
/* Launch-shape autotuner of the H|psi> kernels for the bound Hamiltonian
 * and precision (sticks per block of the z passes, plane-kernel workgroup
 * size and row-buffer divisor): every candidate timed on n_b bands, the
 * fastest kept for this grid and precision (chs_scf_run tunes on first use
 * of a grid). Outputs the choice and its step time (ms). */
int chs_autotune_hpsi(chs_context* ctx, int precision, int n_b, int* sticks_per_block,
                      int* row_div, int* plane_threads, double* ms);

#ifdef __cplusplus
}
#endif
#endif
