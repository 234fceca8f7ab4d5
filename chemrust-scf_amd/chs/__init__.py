"""Host-side mirror of the chemrust-scf eigensolver surface over libchs_gpu.so.

The names follow the reference (TonyWu20/chemrust-scf): ``Gpu.from_host`` /
``Gpu.sync_to_host`` (src/device/mod.rs:216-265), ``PcieAccount``
(src/device/pcie.rs:21-70), ``apply_full_hamiltonian``
(src/eigensolver/chebyshev.rs:371-403), ``compute_spectral_bounds``
(:214-254), ``chebyshev_filter`` (:626-816), ``rayleigh_ritz``
(src/eigensolver/rayleigh_ritz.rs:48-245), ``pw_coords_to_fft_indices``
(src/scf.rs:387-401) and ``run_scf`` (:358-376). Errors are raised as
``ChsError`` subclasses named after the reference's ``Error`` variants
(src/types.rs:225-248).

Every compute call goes to hand-written gfx950 kernels in libchs_gpu.so;
there is no CPU fallback: if the library or a GPU is missing, the calls
raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libchs_gpu.so")

C128, C64 = 0, 1
OK = 0
_ERR = {
    -1: "NotImplemented", -2: "ChebyshevDiverged", -3: "RayleighRitzFailed", -4: "Solver",
    -5: "Blas", -6: "Fft", -7: "Device", -8: "KernelBuild", -9: "InvalidArgument",
    -10: "OutOfMemory", -11: "NotConverged",
}


class ChsError(RuntimeError):
    def __init__(self, code, msg, diag=None):
        super().__init__(f"{_ERR.get(code, code)}: {msg}")
        self.code = code
        self.diag = diag


class ChebyshevDiverged(ChsError):
    pass


class RayleighRitzFailed(ChsError):
    pass


class Diag(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("norm_previous", C.c_double),
                ("norm_current", C.c_double), ("growth_factor", C.c_double),
                ("solver_info", C.c_int32)]


class ScfParams(C.Structure):
    _fields_ = [
        ("n_rep", C.c_int32), ("a0", C.c_double), ("zval", C.c_double),
        ("rc_loc", C.c_double), ("ecut", C.c_double),
        ("n_grid", C.c_int32), ("n_bands", C.c_int32),
        ("kfrac", C.c_double * 3),
        ("n_proj_per_atom", C.c_int32),
        ("beta_rc", C.c_double), ("q_rc", C.c_double), ("q0", C.c_double),
        ("d_min", C.c_double), ("d_max", C.c_double),
        ("seed_d", C.c_uint64), ("seed_psi", C.c_uint64), ("seed_lanczos", C.c_uint64),
        ("cheb_degree", C.c_int32), ("first_passes", C.c_int32),
        ("lanczos_k", C.c_int32), ("pulay_depth", C.c_int32),
        ("mix_alpha", C.c_double), ("kerker_q", C.c_double), ("etol", C.c_double),
        ("smear_kt", C.c_double),
        ("max_iter", C.c_int32), ("precision", C.c_int32), ("fp64_final", C.c_int32),
        ("fp32_until", C.c_double), ("fp64_refine", C.c_int32),
    ]


class ScfResult(C.Structure):
    _fields_ = [("n_iter", C.c_int32), ("converged", C.c_int32), ("n_pw", C.c_int32),
                ("n_proj", C.c_int32), ("energy", C.c_double), ("mu", C.c_double),
                ("seconds", C.c_double), ("t_filter", C.c_double), ("t_rr", C.c_double),
                ("t_density", C.c_double), ("t_veff", C.c_double), ("t_lanczos", C.c_double),
                ("n_iter_fp32", C.c_int32), ("drho", C.c_double)]


class ScfSystem(C.Structure):
    """chs_scf_system: a caller-supplied cell (scf_context_create,
    PROJECT_ROOT_PLAN.md:880-1029)."""
    _fields_ = [
        ("grid", C.c_int32 * 3), ("real_lattice", C.c_double * 9),
        ("n_atoms", C.c_int32), ("positions", C.c_void_p), ("n_electrons", C.c_double),
        ("v_ion", C.c_void_p), ("n_bands", C.c_int32), ("n_k", C.c_int32),
        ("k_frac", C.c_void_p), ("k_weight", C.c_void_p), ("n_pw", C.c_void_p),
        ("miller", C.c_void_p), ("proj_per_atom", C.c_int32), ("beta", C.c_void_p),
        ("d0", C.c_void_p), ("q0", C.c_double), ("q_rc", C.c_double), ("rho0", C.c_void_p),
        ("psi0", C.c_void_p),
    ]


_LIB = None
vp = C.c_void_p


def lib():
    """Load libchs_gpu.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        i, d, sz, u64 = C.c_int, C.c_double, C.c_size_t, C.c_uint64
        dp = C.POINTER(C.c_double)
        sig = {
            "chs_context_create": [i, C.POINTER(vp)],
            "chs_context_destroy": [vp],
            "chs_synchronize": [vp],
            "chs_diag_get": [vp, C.POINTER(Diag)],
            "chs_transfer_bytes": [vp, C.POINTER(u64), C.POINTER(u64)],
            "chs_transfer_reset": [vp],
            "chs_tuning_entries": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)],
            "chs_alloc": [vp, sz, C.POINTER(vp)],
            "chs_free": [vp, vp],
            "chs_h2d": [vp, vp, vp, sz],
            "chs_d2h": [vp, vp, vp, sz],
            "chs_d2d": [vp, vp, vp, sz],
            "chs_memset": [vp, vp, i, sz],
            "chs_pw_coords_to_fft_indices": [vp, vp, i, i, i, i, vp],
            "chs_gsphere": [vp, d, d, dp, i, i, C.POINTER(i), vp, vp, vp],
            "chs_fft3d_c2c": [vp, i, vp, i, i, i, i, i],
            "chs_fft3d_z2d": [vp, vp, i, i, i, i, vp],
            "chs_downsample_to_wave_grid": [vp, vp, C.POINTER(i), C.POINTER(i), vp],
            "chs_bind_hamiltonian": [vp, i, i, vp, vp, vp],
            "chs_set_veff": [vp, vp],
            "chs_bind_projectors": [vp, i, vp, vp, vp],
            "chs_set_dion": [vp, vp],
            "chs_apply_h": [vp, i, vp, i, vp, vp],
            "chs_spectral_bounds": [dp, i, d, dp],
            "chs_grid_lambda_max": [vp, d, dp],
            "chs_lanczos_bound": [vp, i, u64, dp, dp],
            "chs_chebyshev_filter": [vp, i, vp, i, i, dp, vp],
            "chs_chebyshev_filter_residual": [vp, i, vp, i, dp, i, dp, vp],
            "chs_rayleigh_ritz": [vp, vp, i, dp, vp],
            "chs_zhegvd": [vp, i, vp, vp, dp],
            "chs_cholqr": [vp, vp, i],
            "chs_zgemm": [vp, C.c_char, C.c_char, i, i, i, dp, vp, i, vp, i, dp, vp, i],
            "chs_zdotc": [vp, i, vp, vp, dp],
            "chs_zaxpy": [vp, i, dp, vp, vp],
            "chs_zscal": [vp, i, dp, vp],
            "chs_density": [vp, vp, i, dp, d, vp],
            "chs_scf_run": [vp, C.POINTER(ScfParams), dp, C.POINTER(ScfResult), dp, dp, vp],
            "chs_scf_run_k": [vp, C.POINTER(ScfParams), dp, i, dp, dp, C.POINTER(ScfResult), dp,
                              dp, dp, vp],
            "chs_scf_context_create": [vp, C.POINTER(ScfSystem), C.POINTER(ScfParams),
                                       C.POINTER(vp)],
            "chs_scf_context_run": [vp, C.POINTER(ScfResult), dp, dp, dp, vp],
            "chs_scf_get_psi": [vp, i, vp],
            "chs_initial_psi": [vp, vp, i, i, u64, vp],
            "chs_scf_context_destroy": [vp],
            "chs_profile_hpsi": [vp, i, i, i, dp, dp, vp],
            "chs_band_cg": [vp, vp, i, i, d, i, vp, dp, vp, vp],
            "chs_cgemm": [vp, C.c_char, C.c_char, i, i, i, C.POINTER(C.c_float), vp, i, vp, i,
                          C.POINTER(C.c_float), vp, i],
            "chs_profile_gemm": [vp, i, C.c_char, C.c_char, i, i, i, i, dp],
            "chs_profile_projector_products": [vp, i, i, i, dp],
            "chs_kset_bind": [vp, i, i, vp, vp, vp, vp, i, vp, vp, vp],
            "chs_kset_set_veff": [vp, vp],
            "chs_kset_select": [vp, i],
            "chs_kset_filter": [vp, i, vp, i, i, dp],
            "chs_autotune_hpsi": [vp, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i), dp],
            "chs_build_system": [vp, C.POINTER(ScfParams), dp, C.POINTER(i), C.POINTER(i), vp, vp,
                                 vp, vp, vp, vp],
            "chs_mlp_create": [vp, u64, C.POINTER(vp)],
            "chs_mlp_destroy": [vp],
            "chs_mlp_set_params": [vp, vp, vp],
            "chs_mlp_get_params": [vp, vp, vp],
            "chs_mlp_loss_grad": [vp, vp, vp, vp, vp, vp, C.c_int64, i, dp, vp],
            "chs_mlp_train": [vp, vp, vp, vp, vp, vp, C.c_int64, i, i, i, C.c_float, dp],
            "chs_mlp_scale_table": [vp, vp, vp, i, vp],
            "chs_mlp_pairs": [vp, vp, vp, dp, i, d, vp, vp, vp, vp, C.POINTER(i)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.chs_last_error.restype = C.c_char_p
        L.chs_last_error.argtypes = [vp]
        _LIB = L
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One device context = one HIP stream (CudaContext::new + default stream)."""

    def __init__(self, device=0):
        h = vp()
        rc = lib().chs_context_create(device, C.byref(h))
        if rc != OK:
            raise ChsError(rc, f"cannot create a context on device {device}")
        self.h = h
        self._bufs = []

    def close(self):
        if getattr(self, "h", None):
            lib().chs_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, rc):
        if rc == OK:
            return
        msg = lib().chs_last_error(self.h).decode(errors="replace")
        dg = Diag()
        lib().chs_diag_get(self.h, C.byref(dg))
        cls = {-2: ChebyshevDiverged, -3: RayleighRitzFailed}.get(rc, ChsError)
        raise cls(rc, msg, dg)

    def synchronize(self):
        self.check(lib().chs_synchronize(self.h))

    def transfer_bytes(self):
        """PcieAccount {h2d_bytes, d2h_bytes} of this context."""
        a, b = C.c_uint64(), C.c_uint64()
        self.check(lib().chs_transfer_bytes(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def transfer_reset(self):
        self.check(lib().chs_transfer_reset(self.h))

    def tuning_entries(self):
        """(H|psi> launch shapes, GEMM variants, density workgroup sizes) held."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.check(lib().chs_tuning_entries(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


class Gpu:
    """Device array owned by the caller (Gpu<T>, src/device/mod.rs:50-58)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = vp()
        ctx.check(lib().chs_alloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, arr, dtype=None):
        arr = np.ascontiguousarray(arr, dtype=dtype)
        g = cls(ctx, arr.shape, arr.dtype)
        ctx.check(lib().chs_h2d(ctx.h, g.ptr, arr.ctypes.data_as(vp), g.nbytes))
        return g

    def sync_to_host(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib().chs_d2h(self.ctx.h, out.ctypes.data_as(vp), self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ctx.check(lib().chs_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            lib().chs_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def precision_of(dtype):
    return C128 if np.dtype(dtype) == np.complex128 else C64


# ---- setup ---------------------------------------------------------------

def pw_coords_to_fft_indices(ctx, miller, grid):
    """src/scf.rs:387-401 on the device: Miller (n_pw, 3) int32 -> int32 indices."""
    miller = np.ascontiguousarray(miller, np.int32).reshape(-1, 3)
    n_pw = miller.shape[0]
    dm = Gpu.from_host(ctx, miller)
    di = Gpu(ctx, (n_pw,), np.int32)
    ngx, ngy, ngz = grid
    ctx.check(lib().chs_pw_coords_to_fft_indices(ctx.h, dm.ptr, n_pw, ngx, ngy, ngz, di.ptr))
    return di.sync_to_host()


def gsphere(ctx, L, ecut, n_grid, kfrac=(0.0, 0.0, 0.0)):
    """Plane-wave sphere on the device: returns (miller, fft_idx, kin) as device arrays."""
    kf = np.ascontiguousarray(kfrac, np.float64)
    n = C.c_int()
    ctx.check(lib().chs_gsphere(ctx.h, L, ecut, _dptr(kf), n_grid, 0, C.byref(n), None, None, None))
    n_pw = n.value
    mil = Gpu(ctx, (n_pw, 3), np.int32)
    idx = Gpu(ctx, (n_pw,), np.int32)
    kin = Gpu(ctx, (n_pw,), np.float64)
    ctx.check(lib().chs_gsphere(ctx.h, L, ecut, _dptr(kf), n_grid, n_pw, C.byref(n), mil.ptr,
                                idx.ptr, kin.ptr))
    return mil, idx, kin


def fft3d_c2c(ctx, grid, sign):
    """BatchedFftPlan3d c2c (plan_batched_c2c(nx, ny, nz, batch); sign +1
    inverse, -1 forward) in place on a device array of shape
    (batch, nz, ny, nx) (x fastest)."""
    batch, nz, ny, nx = grid.shape
    ctx.check(lib().chs_fft3d_c2c(ctx.h, precision_of(grid.dtype), grid.ptr, nx, ny, nz, batch,
                                  sign))


def downsample_to_wave_grid(ctx, fine, wave_dims):
    """downsample_array_to_wave_grid (src/scf.rs:413-454): device f64 array of
    shape (nz, ny, nx) on the fine grid -> a new device array on the wave
    grid wave_dims = (wx, wy, wz)."""
    fz, fy, fx = fine.shape
    fd = (C.c_int * 3)(fx, fy, fz)
    wd = (C.c_int * 3)(*wave_dims)
    out = Gpu(ctx, (wave_dims[2], wave_dims[1], wave_dims[0]), np.float64)
    ctx.check(lib().chs_downsample_to_wave_grid(ctx.h, fine.ptr, fd, wd, out.ptr))
    return out


def fft3d_z2d(ctx, half, nx):
    """BatchedFftPlan3d z2d (plan_batched_z2d): Hermitian half grids, device
    complex128 of shape (batch, nz, ny, nx // 2 + 1), to real grids
    (batch, nz, ny, nx), unnormalised inverse."""
    batch, nz, ny, nh = half.shape
    assert nh == nx // 2 + 1
    out = Gpu(ctx, (batch, nz, ny, nx), np.float64)
    ctx.check(lib().chs_fft3d_z2d(ctx.h, half.ptr, nx, ny, nz, batch, out.ptr))
    return out


# ---- Hamiltonian and eigensolver ----------------------------------------

class Hamiltonian:
    """Bound plane-wave set + V_eff (+ projectors) on a context."""

    def __init__(self, ctx, n_grid, miller, kin, veff):
        self.ctx = ctx
        self.n_grid = n_grid
        self.miller = miller if isinstance(miller, Gpu) else Gpu.from_host(ctx, np.asarray(miller, np.int32))
        self.kin = kin if isinstance(kin, Gpu) else Gpu.from_host(ctx, np.asarray(kin, np.float64))
        self.veff = veff if isinstance(veff, Gpu) else Gpu.from_host(ctx, np.asarray(veff, np.float64))
        self.n_pw = self.kin.shape[0]
        ctx.check(lib().chs_bind_hamiltonian(ctx.h, n_grid, self.n_pw, self.miller.ptr,
                                             self.kin.ptr, self.veff.ptr))
        self.n_proj = 0
        # the context reads these device arrays until another binding replaces them
        ctx._bound_hamiltonian = self

    def set_veff(self, veff):
        self.veff = veff if isinstance(veff, Gpu) else Gpu.from_host(self.ctx, np.asarray(veff, np.float64))
        self.ctx.check(lib().chs_set_veff(self.ctx.h, self.veff.ptr))

    def bind_projectors(self, beta, d, q):
        self.beta = beta if isinstance(beta, Gpu) else Gpu.from_host(self.ctx, np.asarray(beta, np.complex128))
        self.d = d if isinstance(d, Gpu) else Gpu.from_host(self.ctx, np.asarray(d, np.float64))
        self.q = q if isinstance(q, Gpu) else Gpu.from_host(self.ctx, np.asarray(q, np.float64))
        self.n_proj = self.beta.shape[0]
        self.ctx.check(lib().chs_bind_projectors(self.ctx.h, self.n_proj, self.beta.ptr,
                                                 self.d.ptr, self.q.ptr))

    def set_dion(self, d):
        self.d = d if isinstance(d, Gpu) else Gpu.from_host(self.ctx, np.asarray(d, np.float64))
        self.ctx.check(lib().chs_set_dion(self.ctx.h, self.d.ptr))


def apply_full_hamiltonian(ctx, psi, want_s=False):
    """H|psi> (and S|psi>) for a device block psi of shape (n_b, n_pw)."""
    hpsi = Gpu(ctx, psi.shape, psi.dtype)
    spsi = Gpu(ctx, psi.shape, psi.dtype) if want_s else None
    ctx.check(lib().chs_apply_h(ctx.h, precision_of(psi.dtype), psi.ptr, psi.shape[0], hpsi.ptr,
                                spsi.ptr if spsi else None))
    return (hpsi, spsi) if want_s else hpsi


def compute_spectral_bounds(eigenvalues, lambda_max):
    """{lambda_max, eps_cut, center, half_width} (chebyshev.rs:214-254)."""
    eig = np.ascontiguousarray(eigenvalues if eigenvalues is not None else [], np.float64)
    out = np.zeros(4)
    rc = lib().chs_spectral_bounds(_dptr(eig), len(eig), lambda_max, _dptr(out))
    if rc:
        raise ChsError(rc, "spectral bounds")
    return out


def grid_lambda_max(ctx, L):
    out = C.c_double()
    ctx.check(lib().chs_grid_lambda_max(ctx.h, L, C.byref(out)))
    return out.value


def lanczos_bound(ctx, k=10, seed=4):
    up, lo = C.c_double(), C.c_double()
    ctx.check(lib().chs_lanczos_bound(ctx.h, k, seed, C.byref(up), C.byref(lo)))
    return up.value, lo.value


def chebyshev_filter(ctx, psi, degree, bounds, out=None):
    """Degree-m filter of device block psi (n_b, n_pw); returns the filtered block."""
    out = out if out is not None else Gpu(ctx, psi.shape, psi.dtype)
    b4 = np.ascontiguousarray(bounds, np.float64)
    ctx.check(lib().chs_chebyshev_filter(ctx.h, precision_of(psi.dtype), psi.ptr, psi.shape[0],
                                         degree, _dptr(b4), out.ptr))
    return out


def chebyshev_filter_residual(ctx, x, eps, degree, bounds, precision=C64, out=None):
    """Residual-based filter (R-ChFSI) of the c128 Rayleigh-Ritz block x with
    eigenvalues eps; recurrence in `precision`, residual and assembly FP64."""
    out = out if out is not None else Gpu(ctx, x.shape, np.complex128)
    b4 = np.ascontiguousarray(bounds, np.float64)
    e = np.ascontiguousarray(eps, np.float64)
    ctx.check(lib().chs_chebyshev_filter_residual(ctx.h, precision, x.ptr, x.shape[0], _dptr(e),
                                                  degree, _dptr(b4), out.ptr))
    return out


def zgemm(ctx, opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """BlasHandle::gemm_c64 (src/device/blas.rs:24-61), column-major."""
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    ctx.check(lib().chs_zgemm(ctx.h, opa.encode(), opb.encode(), m, n, k, _dptr(al), a.ptr, lda,
                              b.ptr, ldb, _dptr(be), c.ptr, ldc))


def zhegvd(ctx, a, b=None):
    """SolverHandle::zhegvd (src/device/solver.rs:47-95): itype 1, lower, vectors.
    a, b: device (n, n) c128 column-major; returns eigenvalues; a <- eigenvectors."""
    n = a.shape[0]
    w = np.zeros(n)
    ctx.check(lib().chs_zhegvd(ctx.h, n, a.ptr, b.ptr if b is not None else None, _dptr(w)))
    return w


def cholqr(ctx, y):
    """Cholesky-QR orthonormalisation of a c128 device block (in place)."""
    ctx.check(lib().chs_cholqr(ctx.h, y.ptr, y.shape[0]))


def density(ctx, psi, occ, omega, n_grid):
    """rho(r) = sum_b occ_b |psi_b(r)|^2 / Omega for a c128 device block on the
    bound n_grid^3 grid (construct_density, src/scf.rs:272-280)."""
    occ = np.ascontiguousarray(occ, np.float64)
    rho = Gpu(ctx, (n_grid ** 3,), np.float64)
    ctx.check(lib().chs_density(ctx.h, psi.ptr, psi.shape[0], _dptr(occ), omega, rho.ptr))
    return rho


def profile_hpsi(ctx, n_b, precision=C128, reps=20):
    """Mean duration (ms) and algorithmic bytes of the three H|psi> kernels of
    one Chebyshev step on the bound Hamiltonian (HIP events, context stream)."""
    ms = np.zeros(3)
    by = np.zeros(3)
    dims = np.zeros(4, np.int32)
    ctx.check(lib().chs_profile_hpsi(ctx.h, precision, n_b, reps, _dptr(ms), _dptr(by),
                                     dims.ctypes.data_as(vp)))
    return dict(ms=ms, bytes=by, n_grid=int(dims[0]), n_pw=int(dims[1]), sticks=int(dims[2]),
                active_x=int(dims[3]))


def synth_positions(n_rep, a0=10.26, sigma=0.05, seed=1):
    """Diamond-cubic supercell (n_rep^3 x 8 atoms, replica-major then basis
    order) with every Cartesian component displaced by sigma * N(0,1) from
    the 48-bit LCG (src/eigensolver/subspace_diag.rs:434-447) and
    Box-Muller sqrt(-2 ln(1-u1)) cos(2 pi u2)."""
    basis = [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0),
             (.25, .25, .25), (.25, .75, .75), (.75, .25, .75), (.75, .75, .25)]
    state = seed

    def uni():
        nonlocal state
        state = (state * 25214903917 + 11) & 0xFFFFFFFFFFFF
        return state / 281474976710656.0

    def normal():
        u1 = 1.0 - uni()
        u2 = uni()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    pos = []
    for ix in range(n_rep):
        for iy in range(n_rep):
            for iz in range(n_rep):
                for b in basis:
                    for c, cell in enumerate((ix, iy, iz)):
                        pos.append(a0 * (cell + b[c]) + sigma * normal())
    return np.array(pos)


def scf_params(**kw):
    """chs_scf_params with the defaults of the reference-sized config A."""
    p = ScfParams()
    d = dict(n_rep=1, a0=10.26, zval=4.0, rc_loc=1.0, ecut=15.0, n_grid=36, n_bands=20,
             n_proj_per_atom=0, beta_rc=1.3, q_rc=1.2, q0=0.0, d_min=0.5, d_max=2.0, seed_d=3,
             seed_psi=2, seed_lanczos=4, cheb_degree=10, first_passes=4, lanczos_k=10,
             pulay_depth=6, mix_alpha=0.5, kerker_q=0.8, etol=1e-6, smear_kt=0.01,
             max_iter=60, precision=C128, fp64_final=2, fp32_until=0.0, fp64_refine=0)
    d.update(kw)
    for k, v in d.items():
        if k == "kfrac":
            for i in range(3):
                p.kfrac[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def run_scf(ctx, params, pos, want_rho=False):
    """run_scf (src/scf.rs:358-376): full SCF on the device. Returns a dict
    with n_iter, converged, energies, eps, timings (and rho)."""
    pos = np.ascontiguousarray(pos, np.float64)
    ctx._bound_hamiltonian = None  # the SCF binds (and releases) its own arrays
    res = ScfResult()
    energies = np.zeros(params.max_iter)
    eps = np.zeros(params.n_bands)
    rho = np.zeros(params.n_grid ** 3) if want_rho else None
    rc = lib().chs_scf_run(ctx.h, C.byref(params), _dptr(pos), C.byref(res), _dptr(energies),
                           _dptr(eps), rho.ctypes.data_as(vp) if rho is not None else None)
    if rc not in (OK, -11):
        ctx.check(rc)
    out = {f: getattr(res, f) for f, _ in ScfResult._fields_}
    out["energies"] = energies[: res.n_iter]
    out["eps"] = eps
    out["rho"] = rho
    out["status"] = rc
    return out


def run_scf_k(ctx, params, pos, kpts, weights, want_rho=False):
    """k-point SCF of the synthetic cell (the k loop of ScfIteration::
    diagonalize, src/scf.rs:169-265): one Fermi level, rho = sum_k w_k rho_k,
    all k-points' filters in one graph launch. eps is (n_k, n_bands)."""
    pos = np.ascontiguousarray(pos, np.float64)
    kpts = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(weights, np.float64)
    ctx._bound_hamiltonian = None
    ctx._bound_kset = None
    res = ScfResult()
    energies = np.zeros(params.max_iter)
    drho = np.zeros(params.max_iter)
    eps = np.zeros((len(kpts), params.n_bands))
    rho = np.zeros(params.n_grid ** 3) if want_rho else None
    rc = lib().chs_scf_run_k(ctx.h, C.byref(params), _dptr(pos), len(kpts), _dptr(kpts), _dptr(w),
                             C.byref(res), _dptr(energies), _dptr(drho), _dptr(eps),
                             rho.ctypes.data_as(vp) if rho is not None else None)
    if rc not in (OK, -11):
        ctx.check(rc)
    out = {f: getattr(res, f) for f, _ in ScfResult._fields_}
    out["energies"] = energies[: res.n_iter]
    out["drho_trace"] = drho[: res.n_iter]
    out["eps"] = eps
    out["rho"] = rho
    out["status"] = rc
    return out


class ScfContext:
    """A caller-supplied system on the device (chs_scf_context_create /
    chs_scf_context_run / chs_scf_context_destroy; the reference's planned
    scf_context_create / scf_run / scf_context_destroy,
    PROJECT_ROOT_PLAN.md:880-1029). Arrays are host numpy arrays:
    lattice (3, 3), positions (n_atoms, 3), v_ion (n^3), miller per k
    (n_pw_k, 3), beta per k (P, n_pw_k) complex, d0 (P), rho0 (n^3),
    psi0 per k (n_bands, n_pw_k) complex. `solver` is a scf_params() whose
    solver fields are used."""

    def __init__(self, ctx, grid, lattice, positions, n_electrons, v_ion, n_bands, kpts,
                 weights, miller, solver, proj_per_atom=0, beta=None, d0=None, q0=0.0,
                 q_rc=1.2, rho0=None, psi0=None):
        self.ctx = ctx
        self.n_bands = n_bands
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data

        def ptrs(lst, dt):
            if lst is None:
                return None
            ps = (C.c_void_p * len(lst))(*[arr(a, dt) for a in lst])
            keep.append(ps)
            return C.cast(ps, C.c_void_p)

        kpts = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
        self.n_pw = [int(np.asarray(m).shape[0]) for m in miller]
        sy = ScfSystem()
        for i in range(3):
            sy.grid[i] = grid[i]
        for i, v in enumerate(np.asarray(lattice, np.float64).reshape(9)):
            sy.real_lattice[i] = v
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        sy.n_atoms = pos.shape[0]
        sy.positions = arr(pos, np.float64)
        sy.n_electrons = n_electrons
        sy.v_ion = arr(v_ion, np.float64)
        sy.n_bands = n_bands
        sy.n_k = len(kpts)
        sy.k_frac = arr(kpts, np.float64)
        sy.k_weight = arr(weights, np.float64)
        sy.n_pw = arr(self.n_pw, np.int32)
        sy.miller = ptrs(miller, np.int32)
        sy.proj_per_atom = proj_per_atom
        sy.beta = ptrs(beta, np.complex128)
        sy.d0 = arr(d0, np.float64) if d0 is not None else None
        sy.q0 = q0
        sy.q_rc = q_rc
        sy.rho0 = arr(rho0, np.float64) if rho0 is not None else None
        sy.psi0 = ptrs(psi0, np.complex128)
        self.n_grid = grid[0]
        self.max_iter = solver.max_iter
        h = vp()
        ctx.check(lib().chs_scf_context_create(ctx.h, C.byref(sy), C.byref(solver), C.byref(h)))
        self.h = h

    def run(self, want_rho=False):
        res = ScfResult()
        energies = np.zeros(self.max_iter)
        drho = np.zeros(self.max_iter)
        eps = np.zeros((len(self.n_pw), self.n_bands))
        rho = np.zeros(self.n_grid ** 3) if want_rho else None
        self.ctx._bound_hamiltonian = None
        self.ctx._bound_kset = None
        rc = lib().chs_scf_context_run(self.h, C.byref(res), _dptr(energies), _dptr(drho),
                                       _dptr(eps), rho.ctypes.data_as(vp) if rho is not None else None)
        if rc not in (OK, -11):
            self.ctx.check(rc)
        out = {f: getattr(res, f) for f, _ in ScfResult._fields_}
        out["energies"] = energies[: res.n_iter]
        out["drho_trace"] = drho[: res.n_iter]
        out["eps"] = eps
        out["rho"] = rho
        out["status"] = rc
        return out

    def psi(self, k=0):
        out = np.zeros((self.n_bands, self.n_pw[k]), np.complex128)
        self.ctx.check(lib().chs_scf_get_psi(self.h, k, out.ctypes.data_as(vp)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().chs_scf_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def profile_projector_products(ctx, n_b, precision=C64, reps=5):
    """Mean ms of the filter's two projector products (G = beta^H O,
    O + beta C) on the bound projectors, in the form the filter uses."""
    ms = np.zeros(2)
    ctx.check(lib().chs_profile_projector_products(ctx.h, precision, n_b, reps, _dptr(ms)))
    return ms


def profile_gemm(ctx, precision, opa, opb, m, n, k, reps=5):
    """Mean ms of a complex GEMM of the given shape (HIP events, context stream)."""
    ms = C.c_double()
    ctx.check(lib().chs_profile_gemm(ctx.h, precision, opa.encode(), opb.encode(), m, n, k, reps,
                                     C.byref(ms)))
    return ms.value


def initial_psi(ctx, kin, n_b, seed):
    """The SCF's seeded start block (n_b, n_pw) for device kinetic energies
    `kin` (the LCG + Box-Muller stream of the driver)."""
    out = Gpu(ctx, (n_b, kin.shape[0]), np.complex128)
    ctx.check(lib().chs_initial_psi(ctx.h, kin.ptr, kin.shape[0], n_b, seed, out.ptr))
    return out


def build_system(ctx, params, pos):
    """Plane waves, V_ion and (USPP) projectors of a synthetic cell on the
    device. Returns a dict of device arrays: miller, kin, vion and, with
    projectors, beta (P, n_pw), d0, q."""
    pos = np.ascontiguousarray(pos, np.float64)
    n_pw, n_proj = C.c_int(), C.c_int()
    ctx.check(lib().chs_build_system(ctx.h, C.byref(params), _dptr(pos), C.byref(n_pw),
                                     C.byref(n_proj), None, None, None, None, None, None))
    npw, P = n_pw.value, n_proj.value
    out = {"miller": Gpu(ctx, (npw, 3), np.int32), "kin": Gpu(ctx, (npw,), np.float64),
           "vion": Gpu(ctx, (params.n_grid ** 3,), np.float64), "n_pw": npw, "n_proj": P}
    if P:
        out.update(beta=Gpu(ctx, (P, npw), np.complex128), d0=Gpu(ctx, (P,), np.float64),
                   q=Gpu(ctx, (P,), np.float64))
    ptr = lambda k: out[k].ptr if k in out else None
    ctx.check(lib().chs_build_system(ctx.h, C.byref(params), _dptr(pos), C.byref(n_pw),
                                     C.byref(n_proj), out["miller"].ptr, out["kin"].ptr,
                                     out["vion"].ptr, ptr("beta"), ptr("d0"), ptr("q")))
    return out


class Mlp:
    """The learned band-CG preconditioner network on a context (chs_mlp_*)."""
    N_PARAMS = 8705

    def __init__(self, ctx, seed=11):
        self.ctx = ctx
        h = vp()
        ctx.check(lib().chs_mlp_create(ctx.h, seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().chs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        th = np.zeros(self.N_PARAMS, np.float32)
        self.ctx.check(lib().chs_mlp_get_params(self.ctx.h, self.h, th.ctypes.data_as(vp)))
        return th

    @params.setter
    def params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == self.N_PARAMS
        self.ctx.check(lib().chs_mlp_set_params(self.ctx.h, self.h, th.ctypes.data_as(vp)))

    def loss_grad(self, pairs, n=None):
        """Band loss and its gradient over the first n pairs (whole bands)."""
        n = pairs["n"] if n is None else n
        loss = C.c_double()
        g = np.zeros(self.N_PARAMS, np.float32)
        self.ctx.check(lib().chs_mlp_loss_grad(self.ctx.h, self.h, pairs["feat"].ptr,
                                               pairs["r"].ptr, pairs["e"].ptr, pairs["w"].ptr, n,
                                               pairs["band_len"], C.byref(loss),
                                               g.ctypes.data_as(vp)))
        return loss.value, g

    def train(self, pairs, steps, batch_bands, lr, n=None):
        """Adam steps on minibatches of whole bands; returns the loss trace."""
        n = pairs["n"] if n is None else n
        trace = np.zeros(max(steps, 1))
        self.ctx.check(lib().chs_mlp_train(self.ctx.h, self.h, pairs["feat"].ptr, pairs["r"].ptr,
                                           pairs["e"].ptr, pairs["w"].ptr, n, pairs["band_len"],
                                           steps, batch_bands, lr, _dptr(trace)))
        return trace[:steps]

    def scale_table(self, psi, out=None):
        out = out if out is not None else Gpu(self.ctx, psi.shape, np.float32)
        self.ctx.check(lib().chs_mlp_scale_table(self.ctx.h, self.h, psi.ptr, psi.shape[0],
                                                 out.ptr))
        return out


def mlp_pair_buffers(ctx, n):
    """Device arrays for n training pairs (feat n x 4, r / e complex64, w)."""
    return {"feat": Gpu(ctx, (n, 4), np.float32), "r": Gpu(ctx, (n,), np.complex64),
            "e": Gpu(ctx, (n,), np.complex64), "w": Gpu(ctx, (n,), np.float32), "n": 0,
            "cap": n, "band_len": 0}


def mlp_pairs(ctx, psi, psi_ref, eps_ref, pairs, gap_min=1e-3):
    """Append the training pairs of block psi against the converged psi_ref to
    the buffers of mlp_pair_buffers; returns the number of bands used."""
    n_b, n_pw = psi.shape
    o = pairs["n"]
    if o + n_b * n_pw > pairs["cap"]:
        raise ValueError("pair buffers full")
    if pairs["band_len"] not in (0, n_pw):
        raise ValueError("pairs of one buffer set must share the band length")
    pairs["band_len"] = n_pw
    used = C.c_int()
    er = np.ascontiguousarray(eps_ref, np.float64)
    at = lambda k, sz: C.c_void_p(pairs[k].ptr.value + o * sz)
    ctx.check(lib().chs_mlp_pairs(ctx.h, psi.ptr, psi_ref.ptr, _dptr(er), n_b, gap_min,
                                  at("feat", 16), at("r", 8), at("e", 8), at("w", 4),
                                  C.byref(used)))
    pairs["n"] = o + n_b * n_pw
    return used.value


class KSet:
    """k-point set of one cell on a context (chs_kset_*): per k a plane-wave
    set (miller, kin device arrays), optional projectors beta[k]; shared
    V_eff, D and Q."""

    def __init__(self, ctx, n_grid, millers, kins, veff, betas=None, d=None, q=None):
        self.ctx = ctx
        self.millers, self.kins, self.veff = list(millers), list(kins), veff
        self.betas, self.d, self.q = (list(betas) if betas is not None else None), d, q
        K = len(self.kins)
        npw = np.array([k.shape[0] for k in self.kins], np.int32)
        P = self.betas[0].shape[0] if self.betas else 0
        arr = lambda xs: (vp * len(xs))(*[x.ptr for x in xs])
        self.n_pw = npw
        ctx.check(lib().chs_kset_bind(ctx.h, K, n_grid, npw.ctypes.data_as(vp), arr(self.millers),
                                      arr(self.kins), veff.ptr, P,
                                      arr(self.betas) if P else None,
                                      d.ptr if P else None, q.ptr if P else None))
        ctx._bound_kset = self

    def filter(self, psis, degree, bounds, precision=None):
        """In-place filter of every k-point's device block; bounds (K, 4)."""
        b = np.ascontiguousarray(bounds, np.float64).reshape(-1)
        prec = precision_of(psis[0].dtype) if precision is None else precision
        ptrs = (vp * len(psis))(*[p.ptr for p in psis])
        self.ctx.check(lib().chs_kset_filter(self.ctx.h, prec, ptrs, psis[0].shape[0], degree,
                                             _dptr(b)))

    def select(self, k):
        self.ctx.check(lib().chs_kset_select(self.ctx.h, k))


def autotune_hpsi(ctx, n_b, precision=C128):
    """Tune the H|psi> launch shape for the bound Hamiltonian; returns
    (sticks per block, row-buffer divisor, plane threads, step ms)."""
    sb, rd, th, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    ctx.check(lib().chs_autotune_hpsi(ctx.h, precision, n_b, C.byref(sb), C.byref(rd),
                                      C.byref(th), C.byref(ms)))
    return sb.value, rd.value, th.value, ms.value


def band_cg_minimize(ctx, psi, max_steps, tol, precond=0, mlp_scale=None, full_overlap=False):
    """band_cg_minimize (src/eigensolver/band_cg.rs:157-437) for every band of the
    c128 device block psi in place; band i against bands j < i of the block as
    handed in. precond: 0 TPA, 1 TPA + USPP term, 2 MLP scale table (device f32,
    n_b x n_pw). full_overlap: remove the complex lower-band overlap instead of
    its real part (cg_helpers.rs:188; not in the reference). Returns (eps,
    n_steps, converged)."""
    n_b = psi.shape[0]
    eps = np.zeros(n_b)
    steps = np.zeros(n_b, np.int32)
    conv = np.zeros(n_b, np.int32)
    ctx.check(lib().chs_band_cg(ctx.h, psi.ptr, n_b, max_steps, tol,
                                precond | (256 if full_overlap else 0),
                                mlp_scale.ptr if mlp_scale is not None else None, _dptr(eps),
                                steps.ctypes.data_as(vp), conv.ctypes.data_as(vp)))
    return eps, steps, conv


def rayleigh_ritz(ctx, y, want_hy=False):
    """In-place Rayleigh-Ritz of a c128 device block; returns (eps, hy or None)."""
    eps = np.zeros(y.shape[0])
    hy = Gpu(ctx, y.shape, y.dtype) if want_hy else None
    ctx.check(lib().chs_rayleigh_ritz(ctx.h, y.ptr, y.shape[0], _dptr(eps), hy.ptr if hy else None))
    return eps, hy
