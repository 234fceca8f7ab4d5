"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


class Params(C.Structure):
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
        ("max_iter", C.c_int32),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        vp = C.c_void_p
        sig = {
            "orc_default_params": (None, [C.POINTER(Params)]),
            "orc_set_num_threads": (None, [C.c_int]),
            "orc_downsample": (None, [f64p, C.c_int, C.c_int, f64p]),
            "orc_get_max_threads": (C.c_int, []),
            "orc_lcg_uniforms": (C.c_double, [C.c_uint64, C.c_int, f64p]),
            "orc_lcg_normals": (None, [C.c_uint64, C.c_int, f64p]),
            "orc_synth_positions": (None, [C.c_int, C.c_double, C.c_double, C.c_uint64, f64p]),
            "orc_gsphere": (C.c_int, [C.c_double, C.c_double, f64p, C.c_int, C.c_int, vp, vp, vp]),
            "orc_pw_coords_to_fft_indices": (None, [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p]),
            "orc_fft3d": (None, [c128p, C.c_int, C.c_int]),
            "orc_model_create": (vp, [C.POINTER(Params), f64p]),
            "orc_model_destroy": (None, [vp]),
            "orc_model_n_pw": (C.c_int, [vp]),
            "orc_model_n_proj": (C.c_int, [vp]),
            "orc_model_get_pw": (None, [vp, i32p, i32p, f64p]),
            "orc_model_get_vion": (None, [vp, f64p]),
            "orc_model_get_beta": (None, [vp, c128p]),
            "orc_model_get_d": (None, [vp, f64p]),
            "orc_model_set_d": (None, [vp, f64p]),
            "orc_model_set_veff": (None, [vp, f64p]),
            "orc_model_set_sinv_filter": (None, [vp, C.c_int]),
            "orc_initial_psi": (None, [vp, C.c_int, C.c_uint64, c128p]),
            "orc_apply_h": (None, [vp, c128p, C.c_int, c128p, vp]),
            "orc_apply_sinv": (None, [vp, c128p, C.c_int]),
            "orc_grid_lambda_max": (C.c_double, [vp]),
            "orc_lanczos_bound": (C.c_double, [vp, C.c_int, C.c_uint64, C.POINTER(C.c_double)]),
            "orc_spectral_bounds": (None, [f64p, C.c_int, C.c_double, f64p]),
            "orc_chebyshev_filter": (C.c_int, [vp, c128p, C.c_int, C.c_int, f64p, c128p, f64p]),
            "orc_rayleigh_ritz": (C.c_int, [vp, c128p, C.c_int, f64p, vp]),
            "orc_density": (None, [vp, c128p, C.c_int, f64p, f64p]),
            "orc_fermi_occupations": (C.c_double, [f64p, C.c_int, C.c_double, C.c_double, f64p,
                                                   C.POINTER(C.c_double)]),
            "orc_vhxc": (None, [vp, f64p, f64p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_screen_d": (None, [vp, f64p]),
            "orc_lda_pz": (None, [C.c_int, f64p, f64p, f64p]),
            "orc_cholesky_lower": (C.c_int, [c128p, C.c_int]),
            "orc_jacobi_eigh": (None, [c128p, C.c_int, f64p, c128p]),
            "orc_subspace_matrix": (None, [c128p, c128p, C.c_int, C.c_int, c128p]),
            "orc_gen_eigh": (C.c_int, [c128p, c128p, C.c_int, f64p, c128p]),
            "orc_scf_run": (C.c_int, [C.POINTER(Params), f64p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      f64p, f64p, f64p, vp]),
            "orc_scf_run_ex": (C.c_int, [C.POINTER(Params), f64p, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), f64p, f64p, f64p, vp, vp, vp,
                                         C.POINTER(C.c_double)]),
            "orc_scf_run_k": (C.c_int, [C.POINTER(Params), f64p, C.c_int, f64p, f64p,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), f64p, f64p, f64p,
                                        vp, i32p, C.POINTER(C.c_double)]),
            "orc_fermi_occupations_w": (C.c_double, [f64p, f64p, C.c_int, C.c_double, C.c_double,
                                                     f64p, C.POINTER(C.c_double)]),
            "orc_pulay_create": (vp, [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double]),
            "orc_pulay_mix": (None, [vp, f64p, f64p, f64p]),
            "orc_pulay_destroy": (None, [vp]),
            "orc_line_search": (None, [C.c_double] * 4 + [f64p, C.POINTER(C.c_int)]),
            "orc_compute_residual": (C.c_int, [c128p, c128p, c128p, C.c_int, c128p, C.POINTER(C.c_double)]),
            "orc_residual_bare_norm": (C.c_double, [c128p, C.c_int]),
            "orc_s_normalize": (C.c_int, [c128p, c128p, C.c_int, c128p]),
            "orc_s_orthogonalize_against": (C.c_int, [c128p, c128p, C.c_int, C.c_int, vp, vp, c128p, c128p]),
            "orc_tpa_diag": (None, [f64p, C.c_int, f64p]),
            "orc_uspp_precond_apply": (C.c_int, [c128p, C.c_int, C.c_int, c128p, f64p, c128p, C.c_int,
                                                 c128p, vp, vp]),
            "orc_band_cg_model": (C.c_int, [vp, c128p, C.c_int, C.c_int, C.c_double, C.c_int, f64p,
                                            i32p, i32p]),
            "orc_band_cg_dense": (C.c_int, [c128p, C.c_int, c128p, c128p, c128p, C.c_int, c128p, f64p,
                                            C.c_int, vp, vp, C.c_int, C.c_double, c128p, f64p,
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def downsample(fine, nw):
    """Fine n_f^3 real array -> wave n_w^3 by G-space truncation
    (downsample_array_to_wave_grid, src/scf.rs:413-454)."""
    fine = np.ascontiguousarray(fine, np.float64).reshape(-1)
    nf = round(len(fine) ** (1 / 3))
    out = np.zeros(nw ** 3)
    lib().orc_downsample(fine, nf, nw, out)
    return out


def set_num_threads(n):
    """OpenMP threads of the oracle's parallel loops."""
    lib().orc_set_num_threads(int(n))


def max_threads():
    return int(lib().orc_get_max_threads())


def default_params(**kw):
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "kfrac":
            for i in range(3):
                p.kfrac[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def positions(n_rep, a0=10.26, sigma=0.05, seed=1):
    pos = np.zeros(3 * 8 * n_rep ** 3)
    lib().orc_synth_positions(n_rep, a0, sigma, seed, pos)
    return pos


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Model:
    """Oracle Hamiltonian: cell + plane waves + projectors + V_eff."""

    def __init__(self, params, pos):
        self.params = params
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.h = lib().orc_model_create(C.byref(params), self.pos)
        self.n_pw = lib().orc_model_n_pw(self.h)
        self.n_proj = lib().orc_model_n_proj(self.h)
        self.n_grid = params.n_grid

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_destroy(self.h)
            self.h = None

    @property
    def n3(self):
        return self.n_grid ** 3

    def pw(self):
        mil = np.zeros(3 * self.n_pw, np.int32)
        idx = np.zeros(self.n_pw, np.int32)
        kin = np.zeros(self.n_pw)
        lib().orc_model_get_pw(self.h, mil, idx, kin)
        return mil.reshape(-1, 3), idx, kin

    def vion(self):
        v = np.zeros(self.n3)
        lib().orc_model_get_vion(self.h, v)
        return v

    def beta(self):
        b = np.zeros(self.n_proj * self.n_pw, np.complex128)
        lib().orc_model_get_beta(self.h, b)
        return b.reshape(self.n_proj, self.n_pw)

    def d(self):
        d = np.zeros(self.n_proj)
        lib().orc_model_get_d(self.h, d)
        return d

    def set_d(self, d):
        lib().orc_model_set_d(self.h, np.ascontiguousarray(d, np.float64))

    def set_veff(self, v):
        lib().orc_model_set_veff(self.h, np.ascontiguousarray(v, np.float64))

    def set_sinv_filter(self, on):
        lib().orc_model_set_sinv_filter(self.h, int(on))

    def initial_psi(self, n_b, seed=2):
        psi = np.zeros(n_b * self.n_pw, np.complex128)
        lib().orc_initial_psi(self.h, n_b, seed, psi)
        return psi.reshape(n_b, self.n_pw)

    def apply_h(self, psi, want_s=False):
        psi = np.ascontiguousarray(psi, np.complex128)
        n_b = psi.shape[0]
        h = np.zeros_like(psi)
        s = np.zeros_like(psi) if want_s else None
        lib().orc_apply_h(self.h, psi, n_b, h, _ptr(s))
        return (h, s) if want_s else h

    def apply_sinv(self, x):
        x = np.ascontiguousarray(x, np.complex128).copy()
        lib().orc_apply_sinv(self.h, x, x.shape[0])
        return x

    def grid_lambda_max(self):
        return lib().orc_grid_lambda_max(self.h)

    def lanczos_bound(self, k=10, seed=4):
        rmin = C.c_double()
        b = lib().orc_lanczos_bound(self.h, k, seed, C.byref(rmin))
        return b, rmin.value

    def chebyshev_filter(self, psi, degree, bounds4):
        psi = np.ascontiguousarray(psi, np.complex128)
        out = np.zeros_like(psi)
        diag = np.zeros(4)
        rc = lib().orc_chebyshev_filter(self.h, psi, psi.shape[0], degree,
                                        np.ascontiguousarray(bounds4, np.float64), out, diag)
        return rc, out, diag

    def rayleigh_ritz(self, psi, want_h=False):
        psi = np.ascontiguousarray(psi, np.complex128).copy()
        eps = np.zeros(psi.shape[0])
        h = np.zeros_like(psi) if want_h else None
        rc = lib().orc_rayleigh_ritz(self.h, psi, psi.shape[0], eps, _ptr(h))
        if rc:
            raise RuntimeError(f"oracle rayleigh_ritz failed: {rc}")
        return (psi, eps, h) if want_h else (psi, eps)

    def density(self, psi, occ):
        psi = np.ascontiguousarray(psi, np.complex128)
        occ = np.ascontiguousarray(occ, np.float64)
        rho = np.zeros(self.n3)
        lib().orc_density(self.h, psi, psi.shape[0], occ, rho)
        return rho

    def vhxc(self, rho):
        v = np.zeros(self.n3)
        eh, exc = C.c_double(), C.c_double()
        lib().orc_vhxc(self.h, np.ascontiguousarray(rho, np.float64), v, C.byref(eh), C.byref(exc))
        return v, eh.value, exc.value

    def band_cg(self, psi, max_steps, tol, precond=0):
        psi = np.ascontiguousarray(psi, np.complex128).copy()
        n_b = psi.shape[0]
        eps = np.zeros(n_b)
        steps = np.zeros(n_b, np.int32)
        conv = np.zeros(n_b, np.int32)
        rc = lib().orc_band_cg_model(self.h, psi, n_b, max_steps, tol, precond, eps, steps, conv)
        if rc:
            raise RuntimeError(f"oracle band_cg failed: {rc}")
        return psi, eps, steps, conv

    def screen_d(self):
        d = np.zeros(self.n_proj)
        lib().orc_screen_d(self.h, d)
        return d


def fermi_occupations(eps, n_e, kt):
    eps = np.ascontiguousarray(eps, np.float64)
    occ = np.zeros(len(eps))
    ts = C.c_double()
    mu = lib().orc_fermi_occupations(eps, len(eps), n_e, kt, occ, C.byref(ts))
    return occ, mu, ts.value


def spectral_bounds(eig, lambda_max):
    eig = np.ascontiguousarray(eig if eig is not None else np.zeros(0), np.float64)
    out = np.zeros(4)
    lib().orc_spectral_bounds(eig, len(eig), lambda_max, out)
    return out


def scf_run(params, pos, with_rho=False):
    n_iter, conv = C.c_int(), C.c_int()
    energies = np.zeros(params.max_iter)
    drho = np.zeros(params.max_iter)
    eps = np.zeros(params.n_bands)
    rho = np.zeros(params.n_grid ** 3) if with_rho else None
    rc = lib().orc_scf_run(C.byref(params), np.ascontiguousarray(pos, np.float64), C.byref(n_iter),
                           C.byref(conv), energies, drho, eps, _ptr(rho))
    if rc:
        raise RuntimeError(f"oracle scf failed: {rc}")
    n = n_iter.value
    return dict(n_iter=n, converged=bool(conv.value), energies=energies[:n], drho=drho[:n],
                eps=eps, rho=rho)


def scf_run_ex(params, pos):
    """Full SCF with the final block, occupations and chemical potential."""
    n_iter, conv, mu = C.c_int(), C.c_int(), C.c_double()
    m = Model(params, pos)
    energies = np.zeros(params.max_iter)
    drho = np.zeros(params.max_iter)
    eps = np.zeros(params.n_bands)
    rho = np.zeros(params.n_grid ** 3)
    psi = np.zeros((params.n_bands, m.n_pw), np.complex128)
    occ = np.zeros(params.n_bands)
    rc = lib().orc_scf_run_ex(C.byref(params), np.ascontiguousarray(pos, np.float64),
                              C.byref(n_iter), C.byref(conv), energies, drho, eps, _ptr(rho),
                              _ptr(psi), _ptr(occ), C.byref(mu))
    if rc:
        raise RuntimeError(f"oracle scf failed: {rc}")
    n = n_iter.value
    return dict(n_iter=n, converged=bool(conv.value), energies=energies[:n], drho=drho[:n],
                eps=eps, rho=rho, psi=psi, occ=occ, mu=mu.value, model=m)


def scf_run_k(params, pos, kpts, weights):
    """k-point SCF: kpts (n_k x 3 fractional), weights (n_k, sum 1)."""
    kpts = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
    n_k = len(kpts)
    n_iter, conv, mu = C.c_int(), C.c_int(), C.c_double()
    energies = np.zeros(params.max_iter)
    drho = np.zeros(params.max_iter)
    eps = np.zeros(n_k * params.n_bands)
    rho = np.zeros(params.n_grid ** 3)
    n_pw = np.zeros(n_k, np.int32)
    rc = lib().orc_scf_run_k(C.byref(params), np.ascontiguousarray(pos, np.float64), n_k,
                             kpts.reshape(-1), np.ascontiguousarray(weights, np.float64),
                             C.byref(n_iter), C.byref(conv), energies, drho, eps, _ptr(rho), n_pw,
                             C.byref(mu))
    if rc:
        raise RuntimeError(f"oracle k-point scf failed: {rc}")
    n = n_iter.value
    return dict(n_iter=n, converged=bool(conv.value), energies=energies[:n], drho=drho[:n],
                eps=eps.reshape(n_k, params.n_bands), rho=rho, n_pw=n_pw, mu=mu.value)


def fermi_occupations_w(eps, w, n_e, kt):
    eps = np.ascontiguousarray(eps, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    occ = np.zeros(len(eps))
    ts = C.c_double()
    mu = lib().orc_fermi_occupations_w(eps, w, len(eps), n_e, kt, occ, C.byref(ts))
    return occ, mu, ts.value


class Pulay:
    """The oracle's Pulay/Kerker density mixer on an n^3 grid of edge L."""

    def __init__(self, n_grid, L, depth=6, alpha=0.5, kerker_q=0.8):
        self.n3 = n_grid ** 3
        self.h = lib().orc_pulay_create(n_grid, L, depth, alpha, kerker_q)

    def mix(self, rin, rout):
        out = np.zeros(self.n3)
        lib().orc_pulay_mix(self.h, np.ascontiguousarray(rin, np.float64),
                            np.ascontiguousarray(rout, np.float64), out)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pulay_destroy(self.h)
            self.h = None
