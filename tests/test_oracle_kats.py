"""Pins the FP64 oracle (CPU) against the reference's own known-answer tests
(src/eigensolver/{line_search,cg_helpers,uspp_preconditioner,band_cg,
subspace_diag}.rs and src/device/{fft,blas,solver}.rs) and against closed
forms and independent numpy/scipy computations. No GPU needed."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

import pyoracle as O

L = O.lib()


def cz(a):
    return np.ascontiguousarray(a, np.complex128)


# ---------------------------------------------------------------------------
# generator and cell
# ---------------------------------------------------------------------------
def test_lcg_matches_exact_integer_arithmetic():
    # 48-bit LCG of src/eigensolver/subspace_diag.rs:434-447, seed 42
    out = np.zeros(1000)
    L.orc_lcg_uniforms(42, 1000, out)
    s = 42
    for i in range(1000):
        s = (s * 25214903917 + 11) % (1 << 48)
        assert out[i] == s / 2 ** 48


def test_box_muller_statistics():
    z = np.zeros(200000)
    L.orc_lcg_normals(1, len(z), z)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_positions_are_diamond_cubic_without_displacement():
    pos = np.zeros(3 * 64)
    L.orc_synth_positions(2, 10.26, 0.0, 1, pos)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25],
                      [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array([(i, j, k) for i in range(2) for j in range(2) for k in range(2)])
    ref = ((cells[:, None, :] + basis[None]) * 10.26).reshape(-1)
    assert np.array_equal(pos, ref)
    # nearest-neighbour distance of diamond: a sqrt(3)/4
    p = pos.reshape(-1, 3)
    d = np.linalg.norm(p[4] - p[0])
    assert abs(d - 10.26 * np.sqrt(3) / 4) < 1e-12


def test_positions_displacement_scale():
    p0 = np.zeros(3 * 216)
    p1 = np.zeros(3 * 216)
    L.orc_synth_positions(3, 10.26, 0.0, 1, p0)
    L.orc_synth_positions(3, 10.26, 0.05, 1, p1)
    d = p1 - p0
    assert abs(d.std() - 0.05) < 0.005 and abs(d.mean()) < 0.01


# ---------------------------------------------------------------------------
# plane-wave sphere and index map (src/scf.rs:387-401)
# ---------------------------------------------------------------------------
def gs(Lc, n, kf=(0, 0, 0)):
    kf = np.array(kf, float)
    cnt = L.orc_gsphere(Lc, 15.0, kf, n, 0, None, None, None)
    mil = np.zeros(3 * cnt, np.int32)
    idx = np.zeros(cnt, np.int32)
    kin = np.zeros(cnt)
    L.orc_gsphere(Lc, 15.0, kf, n, cnt, mil.ctypes.data, idx.ctypes.data, kin.ctypes.data)
    return mil.reshape(-1, 3), idx, kin


@pytest.mark.parametrize("n_rep,n,count", [(1, 36, 2945), (2, 72, 23847), (3, 108, 80725)])
def test_gsphere_counts_match_brute_force(n_rep, n, count):
    Lc = 10.26 * n_rep
    mil, idx, kin = gs(Lc, n)
    assert len(kin) == count
    b = 2 * np.pi / Lc
    h = np.arange(-30, 31)
    H, K, Lm = np.meshgrid(h, h, h, indexing="ij")
    sel = 0.5 * b * b * (H ** 2 + K ** 2 + Lm ** 2) <= 15.0
    ref = np.stack([H[sel], K[sel], Lm[sel]], 1)  # meshgrid 'ij' order is lexicographic
    assert np.array_equal(mil, ref.astype(np.int32))
    assert np.allclose(kin, 0.5 * b * b * (mil ** 2).sum(1), rtol=1e-15, atol=0)


def test_gsphere_shifted_kpoint():
    Lc = 20.52
    mil, idx, kin = gs(Lc, 72, (0.25, 0.25, 0.25))
    b = 2 * np.pi / Lc
    g = b * (mil + 0.25)
    assert np.all(0.5 * (g ** 2).sum(1) <= 15.0)
    assert len(kin) == 23976  # shifted MP point count


def test_index_map_formula():
    rng = np.random.default_rng(3)
    mil = rng.integers(-17, 18, size=(4000, 3)).astype(np.int32)
    idx = np.zeros(4000, np.int32)
    L.orc_pw_coords_to_fft_indices(mil.ravel().copy(), 4000, 72, 64, 36, idx)
    ix, iy, iz = (mil[:, 0] % 72, mil[:, 1] % 64, mil[:, 2] % 36)
    assert np.array_equal(idx, (ix + 72 * (iy + 64 * iz)).astype(np.int32))


# ---------------------------------------------------------------------------
# FFT (src/device/fft.rs:240-401 conventions)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [12, 36, 72])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    f = x.copy()
    L.orc_fft3d(f, n, -1)
    ref = np.fft.fftn(x.reshape(n, n, n)).ravel()
    assert np.abs(f - ref).max() / np.abs(ref).max() < 1e-13
    L.orc_fft3d(f, n, +1)
    assert np.abs(f / n ** 3 - x).max() < 1e-13  # unnormalised round trip = N x


def test_fft_dc_only_is_constant():
    n = 12
    g = np.zeros(n ** 3, complex)
    g[0] = 3.0
    L.orc_fft3d(g, n, +1)
    assert np.abs(g - 3.0).max() < 1e-15


# ---------------------------------------------------------------------------
# Hamiltonian
# ---------------------------------------------------------------------------
def small_model(ecut=2.0, n_grid=12, n_bands=4):
    p = O.default_params(n_rep=1, n_grid=n_grid, n_bands=n_bands, ecut=ecut)
    return O.Model(p, O.positions(1))


def dense_h(m, veff):
    """H_{GG'} = 0.5|G|^2 delta + V(G - G') from numpy's FFT of V(r)."""
    mil, idx, kin = m.pw()
    n = m.n_grid
    vg = np.fft.fftn(veff.reshape(n, n, n)) / n ** 3  # vg[iz, iy, ix]
    d = mil[:, None, :] - mil[None, :, :]
    Hm = vg[d[..., 2] % n, d[..., 1] % n, d[..., 0] % n]
    return Hm + np.diag(kin)


def test_apply_h_matches_dense_matrix():
    m = small_model()
    rng = np.random.default_rng(0)
    v = rng.standard_normal(m.n3) * 0.3
    m.set_veff(v)
    Hm = dense_h(m, v)
    psi = rng.standard_normal((3, m.n_pw)) + 1j * rng.standard_normal((3, m.n_pw))
    got = m.apply_h(psi)
    assert np.abs(got - psi @ Hm.T).max() < 1e-12


def test_apply_h_closed_forms():
    m = small_model()
    mil, idx, kin = m.pw()
    psi = m.initial_psi(2)
    m.set_veff(np.zeros(m.n3))
    assert np.abs(m.apply_h(psi) - kin * psi).max() < 1e-14
    m.set_veff(np.full(m.n3, -0.7))
    assert np.abs(m.apply_h(psi) - (kin - 0.7) * psi).max() < 1e-14


def test_rayleigh_ritz_full_basis_equals_dense_eigh():
    m = small_model()
    rng = np.random.default_rng(1)
    v = rng.standard_normal(m.n3) * 0.3
    m.set_veff(v)
    Hm = dense_h(m, v)
    nb = m.n_pw
    y = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    _, eps = m.rayleigh_ritz(y)
    assert np.abs(eps - np.linalg.eigvalsh(Hm)).max() < 1e-11


def test_hartree_single_fourier_mode():
    """V_H of rho = cos(G0 x) is 4 pi cos(G0 x) / G0^2."""
    m = small_model(n_grid=36)
    n, Lc = 36, 10.26
    x = np.arange(n) * Lc / n
    g0 = 2 * np.pi / Lc * 3
    rho = np.broadcast_to(np.cos(g0 * x)[None, None, :], (n, n, n)).ravel().copy()
    lib = O.lib()
    v = np.zeros(m.n3)
    eh, exc = C.c_double(), C.c_double()
    # V_Hxc minus the LDA part on a density shifted to stay positive
    shift = 2.0
    lib.orc_vhxc(m.h, rho + shift, v, C.byref(eh), C.byref(exc))
    vx = np.zeros(m.n3)
    e = np.zeros(m.n3)
    lib.orc_lda_pz(m.n3, rho + shift, e, vx)
    assert np.abs((v - vx) - 4 * np.pi * rho / g0 ** 2).max() < 1e-10
    # E_H = Omega/2 * 4 pi |rho(G0)|^2/G0^2 * 2 (two modes, amplitude 1/2)
    assert abs(eh.value - Lc ** 3 / 2 * 4 * np.pi * 2 * 0.25 / g0 ** 2) < 1e-9


def test_vion_is_the_potential_of_gaussian_ions():
    """-Z erf(r)/r is the potential of a normalised Gaussian charge
    exp(-r^2)/pi^1.5 of charge Z: V_ion = -V_H[Z sum_a g(r - R_a)] up to the
    G = 0 constant."""
    p = O.default_params(n_rep=1, n_grid=36)
    pos = O.positions(1)
    m = O.Model(p, pos)
    n, Lc = 36, 10.26
    ax = np.arange(n) * Lc / n
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    rho = np.zeros((n, n, n))
    for a in pos.reshape(-1, 3):
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                for sz in (-1, 0, 1):
                    r2 = ((X - a[0] - sx * Lc) ** 2 + (Y - a[1] - sy * Lc) ** 2 +
                          (Z - a[2] - sz * Lc) ** 2)
                    rho += 4.0 * np.exp(-r2) / np.pi ** 1.5
    vh = np.zeros(m.n3)
    O.lib().orc_vhxc  # V_H via the oracle's Hartree solver: subtract the LDA part
    eh, exc = C.c_double(), C.c_double()
    O.lib().orc_vhxc(m.h, rho.ravel().copy(), vh, C.byref(eh), C.byref(exc))
    vx = np.zeros(m.n3)
    e = np.zeros(m.n3)
    O.lib().orc_lda_pz(m.n3, rho.ravel().copy(), e, vx)
    vh -= vx
    vion = m.vion()
    d = (vion - vion.mean()) + (vh - vh.mean())
    assert np.abs(d).max() < 1e-6 * np.abs(vion).max()
    # the G = 0 term: pi Z rc^2 / Omega per atom
    assert abs(vion.mean() - 8 * np.pi * 4.0 / Lc ** 3) < 1e-12


def test_lda_potential_is_energy_derivative():
    rho = np.logspace(-6, 1, 400)
    e = np.zeros_like(rho)
    v = np.zeros_like(rho)
    O.lib().orc_lda_pz(len(rho), rho, e, v)
    h = rho * 1e-6
    ep, em = np.zeros_like(rho), np.zeros_like(rho)
    d = np.zeros_like(rho)
    O.lib().orc_lda_pz(len(rho), rho + h, ep, d)
    O.lib().orc_lda_pz(len(rho), rho - h, em, d)
    dnum = ((rho + h) * ep - (rho - h) * em) / (2 * h)
    assert np.abs(dnum - v).max() < 1e-7 * np.abs(v).max()
    # the two Perdew-Zunger branches meet at rs = 1 to their published 4-digit
    # precision: -0.1423/(1 + 1.0529 + 0.3334) = -0.05963 vs B + D = -0.0596
    rs1 = 3 / (4 * np.pi)
    e1, e2 = np.zeros(2), np.zeros(2)
    O.lib().orc_lda_pz(2, np.array([rs1 * (1 - 1e-12), rs1 * (1 + 1e-12)]), e1, e2)
    assert abs(abs(e1[0] - e1[1]) - abs(-0.1423 / 2.3863 + 0.0596)) < 1e-9
    # high-density limit: exchange dominates, e_x = -(3/4)(3 rho/pi)^(1/3)
    big = np.array([1e6])
    e, v = np.zeros(1), np.zeros(1)
    O.lib().orc_lda_pz(1, big, e, v)
    ex = -0.75 * (3 * big[0] / np.pi) ** (1 / 3)
    assert abs(e[0] - ex) / abs(ex) < 5e-3


# ---------------------------------------------------------------------------
# Chebyshev filter, spectral bounds, Lanczos (chebyshev.rs:214-254, 626-816)
# ---------------------------------------------------------------------------
def test_spectral_bounds_formulas():
    b = O.spectral_bounds(None, 30.0)
    assert np.allclose(b, [30.0, 10.0, 20.0, 10.0])
    b = O.spectral_bounds(np.array([-1.0, 0.5, 2.0]), 30.0)
    eps_cut = 2.0 + 0.2 * 3.0
    assert np.allclose(b, [30.0, eps_cut, (30 + eps_cut) / 2, (30 - eps_cut) / 2])
    b = O.spectral_bounds(np.array([0.0, 40.0]), 30.0)
    assert abs(b[1] - 28.5) < 1e-14  # capped at 0.95 lambda_max


def test_chebyshev_filter_diagonal_closed_form():
    """V = const makes H diagonal: Y_m = T_m((lambda - c)/e) psi per plane wave."""
    m = small_model(ecut=4.0, n_grid=16)
    mil, idx, kin = m.pw()
    m.set_veff(np.full(m.n3, 0.3))
    psi = m.initial_psi(3)
    b4 = O.spectral_bounds(None, 6.0)
    for deg in (1, 2, 5, 9):
        rc, y, _ = m.chebyshev_filter(psi, deg, b4)
        if rc:
            continue
        x = (kin + 0.3 - b4[2]) / b4[3]
        t = np.polynomial.chebyshev.chebval(x, [0] * deg + [1])
        assert np.abs(y - t * psi).max() < 1e-9 * np.abs(t * psi).max()


def test_chebyshev_divergence_guard():
    m = small_model()
    psi = m.initial_psi(2)
    rc, _, diag = m.chebyshev_filter(psi, 4, [1000.0, 999.0, 999.5, 0.5])
    assert rc == -2 and diag[0] == 2 and diag[3] > 10


def test_lanczos_is_upper_bound():
    m = small_model(ecut=3.0)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(m.n3) * 0.5
    m.set_veff(v)
    lmax = np.linalg.eigvalsh(dense_h(m, v)).max()
    prev = -np.inf
    for k in (4, 6, 8, 10):
        up, lo = m.lanczos_bound(k, 4)
        assert up >= lmax - 1e-10
        prev = up
    assert lo <= lmax


# ---------------------------------------------------------------------------
# dense linear algebra and subspace_diag.rs KATs
# ---------------------------------------------------------------------------
def make_synthetic(nb, npw):
    """subspace_diag.rs:417-447: seed-42 LCG, psi/hpsi uniform in [-0.5, 0.5)."""
    u = np.zeros(4 * nb * npw)
    L.orc_lcg_uniforms(42, len(u), u)
    u = u.reshape(-1, 2)
    psi = (u[: nb * npw, 0] - 0.5) + 1j * (u[: nb * npw, 1] - 0.5)
    hpsi = (u[nb * npw:, 0] - 0.5) + 1j * (u[nb * npw:, 1] - 0.5)
    return psi.reshape(nb, npw), hpsi.reshape(nb, npw)


def sub(x, y):
    nb, npw = x.shape
    m = np.zeros(nb * nb, np.complex128)
    L.orc_subspace_matrix(cz(x), cz(y), npw, nb, m)
    return m.reshape(nb, nb).T  # column-major -> numpy


def test_subspace_matrices_hermitian_and_diagonal():
    psi, hpsi = make_synthetic(5, 10)
    Hs, Ss = sub(psi, hpsi), sub(psi, psi)
    for M in (Hs, Ss):
        assert np.abs(M - M.conj().T).max() < 1e-12
    for i in range(5):
        assert abs(Hs[i, i] - np.vdot(psi[i], hpsi[i]).real) < 1e-14


def geneig(H, S):
    n = H.shape[0]
    eps = np.zeros(n)
    v = np.zeros(n * n, np.complex128)
    rc = L.orc_gen_eigh(cz(H.T), cz(S.T), n, eps, v)
    return rc, eps, v.reshape(n, n).T


def test_generalized_4x4_kat():
    rc, eps, _ = geneig(np.diag([1.0, 2, 3, 4]).astype(complex), np.diag([2.0, 1.5, 1.0, 0.8]).astype(complex))
    assert rc == 0
    assert np.abs(eps - [0.5, 4 / 3, 3.0, 5.0]).max() < 1e-12


def test_generalized_vs_scipy_and_s_orthogonality():
    psi, hpsi = make_synthetic(10, 15)
    Hs, Ss = sub(psi, hpsi), sub(psi, psi)
    rc, eps, V = geneig(Hs, Ss)
    assert rc == 0
    assert np.abs(eps - sl.eigh(Hs, Ss, eigvals_only=True)).max() < 1e-10
    assert np.abs(V.conj().T @ Ss @ V - np.eye(10)).max() < 1e-10
    assert np.all(np.diff(eps) >= -1e-10)


def test_cholesky_failure_reports_column():
    rc, _, _ = geneig(np.eye(3, dtype=complex), np.diag([1.0, -1.0, 1.0]).astype(complex))
    assert rc == 2


def test_jacobi_vs_numpy():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    A = X + X.conj().T
    w = np.zeros(40)
    v = np.zeros(1600, np.complex128)
    L.orc_jacobi_eigh(cz(A.T), 40, w, v)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-12
    V = v.reshape(40, 40).T
    assert np.abs(A @ V - V * w).max() < 1e-11


# ---------------------------------------------------------------------------
# density and occupations
# ---------------------------------------------------------------------------
def test_density_integral_and_single_plane_wave():
    m = small_model(ecut=4.0, n_grid=16, n_bands=6)
    psi, _ = m.rayleigh_ritz(m.initial_psi(6))
    occ = np.array([2, 2, 2, 1.5, 0.5, 0.0])
    rho = m.density(psi, occ)
    omega = 10.26 ** 3
    assert abs(rho.sum() * omega / m.n3 - occ.sum()) < 1e-11
    one = np.zeros((1, m.n_pw), complex)
    one[0, 3] = 1.0
    r1 = m.density(one, [1.0])
    assert np.abs(r1 - 1 / omega).max() < 1e-16


def test_fermi_occupations():
    eps = np.linspace(-1, 1, 20)
    occ, mu, ts = O.fermi_occupations(eps, 16.0, 0.01)
    assert abs(occ.sum() - 16.0) < 1e-10
    assert np.all(np.diff(occ) <= 1e-15) and ts <= 0
    occ0, _, _ = O.fermi_occupations(eps, 16.0, 0.0)
    assert np.array_equal(occ0, np.r_[np.full(8, 2.0), np.zeros(12)])


# ---------------------------------------------------------------------------
# line search (src/eigensolver/line_search.rs:218-393)
# ---------------------------------------------------------------------------
def ls(a, hd, dd, sd):
    out = np.zeros(3)
    st = C.c_int()
    L.orc_line_search(a, hd, dd, sd, out, C.byref(st))
    return out, st.value


def test_line_search_parabola():
    out, st = ls(1.0, 1.0, 3.0, 0.0)
    assert abs(out[0] - 1 / 3) < 1e-12 and abs(out[1] - 2 / 3) < 1e-12
    assert abs(out[2] - 1.0) < 1e-12 and st == 0


def test_line_search_rational():
    # line_search.rs:250-283: step sqrt(2) - 1, E at s_min = 1 - sqrt(2)
    out, st = ls(0.0, -0.5, 1.0, 1.0)
    assert abs(out[0] - (np.sqrt(2) - 1)) < 1e-12
    s_min = 1 - np.sqrt(2)
    assert abs(out[1] - (s_min + s_min * s_min) / (1 + s_min * s_min)) < 1e-12
    assert out[1] < 0 and out[1] < (100 + 1e4) / (1 + 1e4)
    assert abs(out[2] - np.sqrt(1 + out[0] ** 2)) < 1e-12


def test_line_search_energy_decrease_and_clamp():
    out, st = ls(-0.4, -0.15, 1.2, 0.8)
    assert out[1] < -0.4 and st in (0, -1)
    out, st = ls(0.0, -5.0, 0.01, 0.001)
    assert out[0] == 15.0 and st == 1
    out, st = ls(0.0, 0.0, -0.1, 0.0)
    assert st == -1
    out, st = ls(-0.38, -0.2, 1.5, 0.7)
    assert out[1] <= -0.38 and out[2] > 0


def test_line_search_tiny_bd_guard():
    out, st = ls(1.0, 1e-20, 1.0, 1e-20)
    assert st == -2 and out[0] == 0.0


# ---------------------------------------------------------------------------
# CG helpers (src/eigensolver/cg_helpers.rs:234-715)
# ---------------------------------------------------------------------------
def test_residual_kats():
    r = np.zeros(4, complex)
    e = C.c_double()
    psi = cz([1, 0, 0, 0]); hpsi = cz([2, 3, 4, 5]); spsi = cz([1.5, 0.5, 0, 0])
    assert L.orc_compute_residual(psi, hpsi, spsi, 4, r, C.byref(e)) == 0
    assert abs(e.value - 4 / 3) < 1e-14 and abs(np.vdot(r, psi)) < 1e-14
    psi = cz([2, 0, -1, 0]); hpsi = cz([3, 1, 2, 0]); spsi = cz([1, 0.5, 0, 0.5])
    L.orc_compute_residual(psi, hpsi, spsi, 4, r, C.byref(e))
    assert abs(e.value - 2.0) < 1e-14
    psi = cz([1, 0, 1]); hpsi = cz([2, 3, 4]); spsi = cz([1, 0, 1])
    r3 = np.zeros(3, complex)
    L.orc_compute_residual(psi, hpsi, spsi, 3, r3, C.byref(e))
    assert abs(e.value - 3.0) < 1e-14 and abs(np.vdot(r3, psi)) < 1e-14
    assert L.orc_compute_residual(cz([1]), cz([1]), cz([0]), 1, np.zeros(1, complex), C.byref(e)) == -1


def test_norm_and_inner_product_kats():
    assert abs(L.orc_residual_bare_norm(cz([3, 4]), 2) - 5.0) < 1e-15
    assert L.orc_residual_bare_norm(cz(np.zeros(5)), 5) == 0.0
    # <a|b> = -4i (cg_helpers.rs inner_product_with_complex_coefficients)
    a, b = cz([1 + 1j, 2j]), cz([2, 1 - 1j])
    assert abs(np.vdot(a, b) - (-4j)) < 1e-14


def test_s_normalize_kats():
    psi = cz([1, 0.5, 0.25])
    spsi = psi * (0.5 / np.vdot(psi, psi).real)
    out = np.zeros(3, complex)
    assert L.orc_s_normalize(psi, spsi, 3, out) == 0
    assert abs(np.vdot(out, spsi / np.sqrt(0.5)).real - 1) < 1e-14
    psi = cz([1, 0])
    L.orc_s_normalize(psi, psi, 2, out[:2])
    assert np.array_equal(out[:2], psi)
    assert L.orc_s_normalize(cz([1]), cz([-0.5]), 1, np.zeros(1, complex)) == -1


def orthogonalize(psi, spsi, conv):
    n = len(psi)
    k = len(conv)
    cp = cz(np.array([c[0] for c in conv]).ravel()) if k else None
    cs = cz(np.array([c[1] for c in conv]).ravel()) if k else None
    op, osp = np.zeros(n, complex), np.zeros(n, complex)
    rc = L.orc_s_orthogonalize_against(cz(psi), cz(spsi), n, k,
                                       cp.ctypes.data if k else None,
                                       cs.ctypes.data if k else None, op, osp)
    return rc, op, osp


def test_s_orthogonalize_kats():
    conv = [(np.eye(6)[i].astype(complex), np.eye(6)[i].astype(complex)) for i in range(3)]
    rc, op, osp = orthogonalize([0.5, -0.3, 0.2, 0.8, 0.1, -0.4], [0.5, -0.3, 0.2, 0.8, 0.1, -0.4], conv)
    assert rc == 0
    for p, _ in conv:
        assert abs(np.vdot(p, osp)) < 1e-10
    assert abs(np.vdot(op, osp).real - 1) < 1e-12
    s_diag = np.array([2.0, 1.5, 0.8, 1.0, 1.2])
    conv = []
    for i in range(2):
        p = np.zeros(5, complex)
        p[i] = 1 / np.sqrt(s_diag[i])
        conv.append((p, p * s_diag))
    t = np.array([2.0, 1.0, 0.5, 0.3, 0.1], complex)
    rc, op, osp = orthogonalize(t, t * s_diag, conv)
    for p, _ in conv:
        assert abs(np.vdot(p, osp)) < 1e-10
    rc, op, osp = orthogonalize([2, 0], [0.5, 0], [])
    assert abs(np.vdot(op, osp).real - 1) < 1e-14


# ---------------------------------------------------------------------------
# USPP preconditioner (src/eigensolver/uspp_preconditioner.rs:250-520)
# ---------------------------------------------------------------------------
def precond_fixture():
    n_proj, n_pw = 3, 10
    beta = np.zeros((n_pw, n_proj), complex)
    for g in range(n_pw):
        for p in range(n_proj):
            beta[g, p] = (g * 3 + p * 7 + 1) / 100 + 1j * (g * 5 + p * 2 + 3) / 100
    q = np.zeros((n_proj, n_proj), complex)
    for i in range(n_proj):
        for j in range(n_proj):
            q[i, j] = (i + 1) * (j + 1) / 10 + 1j * (0 if i == j else (i + j + 1) / 20)
    for i in range(n_proj):
        for j in range(i):
            avg = (q[i, j] + np.conj(q[j, i])) / 2
            q[i, j] = avg
            q[j, i] = np.conj(avg)
        q[i, i] += n_proj
    kin = np.array([0.5 * (g + 1) ** 2 for g in range(n_pw)])
    return beta, q, kin


def precond_apply(beta, q, kin, res, want_mats=False):
    n_pw, n_proj = beta.shape
    res = cz(np.atleast_2d(res))
    out = np.zeros_like(res)
    r = np.zeros(n_proj * n_proj, complex)
    c = np.zeros(n_proj * n_proj, complex)
    rc = L.orc_uspp_precond_apply(cz(beta), n_pw, n_proj, cz(q), np.ascontiguousarray(kin),
                                  res, res.shape[0], out, r.ctypes.data, c.ctypes.data)
    assert rc == 0
    if want_mats:
        return out, r.reshape(n_proj, n_proj).T, c.reshape(n_proj, n_proj).T
    return out


def test_precond_r_and_c_hermitian_and_contracting():
    beta, q, kin = precond_fixture()
    res = np.array([(i * 3 + 1) / 10 + 1j * (i * 2 + 5) / 10 for i in range(10)])
    v1, R, Cm = precond_apply(beta, q, kin, res, want_mats=True)
    assert np.abs(R - R.conj().T).max() < 1e-12
    assert np.abs(Cm - Cm.conj().T).max() < 1e-12
    v2 = precond_apply(beta, q, kin, v1[0])
    assert np.linalg.norm(v2 - v1) / np.linalg.norm(v1) < 1.0
    e0 = np.zeros(10, complex)
    e0[0] = 1
    assert np.linalg.norm(precond_apply(beta, q, kin, e0)) > 1e-12


def test_precond_reduces_to_tpa_with_zero_beta():
    kin = np.array([0.5 * (g + 1) for g in range(10)])
    res = np.array([(i + 1) * 0.1 + 0.05j * i for i in range(10)])
    out = precond_apply(np.zeros((10, 3), complex), np.eye(3, dtype=complex), kin, res)
    t = np.zeros(10)
    L.orc_tpa_diag(kin, 10, t)
    assert np.abs(out[0] - res * t).max() < 1e-14


def test_tpa_floor():
    kin = np.array([0.5 * g * g for g in range(1, 61)])
    t = np.zeros(60)
    L.orc_tpa_diag(kin, 60, t)
    assert t.min() > 0.01 and np.all(t <= 1.0) and np.all(np.diff(t) < 0)


# ---------------------------------------------------------------------------
# band CG (src/eigensolver/band_cg.rs:474-823)
# ---------------------------------------------------------------------------
def band_cg(h_diag, psi0, max_steps, tol, conv=()):
    n = len(h_diag)
    H = np.diag(h_diag).astype(complex)
    S = np.eye(n, dtype=complex)
    beta = np.zeros((n, 1), complex)
    q = np.eye(1, dtype=complex)
    kin = np.array([0.5 * (g + 1.0) ** 2 for g in range(n)])
    out = np.zeros(n, complex)
    o3 = np.zeros(3)
    ns, cv, stt = C.c_int(), C.c_int(), C.c_int()
    k = len(conv)
    cp = cz(np.array([c[0] for c in conv]).ravel()) if k else None
    cs = cz(np.array([c[1] for c in conv]).ravel()) if k else None
    rc = L.orc_band_cg_dense(cz(psi0), n, cz(H.T), cz(S.T), cz(beta), 1, cz(q), kin, k,
                             cp.ctypes.data if k else None, cs.ctypes.data if k else None,
                             max_steps, tol, out, o3, C.byref(ns), C.byref(cv), C.byref(stt))
    assert rc == 0
    return dict(psi=out, eig=o3[0], res=o3[1], n_steps=ns.value, converged=cv.value,
                step_type="SD" if stt.value == 0 else "CG")


def test_band_cg_sd_first_step():
    h = [1.0 - 0.2 * i for i in range(6)]
    r = band_cg(h, [0.5 ** i for i in range(6)], 1, 1e-10)
    assert r["eig"] > 0 and r["n_steps"] == 1 and r["step_type"] == "SD"


def test_band_cg_monotone_and_bounded():
    h = [0.1 + i * (0.25 + 0.05 * i) for i in range(8)]
    p0 = [1 / np.sqrt(i + 1) for i in range(8)]
    r5 = band_cg(h, p0, 5, 1e-10)
    r10 = band_cg(h, p0, 10, 1e-10)
    assert r10["eig"] <= r5["eig"] + 1e-14
    assert r10["eig"] > h[0] - 1e-5


def test_band_cg_residual_decreases():
    h = [0.3, 0.8, 1.5, 2.4, 3.5, 4.8]
    p0 = [1 - 0.15 * i for i in range(6)]
    r1 = band_cg(h, p0, 1, 1e-14)
    r5 = band_cg(h, p0, 5, 1e-14)
    assert r5["res"] <= r1["res"] * 1.01


def test_band_cg_loose_tolerance_converges_in_one_step():
    r = band_cg([0.5, 1.0, 2.0, 4.0], [1, 1, 1, 1], 10, 100.0)
    assert r["converged"] and r["n_steps"] == 1


def test_band_cg_converged_bands_do_not_alter_independent_band():
    h = [0.2, 0.9, 1.8, 3.1, 4.8]
    p0 = np.zeros(5)
    p0[2] = 1
    conv = [(np.eye(5)[i].astype(complex), np.eye(5)[i].astype(complex)) for i in range(2)]
    r = band_cg(h, p0, 10, 1e-12, conv)
    assert abs(r["eig"] - 1.8) < 1e-10
    assert abs(abs(r["psi"][2]) - 1) < 1e-10


def test_band_cg_converges_to_ground_state():
    h = [0.5, 1.0, 2.0, 4.0, 7.0]
    r = band_cg(h, [1, 0.8, 0.6, 0.4, 0.2], 40, 1e-13)
    assert abs(r["eig"] - 0.5) < 1e-8


def test_band_cg_lower_band_overlap():
    """cg_helpers.rs:188 removes only the real part of <psi_j|S g>. With
    complex bands (the 8-atom cell at Gamma, 12 bands after three degree-8
    filter passes) band CG then lets bands leave the orthogonal complement
    of the lower bands it was handed. With the complex overlap removed
    (precond + 256, an option of this build) every band stays orthogonal
    to them, and the empty bands stay above the eighth Ritz value."""
    m = O.Model(O.default_params(n_rep=1, n_grid=36, n_bands=12), O.positions(1))
    psi = m.initial_psi(12)
    b4 = O.spectral_bounds(None, m.lanczos_bound(10, 4)[0])
    for _ in range(3):
        _, psi, _ = m.chebyshev_filter(psi, 8, b4)
        psi, eps0 = m.rayleigh_ritz(psi)
        b4 = O.spectral_bounds(eps0, b4[0])

    def lower_overlap(p):
        return max((np.abs(psi[:i].conj() @ p[i]) / np.linalg.norm(p[i])).max()
                   for i in range(1, 12))

    p_re, _, _, _ = m.band_cg(psi, 20, 1e-9, 0)
    p_cx, e_cx, _, _ = m.band_cg(psi, 20, 1e-9, 256)
    assert lower_overlap(p_re) > 1e-2
    assert lower_overlap(p_cx) < 1e-10
    assert e_cx[8:].min() > eps0[7]


# ---------------------------------------------------------------------------
# fine -> wave grid resampling (downsample_array_to_wave_grid, scf.rs:413-454)
# ---------------------------------------------------------------------------
def _numpy_downsample(fine, wave_shape):
    """The same truncation written with numpy's FFT (axes z, y, x)."""
    g = np.fft.fftn(fine)
    idx = []
    for n, m in zip(wave_shape, fine.shape):
        i = np.arange(n)
        f = np.where(i <= n // 2, i, i - n)
        idx.append(np.where(f >= 0, f, m + f))
    w = g[np.ix_(*idx)]
    return (np.fft.ifftn(w) * w.size / fine.size).real


@pytest.mark.parametrize("nf,nw", [(24, 12), (36, 18), (36, 36), (72, 36)])
def test_downsample_matches_numpy(nf, nw):
    rng = np.random.default_rng(nf + nw)
    fine = rng.standard_normal((nf, nf, nf))
    got = O.downsample(fine, nw).reshape(nw, nw, nw)
    ref = _numpy_downsample(fine, (nw, nw, nw))
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    if nf == nw:
        assert np.abs(got - fine).max() < 1e-12  # identity on one grid


def test_downsample_band_limited_and_cut_modes():
    """A field whose frequencies lie below the wave grid's Nyquist is sampled
    exactly on the wave grid; a mode above it is removed."""
    nf, nw = 36, 12
    x = np.arange(nf) / nf
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    low = 1.5 + np.cos(2 * np.pi * (2 * xx - 3 * y)) + 0.5 * np.sin(2 * np.pi * (z + 4 * y - xx))
    got = O.downsample(low, nw).reshape(nw, nw, nw)
    xw = np.arange(nw) / nw
    zw, yw, xxw = np.meshgrid(xw, xw, xw, indexing="ij")
    exact = 1.5 + np.cos(2 * np.pi * (2 * xxw - 3 * yw)) + 0.5 * np.sin(2 * np.pi * (zw + 4 * yw - xxw))
    assert np.abs(got - exact).max() < 1e-12
    high = np.cos(2 * np.pi * 9 * xx)  # frequency 9 > 12 / 2
    assert np.abs(O.downsample(high, nw)).max() < 1e-13
