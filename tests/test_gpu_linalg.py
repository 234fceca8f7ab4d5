"""GPU parity of the FP64 dense linear algebra and Rayleigh-Ritz: the
reference's device KATs (src/device/blas.rs:259-310, src/device/solver.rs:
113-163), the subspace_diag.rs KATs (:455-918) and the oracle's RR."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu


def fmat(a):
    """numpy (n, m) -> column-major device layout: store a.T row-major."""
    return np.ascontiguousarray(np.asarray(a, np.complex128).T)


def test_zgemm_kat(ctx):
    # A = [[1+i, 0], [0, 1-i]], B = [[1, 2], [3, 4]] -> C = [[1+i, 2+2i], [3-3i, 4-4i]]
    A = np.array([[1 + 1j, 0], [0, 1 - 1j]])
    B = np.array([[1, 2], [3, 4]], complex)
    a, b = chs.Gpu.from_host(ctx, fmat(A)), chs.Gpu.from_host(ctx, fmat(B))
    c = chs.Gpu(ctx, (2, 2), np.complex128).zero()
    chs.zgemm(ctx, "N", "N", 2, 2, 2, 1.0, a, 2, b, 2, 0.0, c, 2)
    got = c.sync_to_host().T
    assert np.abs(got - np.array([[1 + 1j, 2 + 2j], [3 - 3j, 4 - 4j]])).max() < 1e-10


@pytest.mark.parametrize("opa,opb", [("N", "N"), ("C", "N"), ("N", "C"), ("T", "T"), ("C", "T")])
@pytest.mark.parametrize("m,n,k", [(7, 5, 3), (70, 90, 300), (160, 160, 23847), (24000, 160, 160)])
def test_zgemm_random(ctx, opa, opb, m, n, k):
    if m * k > 5e6 and (opa, opb) != ("N", "N") and (opa, opb) != ("C", "N"):
        pytest.skip("large shape covered by the RR layouts")
    rng = np.random.default_rng(m + n + k)
    def op(x, o):
        return x if o == "N" else (x.T if o == "T" else x.conj().T)
    sa = (m, k) if opa == "N" else (k, m)
    sb = (k, n) if opb == "N" else (n, k)
    A = rng.standard_normal(sa) + 1j * rng.standard_normal(sa)
    B = rng.standard_normal(sb) + 1j * rng.standard_normal(sb)
    C0 = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    al, be = 0.7 - 0.2j, 0.3 + 0.1j
    a, b = chs.Gpu.from_host(ctx, fmat(A)), chs.Gpu.from_host(ctx, fmat(B))
    c = chs.Gpu.from_host(ctx, fmat(C0))
    chs.zgemm(ctx, opa, opb, m, n, k, al, a, sa[0], b, sb[0], be, c, m)
    ref = al * op(A, opa) @ op(B, opb) + be * C0
    got = c.sync_to_host().T
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("choice", [0, 1, 2, 3, 4, 5, 0x31, 0x35, 0x23])
def test_zgemm_every_kernel_choice(monkeypatch, choice):
    """Each GEMM kernel the autotuner can pick (variant in the low bits, split-K
    target index in bits 4-5), pinned with CHS_GEMM_CHOICE, on ragged shapes
    (partial tiles in m, n and the last K slice), every pairing of plain,
    transposed and conjugated operands, both precisions."""
    monkeypatch.setenv("CHS_GEMM_CHOICE", str(choice))
    c = chs.Context(0)
    rng = np.random.default_rng(choice)
    def op(x, o):
        return x if o == "N" else (x.T if o == "T" else x.conj().T)
    for opa, opb, m, n, k in (("C", "N", 150, 77, 1303), ("N", "N", 1303, 77, 150),
                              ("N", "C", 97, 130, 211), ("C", "C", 61, 45, 517),
                              ("T", "C", 70, 33, 190), ("C", "T", 45, 66, 271),
                              ("T", "N", 40, 20, 9)):
        sa = (m, k) if opa == "N" else (k, m)
        sb = (k, n) if opb == "N" else (n, k)
        A = rng.standard_normal(sa) + 1j * rng.standard_normal(sa)
        B = rng.standard_normal(sb) + 1j * rng.standard_normal(sb)
        C0 = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        ref = (0.7 - 0.2j) * op(A, opa) @ op(B, opb) + (0.3 + 0.1j) * C0
        for dt, tol in ((np.complex128, 1e-13), (np.complex64, 3e-6)):
            a, b = chs.Gpu.from_host(c, fmat(A).astype(dt)), chs.Gpu.from_host(c, fmat(B).astype(dt))
            cc = chs.Gpu.from_host(c, fmat(C0).astype(dt))
            if dt == np.complex128:
                chs.zgemm(c, opa, opb, m, n, k, 0.7 - 0.2j, a, sa[0], b, sb[0], 0.3 + 0.1j, cc, m)
            else:
                al = np.array([0.7, -0.2], np.float32)
                be = np.array([0.3, 0.1], np.float32)
                c.check(chs.lib().chs_cgemm(c.h, opa.encode(), opb.encode(), m, n, k,
                                            al.ctypes.data_as(chs.C.POINTER(chs.C.c_float)), a.ptr,
                                            sa[0], b.ptr, sb[0],
                                            be.ctypes.data_as(chs.C.POINTER(chs.C.c_float)), cc.ptr, m))
            got = cc.sync_to_host().T
            assert np.abs(got - ref).max() / np.abs(ref).max() < tol, (choice, opa, opb, dt)


def test_zhegvd_kat_identity(ctx):
    # src/device/solver.rs:113-163: A = diag(1,2,3,4), B = I
    a = chs.Gpu.from_host(ctx, fmat(np.diag([1.0, 2, 3, 4]).astype(complex)))
    b = chs.Gpu.from_host(ctx, fmat(np.eye(4, dtype=complex)))
    w = chs.zhegvd(ctx, a, b)
    assert np.abs(w - [1, 2, 3, 4]).max() < 1e-8


def test_zhegvd_kat_generalized(ctx):
    # src/eigensolver/subspace_diag.rs:534-587: H = diag(1..4), S = diag(2,1.5,1,0.8)
    a = chs.Gpu.from_host(ctx, fmat(np.diag([1.0, 2, 3, 4]).astype(complex)))
    b = chs.Gpu.from_host(ctx, fmat(np.diag([2.0, 1.5, 1.0, 0.8]).astype(complex)))
    w = chs.zhegvd(ctx, a, b)
    assert np.abs(w - [0.5, 4 / 3, 3.0, 5.0]).max() < 1e-12


@pytest.mark.parametrize("n", [3, 6, 33, 65, 66, 160, 512, 1000])
def test_zhegvd_random_vs_scipy(ctx, n):
    """Generalised problems against scipy; n = 3, 65 and 66 are the edges of
    the reduction's single-workgroup tail (trailing matrices up to 64 x 64)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Hm = X + X.conj().T
    Y = rng.standard_normal((n + 8, n)) + 1j * rng.standard_normal((n + 8, n))
    Sm = Y.conj().T @ Y
    a, b = chs.Gpu.from_host(ctx, fmat(Hm)), chs.Gpu.from_host(ctx, fmat(Sm))
    w = chs.zhegvd(ctx, a, b)
    ref = sl.eigh(Hm, Sm, eigvals_only=True)
    assert np.abs(w - ref).max() < 1e-9 * np.abs(ref).max()
    V = a.sync_to_host().T
    # invariants V^H S V = I, V^H H V = diag(w) (subspace_diag.rs:680-858)
    assert np.abs(V.conj().T @ Sm @ V - np.eye(n)).max() < 1e-9
    assert np.abs(V.conj().T @ Hm @ V - np.diag(w)).max() < 1e-8 * np.abs(w).max()


@pytest.mark.parametrize("n", [1, 2, 97, 300])
def test_heev_degenerate_spectrum(ctx, n):
    """Standard problem with eigenvalues of multiplicity up to 4 (the
    degenerate shells of a cubic cell): eigenvalues exact to 1e-12 and an
    orthonormal eigenbasis even inside the degenerate subspaces."""
    rng = np.random.default_rng(7 + n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lev = np.sort(np.repeat(np.linspace(-1.0, 2.0, (n + 3) // 4), 4)[:n])
    Hm = (Q * lev) @ Q.conj().T
    Hm = 0.5 * (Hm + Hm.conj().T)
    a = chs.Gpu.from_host(ctx, fmat(Hm))
    w = chs.zhegvd(ctx, a)
    assert np.abs(w - lev).max() < 1e-12
    U = a.sync_to_host().T
    assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-10
    assert np.abs(Hm @ U - U * w).max() < 1e-11


def test_cholesky_failure_past_first_panel(ctx):
    """Blocked Cholesky reports the failing column (1-based) beyond the first
    32-column panel."""
    n = 100
    Sm = np.eye(n, dtype=complex)
    Sm[69, 69] = -1.0
    a = chs.Gpu.from_host(ctx, fmat(np.eye(n, dtype=complex)))
    b = chs.Gpu.from_host(ctx, fmat(Sm))
    with pytest.raises(chs.RayleighRitzFailed) as e:
        chs.zhegvd(ctx, a, b)
    assert e.value.diag.solver_info == 70


def test_zhegvd_not_positive_definite(ctx):
    a = chs.Gpu.from_host(ctx, fmat(np.eye(3, dtype=complex)))
    b = chs.Gpu.from_host(ctx, fmat(np.diag([1.0, -1.0, 1.0]).astype(complex)))
    with pytest.raises(chs.RayleighRitzFailed) as e:
        chs.zhegvd(ctx, a, b)
    assert e.value.diag.solver_info == 2


def model(n_rep=1, n_grid=36, n_bands=20):
    p = O.default_params(n_rep=n_rep, n_grid=n_grid, n_bands=n_bands)
    return O.Model(p, O.positions(n_rep))


@pytest.mark.parametrize("n_rep,n_grid,nb", [(1, 36, 20), (2, 72, 160)])
def test_rayleigh_ritz_vs_oracle(ctx, n_rep, n_grid, nb):
    m = model(n_rep, n_grid, nb)
    mil, idx, kin = m.pw()
    chs.Hamiltonian(ctx, n_grid, mil, kin, m.vion())
    psi = m.initial_psi(nb)
    lmax = m.grid_lambda_max()
    _, y, _ = m.chebyshev_filter(psi, 6, O.spectral_bounds(None, lmax))
    ry, reps = m.rayleigh_ritz(y)
    gy = chs.Gpu.from_host(ctx, y)
    eps, hy = chs.rayleigh_ritz(ctx, gy, want_hy=True)
    assert np.abs(eps - reps).max() < 1e-10 * max(1, np.abs(reps).max())
    Y = gy.sync_to_host()
    # orthonormal Ritz vectors, and H Y consistent with the rotated block
    assert np.abs(Y.conj() @ Y.T - np.eye(nb)).max() < 1e-10
    HY = hy.sync_to_host()
    assert np.abs(np.einsum("ig,ig->i", Y.conj(), HY).real - eps).max() < 1e-9
    # the spanned subspace equals the oracle's: |<y_i|y_i^ref>| ~ 1 for the
    # lowest, well separated Ritz vector
    assert abs(abs(np.vdot(ry[0], Y[0])) - 1) < 1e-8


def test_cholqr(ctx):
    m = model(1, 36, 20)
    mil, idx, kin = m.pw()
    chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    psi = m.initial_psi(20)
    g = chs.Gpu.from_host(ctx, psi)
    chs.cholqr(ctx, g)
    Y = g.sync_to_host()
    assert np.abs(Y.conj() @ Y.T - np.eye(20)).max() < 1e-12
    # same span: projector onto span(psi) leaves Y unchanged
    Q, _ = np.linalg.qr(psi.T)
    assert np.abs(Q @ (Q.conj().T @ Y.T) - Y.T).max() < 1e-10


def test_tuning_cache_on_disk(tmp_path, monkeypatch):
    """CHS_TUNE_CACHE: one context's tuning decisions (H|psi> launch shape,
    GEMM variants, density launch shape) are written to the file and read
    back by the next context, keyed by this library build and device; a
    file of another build (different header) is ignored."""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("CHS_TUNE_CACHE", str(path))
    c1 = chs.Context(0)
    assert c1.tuning_entries() == (0, 0, 0)
    p = O.default_params(n_rep=1, n_grid=36, n_bands=8)
    m = O.Model(p, O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(c1, 36, mil, kin, m.vion())
    chs.autotune_hpsi(c1, 8, chs.C128)
    a = chs.Gpu.from_host(c1, np.ones((40, 300), complex))
    out = chs.Gpu(c1, (40, 40), np.complex128)
    chs.zgemm(c1, "C", "N", 40, 40, 300, 1.0, a, 300, a, 300, 0.0, out, 40)
    psi = chs.Gpu.from_host(c1, m.initial_psi(8))
    chs.density(c1, psi, np.full(8, 2.0), 10.26 ** 3, 36)
    c1.synchronize()
    n1 = c1.tuning_entries()
    assert n1[0] == 1 and n1[1] >= 1 and n1[2] == 1
    del H, psi
    c1.close()
    lines = path.read_text().splitlines()
    assert lines[0].startswith("chs-tune 1 ") and len(lines) >= 1 + sum(n1)
    c2 = chs.Context(0)
    assert c2.tuning_entries() == n1
    c2.close()
    path.write_text("chs-tune 1 0000000000000000 other\nvloc 36 0 16 1 256\n")
    c3 = chs.Context(0)
    assert c3.tuning_entries() == (0, 0, 0)
    c3.close()
