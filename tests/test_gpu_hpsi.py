"""GPU parity of the plane-wave sphere, index map, 3D FFT, H|psi>, the
Chebyshev filter and the Lanczos bound against the FP64 oracle."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu


def model(n_rep=1, n_grid=36, n_bands=20, **kw):
    p = O.default_params(n_rep=n_rep, n_grid=n_grid, n_bands=n_bands, **kw)
    return O.Model(p, O.positions(n_rep))


@pytest.mark.parametrize("n_rep,n_grid,kf", [(1, 36, (0, 0, 0)), (2, 72, (0, 0, 0)),
                                             (1, 36, (0.25, 0.25, 0.25)), (1, 36, (0.5, 0, 0.5))])
def test_gsphere_bit_exact(ctx, n_rep, n_grid, kf):
    L = 10.26 * n_rep
    mil, idx, kin = chs.gsphere(ctx, L, 15.0, n_grid, kf)
    n = O.lib().orc_gsphere(L, 15.0, np.array(kf, float), n_grid, 0, None, None, None)
    assert mil.shape[0] == n
    om = np.zeros(3 * n, np.int32); oi = np.zeros(n, np.int32); ok = np.zeros(n)
    O.lib().orc_gsphere(L, 15.0, np.array(kf, float), n_grid, n, om.ctypes.data, oi.ctypes.data,
                        ok.ctypes.data)
    assert np.array_equal(mil.sync_to_host().ravel(), om)
    assert np.array_equal(idx.sync_to_host(), oi)
    assert np.array_equal(kin.sync_to_host().view(np.uint64), ok.view(np.uint64))


def test_index_map_bit_exact_random(ctx):
    rng = np.random.default_rng(7)
    mil = rng.integers(-17, 18, size=(5000, 3)).astype(np.int32)
    got = chs.pw_coords_to_fft_indices(ctx, mil, (72, 64, 36))
    ref = np.zeros(5000, np.int32)
    O.lib().orc_pw_coords_to_fft_indices(mil.ravel().copy(), 5000, 72, 64, 36, ref)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [12, 36, 72, 108])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_fft3d_vs_numpy(ctx, n, dtype, tol):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((2, n, n, n)) + 1j * rng.standard_normal((2, n, n, n))).astype(dtype)
    g = chs.Gpu.from_host(ctx, x)
    chs.fft3d_c2c(ctx, g, -1)
    fwd = g.sync_to_host()
    ref = np.fft.fftn(x.astype(np.complex128), axes=(1, 2, 3))
    assert np.abs(fwd - ref).max() / np.abs(ref).max() < tol
    chs.fft3d_c2c(ctx, g, +1)
    back = g.sync_to_host()
    # unnormalised round trip = n^3 x (src/device/fft.rs:240-281)
    assert np.abs(back / n ** 3 - x).max() / np.abs(x).max() < tol


@pytest.mark.parametrize("dims", [(36, 24, 16), (8, 72, 12), (18, 12, 108)])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_fft3d_non_cubic_vs_numpy(ctx, dims, dtype, tol):
    """plan_batched_c2c(nx, ny, nz) (src/device/fft.rs:181-199) on grids with
    three different edges, x fastest: forward = numpy fftn, inverse round
    trip = N x."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * ny + nz)
    shape = (3, nz, ny, nx)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    g = chs.Gpu.from_host(ctx, x)
    chs.fft3d_c2c(ctx, g, -1)
    ref = np.fft.fftn(x.astype(np.complex128), axes=(1, 2, 3))
    assert np.abs(g.sync_to_host() - ref).max() / np.abs(ref).max() < tol
    chs.fft3d_c2c(ctx, g, +1)
    assert np.abs(g.sync_to_host() / (nx * ny * nz) - x).max() / np.abs(x).max() < tol


@pytest.mark.parametrize("dims", [(36, 36, 36), (24, 16, 12), (18, 36, 8)])
def test_fft3d_z2d(ctx, dims):
    """plan_batched_z2d + z2d (src/device/fft.rs:127-174): the unnormalised
    inverse of a Hermitian-symmetric grid is real; equals numpy's irfftn
    times N; a DC-only input gives a constant grid per batch (the
    reference's z2d KAT, fft.rs:284-401)."""
    nx, ny, nz = dims
    n3 = nx * ny * nz
    rng = np.random.default_rng(n3)
    r = rng.standard_normal((2, nz, ny, nx))
    half = np.fft.rfftn(r, axes=(1, 2, 3))  # (2, nz, ny, nx // 2 + 1)
    out = chs.fft3d_z2d(ctx, chs.Gpu.from_host(ctx, half), nx).sync_to_host()
    ref = np.fft.irfftn(half, s=(nz, ny, nx), axes=(1, 2, 3)) * n3
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(out / n3 - r).max() < 1e-12 * np.abs(r).max()
    dc = np.zeros((3, nz, ny, nx // 2 + 1), complex)
    dc[:, 0, 0, 0] = [1.5, -2.0, 0.25]
    got = chs.fft3d_z2d(ctx, chs.Gpu.from_host(ctx, dc), nx).sync_to_host()
    for b, v in enumerate([1.5, -2.0, 0.25]):
        assert np.abs(got[b] - v).max() < 1e-15


@pytest.mark.parametrize("nf,nw", [(72, 36), (36, 36), (144, 72)])
def test_downsample_to_wave_grid_vs_oracle(ctx, nf, nw):
    """downsample_array_to_wave_grid (src/scf.rs:413-454) on the device
    against the oracle's restatement (itself pinned against numpy and band-
    limited closed forms)."""
    rng = np.random.default_rng(nf)
    fine = rng.standard_normal((nf, nf, nf))
    got = chs.downsample_to_wave_grid(ctx, chs.Gpu.from_host(ctx, fine), (nw, nw, nw)).sync_to_host()
    ref = O.downsample(fine, nw).reshape(nw, nw, nw)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_downsample_to_wave_grid_non_cubic(ctx):
    """Three different edges on both grids, against numpy's FFT."""
    fdims, wdims = (48, 36, 24), (24, 18, 12)  # (x, y, z)
    rng = np.random.default_rng(5)
    fine = rng.standard_normal(fdims[::-1])
    got = chs.downsample_to_wave_grid(ctx, chs.Gpu.from_host(ctx, fine), wdims).sync_to_host()
    g = np.fft.fftn(fine)
    idx = []
    for n, m in zip(wdims[::-1], fine.shape):
        i = np.arange(n)
        f = np.where(i <= n // 2, i, i - n)
        idx.append(np.where(f >= 0, f, m + f))
    w = g[np.ix_(*idx)]
    ref = (np.fft.ifftn(w) * w.size / fine.size).real
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def bind(ctx, m, veff=None):
    mil, idx, kin = m.pw()
    v = m.vion() if veff is None else veff
    return chs.Hamiltonian(ctx, m.n_grid, mil, kin, v)


@pytest.mark.parametrize("n_rep,n_grid,nb", [(1, 36, 20), (2, 72, 8)])
def test_apply_h_fp64(ctx, n_rep, n_grid, nb):
    m = model(n_rep, n_grid, nb)
    H = bind(ctx, m)
    psi = m.initial_psi(nb)
    ref = m.apply_h(psi)
    got = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi)).sync_to_host()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


def test_apply_h_fp32(ctx):
    m = model(1, 36, 20)
    H = bind(ctx, m)
    psi = m.initial_psi(20)
    ref = m.apply_h(psi)
    got = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi.astype(np.complex64))).sync_to_host()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


@pytest.mark.parametrize("n_rep,n_grid,kf", [(1, 36, (0, 0, 0)), (2, 72, (0, 0, 0)),
                                             (1, 36, (0.25, 0.25, 0.25))])
@pytest.mark.parametrize("window", ["1", "0"])
def test_apply_h_plane_window(monkeypatch, n_rep, n_grid, kf, window):
    """The plane kernels that skip each line's zero window (the default on
    these grids) and the general ones (CHS_PLANE_WINDOW=0) both match the
    oracle, in both precisions."""
    monkeypatch.setenv("CHS_PLANE_WINDOW", window)
    c = chs.Context(0)
    nb = 9  # these grids have an odd number of active x columns: the paired
    # y-lines of the window kernels take the zero spare column
    m = model(n_rep, n_grid, nb, kfrac=kf)
    H = bind(c, m)
    psi = m.initial_psi(nb)
    ref = m.apply_h(psi)
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-5)):
        got = chs.apply_full_hamiltonian(c, chs.Gpu.from_host(c, psi.astype(dt))).sync_to_host()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < tol, (dt, err)
    del H


@pytest.mark.parametrize("n_rep,n_grid,kf", [(1, 36, (0, 0, 0)), (2, 72, (0, 0, 0)),
                                             (1, 36, (0.25, 0.25, 0.25))])
def test_apply_h_every_plane_launch_shape(tune_header, tmp_path, n_rep, n_grid, kf):
    """Every launch shape of the plane kernel (workgroup size, row-buffer
    divisor, sticks per block of the z passes), forced through a tuning-cache
    file, against the oracle in both precisions: the register and LDS row
    plans and the V_eff prefetch on these grids' 4 x 9 and 8 x 9 splits."""
    import os
    nb = 5
    m = model(n_rep, n_grid, nb, kfrac=kf)
    psi = m.initial_psi(nb)
    ref = m.apply_h(psi)
    shapes = [(sb, thr, div) for sb in (16, 32, 64) for thr in (256, 512, 1024) for div in (1, 2)]
    for dt, prec, tol in ((np.complex128, chs.C128, 1e-12), (np.complex64, chs.C64, 2e-5)):
        for sb, thr, div in shapes:
            if thr == 1024 and dt == np.complex128:
                continue
            f = tmp_path / f"tune_{prec}_{sb}_{thr}_{div}"
            f.write_text(f"{tune_header}\nvloc {n_grid} {prec} {sb} {div} {thr}\n")
            os.environ["CHS_TUNE_CACHE"] = str(f)
            c = chs.Context(0)
            try:
                H = bind(c, m)
                got = chs.apply_full_hamiltonian(c, chs.Gpu.from_host(c, psi.astype(dt)))
                err = np.abs(got.sync_to_host() - ref).max() / np.abs(ref).max()
                assert err < tol, (dt, sb, thr, div, err)
                del H, got
            finally:
                del os.environ["CHS_TUNE_CACHE"]
                c.close()


def test_apply_h_closed_forms(ctx):
    """V = 0 => H psi = 0.5|G|^2 psi; V = const => shift."""
    m = model(1, 36, 4)
    mil, idx, kin = m.pw()
    psi = m.initial_psi(4)
    for c in (0.0, -0.7):
        H = chs.Hamiltonian(ctx, 36, mil, kin, np.full(36 ** 3, c))
        got = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi)).sync_to_host()
        assert np.abs(got - (kin + c) * psi).max() < 1e-13


def test_chebyshev_filter_fp64(ctx):
    m = model(1, 36, 20)
    H = bind(ctx, m)
    psi = m.initial_psi(20)
    lmax = m.grid_lambda_max()
    b4 = O.spectral_bounds(None, lmax)
    assert np.allclose(chs.compute_spectral_bounds(None, lmax), b4, rtol=0, atol=0)
    rc, ref, _ = m.chebyshev_filter(psi, 10, b4)
    assert rc == 0
    got = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 10, b4).sync_to_host()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-11, err
    assert abs(chs.grid_lambda_max(ctx, 10.26) - lmax) < 1e-12 * abs(lmax)


def test_chebyshev_diverged(ctx):
    """A window far above the spectrum makes every step grow > 10x."""
    m = model(1, 36, 4)
    H = bind(ctx, m)
    psi = m.initial_psi(4)
    b4 = np.array([1000.0, 999.0, 999.5, 0.5])
    rc, _, diag = m.chebyshev_filter(psi, 4, b4)
    assert rc == -2
    with pytest.raises(chs.ChebyshevDiverged) as e:
        chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 4, b4)
    d = e.value.diag
    assert d.iteration == int(diag[0]) == 2
    assert abs(d.growth_factor - diag[3]) < 1e-9 * diag[3]


def test_lanczos_bound(ctx):
    m = model(1, 36, 4)
    H = bind(ctx, m)
    up, lo = chs.lanczos_bound(ctx, 10, 4)
    oup, olo = m.lanczos_bound(10, 4)
    assert abs(up - oup) < 1e-9 * abs(oup)
    assert abs(lo - olo) < 1e-9 * max(1.0, abs(olo))
