"""The 216-atom USPP cell (3x3x3 diamond-cubic, 108^3 grid, 80,725 plane
waves, 864 projectors) at operator level against the FP64 oracle: H|psi>
and S|psi> in both precisions on a few bands (the largest plane and stick
tables the kernels see), every launch shape of the plane kernel, and the
complex64 filter against the complex128 one on the device."""
import os

import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cell():
    p = O.default_params(n_rep=3, n_grid=108, n_bands=512, n_proj_per_atom=4, q0=0.25)
    m = O.Model(p, O.positions(3))
    return m


def bind(ctx, m):
    mil, idx, kin = m.pw()
    assert len(kin) == 80725
    H = chs.Hamiltonian(ctx, 108, mil, kin, m.vion())
    H.bind_projectors(m.beta(), m.d(), np.full(m.n_proj, 0.25))
    return H


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 3e-5)])
def test_apply_h_s_config_c(ctx, cell, dtype, tol):
    H = bind(ctx, cell)
    assert H.n_proj == 864
    psi = cell.initial_psi(3)
    rh, rs = cell.apply_h(psi, want_s=True)
    gh, gs = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi.astype(dtype)), want_s=True)
    gh = gh.sync_to_host().astype(np.complex128)
    gs = gs.sync_to_host().astype(np.complex128)
    assert np.abs(gh - rh).max() / np.abs(rh).max() < tol
    assert np.abs(gs - rs).max() / np.abs(rs).max() < tol


def test_filter_c64_vs_c128_config_c(ctx, cell):
    H = bind(ctx, cell)
    rng = np.random.default_rng(0)
    kin = cell.pw()[2]
    nb = 32
    psi = (rng.standard_normal((nb, len(kin))) + 1j * rng.standard_normal((nb, len(kin)))) \
        / (1.0 + 2.0 * kin)
    up, _ = chs.lanczos_bound(ctx, 10, 4)
    b4 = chs.compute_spectral_bounds(None, up)
    a = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 12, b4).sync_to_host()
    b = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi.astype(np.complex64)), 12,
                             b4).sync_to_host().astype(np.complex128)
    assert np.abs(a - b).max() / np.abs(a).max() < 1e-4


# Plane-kernel launch shapes (workgroup size, row-buffer divisor) the tuner
# may choose; each is forced through a tuning-cache file.
# (sticks per z-pass block, plane-kernel threads, row-buffer divisor): every plane shape
# with 16 sticks per block, and the larger z-pass blocks the tuner also tries where their
# LDS fits (complex128 with 64 sticks does not: the library then takes 16)
SHAPES = {np.complex64: [(16, 256, 1), (16, 256, 2), (16, 512, 1), (16, 512, 2), (16, 512, 4),
                         (16, 1024, 1), (16, 1024, 2), (32, 512, 2), (64, 512, 2)],
          np.complex128: [(16, 256, 1), (16, 256, 2), (16, 512, 1), (16, 512, 2), (32, 256, 2)]}


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 3e-5)])
def test_plane_kernel_every_launch_shape(cell, tune_header, tmp_path, dtype, tol):
    psi = cell.initial_psi(3)
    rh = cell.apply_h(psi)
    prec = chs.C64 if dtype == np.complex64 else chs.C128
    for sb, threads, div in SHAPES[dtype]:
        f = tmp_path / f"tune_{sb}_{threads}_{div}"
        f.write_text(f"{tune_header}\nvloc 108 {prec} {sb} {div} {threads}\n")
        os.environ["CHS_TUNE_CACHE"] = str(f)
        c = chs.Context(0)
        try:
            H = bind(c, cell)
            gh = chs.apply_full_hamiltonian(c, chs.Gpu.from_host(c, psi.astype(dtype)))
            gh = gh.sync_to_host().astype(np.complex128)
            err = np.abs(gh - rh).max() / np.abs(rh).max()
            print(f"{sb} sticks per z block, {threads} threads, row-buffer divisor {div}: "
                  f"{err:.1e}")
            assert err < tol, (sb, threads, div, err)
            del H, gh
        finally:
            del os.environ["CHS_TUNE_CACHE"]
            c.close()
