"""GPU parity of the density and of the full SCF (config A: 8-atom
diamond-cubic cell, 36^3 grid, 20 bands, Chebyshev degree 10) against the
FP64 oracle: total energy within 1e-7 Ha/atom, eigenvalues within 1e-8 Ha,
density L1 within 1e-6."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu


def test_density_vs_oracle(ctx):
    p = O.default_params(n_rep=1, n_grid=36, n_bands=20)
    m = O.Model(p, O.positions(1))
    mil, idx, kin = m.pw()
    chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    psi, _ = m.rayleigh_ritz(m.initial_psi(20))
    occ = np.linspace(2.0, 0.0, 20)
    occ[-3:] = 0.0
    ref = m.density(psi, occ)
    omega = 10.26 ** 3
    got = chs.density(ctx, chs.Gpu.from_host(ctx, psi), occ, omega, 36).sync_to_host()
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max() + 1e-16
    # integral of rho = sum of occupations (notes/plans/phase-2/TASKS.md:355)
    assert abs(got.sum() * omega / 36 ** 3 - occ.sum()) < 1e-10


@pytest.mark.parametrize("n_rep,n_grid,nb", [(1, 36, 20), (2, 72, 40)])
@pytest.mark.parametrize("window", ["1", "0"])
def test_density_plane_window(monkeypatch, n_rep, n_grid, nb, window):
    """The zero-window density planes (the default on these grids) and the
    general ones (CHS_PLANE_WINDOW=0) both match the oracle; chunks of 8
    bands with unoccupied bands at the end."""
    monkeypatch.setenv("CHS_PLANE_WINDOW", window)
    c = chs.Context(0)
    p = O.default_params(n_rep=n_rep, n_grid=n_grid, n_bands=nb)
    m = O.Model(p, O.positions(n_rep))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(c, n_grid, mil, kin, m.vion())
    psi = m.initial_psi(nb)
    occ = np.linspace(2.0, 0.1, nb)
    occ[-5:] = 0.0
    ref = m.density(psi, occ)
    omega = (10.26 * n_rep) ** 3
    got = chs.density(c, chs.Gpu.from_host(c, psi), occ, omega, n_grid).sync_to_host()
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max() + 1e-16
    del H


DENSITY_SHAPES = [(256, 1), (512, 1), (512, 2), (512, 4)]  # (threads, row-buffer divisor)


@pytest.mark.parametrize("n_rep,n_grid,nb", [(1, 36, 20), (2, 72, 24)])
def test_density_every_launch_shape(tune_header, tmp_path, n_rep, n_grid, nb):
    """Every launch shape of the zero-window density planes (workgroup size,
    row-buffer divisor), forced through a tuning-cache file, against the
    oracle's density."""
    import os
    p = O.default_params(n_rep=n_rep, n_grid=n_grid, n_bands=nb)
    m = O.Model(p, O.positions(n_rep))
    mil, idx, kin = m.pw()
    psi = m.initial_psi(nb)
    occ = np.linspace(2.0, 0.1, nb)
    occ[-3:] = 0.0
    ref = m.density(psi, occ)
    omega = (10.26 * n_rep) ** 3
    for thr, div in DENSITY_SHAPES:
        f = tmp_path / f"tune_{thr}_{div}"
        f.write_text(f"{tune_header}\ndens {2 * n_grid} {thr | div << 16}\n")
        os.environ["CHS_TUNE_CACHE"] = str(f)
        c = chs.Context(0)
        try:
            H = chs.Hamiltonian(c, n_grid, mil, kin, m.vion())
            got = chs.density(c, chs.Gpu.from_host(c, psi), occ, omega, n_grid).sync_to_host()
            assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max() + 1e-16, (thr, div)
            del H
        finally:
            del os.environ["CHS_TUNE_CACHE"]
            c.close()


def test_complex64_density_shapes_give_identical_scf(tune_header, tmp_path, monkeypatch):
    """The complex64 density planes (the mixed SCF's density) in every launch
    shape: with the GEMM kernel and the H|psi> launch shape pinned, the mixed
    SCF of the 8-atom cell is bit-identical across density shapes (each
    output element sums its bands in the same order) and converges."""
    import os
    monkeypatch.setenv("CHS_GEMM_CHOICE", "1")
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=60, precision=chs.C64)
    pos = O.positions(1)
    runs = []
    for thr, div in DENSITY_SHAPES:
        f = tmp_path / f"tune64_{thr}_{div}"
        f.write_text(f"{tune_header}\nvloc 36 {chs.C64} 16 1 512\nvloc 36 {chs.C128} 16 1 256\n"
                     f"dens {2 * 36 + 1} {thr | div << 16}\ndens {2 * 36} 256\n")
        os.environ["CHS_TUNE_CACHE"] = str(f)
        c = chs.Context(0)
        try:
            r = chs.run_scf(c, chs.scf_params(**kw), pos, want_rho=True)
        finally:
            del os.environ["CHS_TUNE_CACHE"]
            c.close()
        assert r["converged"] and r["n_iter_fp32"] > 0, (thr, div)
        runs.append(r)
    for r, (thr, div) in zip(runs[1:], DENSITY_SHAPES[1:]):
        assert r["n_iter"] == runs[0]["n_iter"], (thr, div)
        assert r["energy"] == runs[0]["energy"], (thr, div)
        assert np.array_equal(r["rho"], runs[0]["rho"]), (thr, div)


def test_density_single_plane_wave(ctx):
    """One plane wave => |psi(r)|^2 is constant = 1/Omega."""
    p = O.default_params(n_rep=1, n_grid=36, n_bands=1)
    m = O.Model(p, O.positions(1))
    mil, idx, kin = m.pw()
    chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    psi = np.zeros((1, m.n_pw), complex)
    psi[0, 17] = 1.0
    omega = 10.26 ** 3
    got = chs.density(ctx, chs.Gpu.from_host(ctx, psi), [1.0], omega, 36).sync_to_host()
    assert np.abs(got - 1.0 / omega).max() < 1e-15


def test_scf_config_a_vs_oracle(ctx):
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=40)
    pos = O.positions(1)
    ref = O.scf_run(O.default_params(**kw), pos, with_rho=True)
    got = chs.run_scf(ctx, chs.scf_params(**kw), pos, want_rho=True)
    assert ref["converged"] and got["converged"]
    assert got["n_iter"] == ref["n_iter"]
    n_at = 8
    assert abs(got["energy"] - ref["energies"][-1]) / n_at < 1e-7
    assert np.abs(got["eps"] - ref["eps"]).max() < 1e-8
    omega = 10.26 ** 3
    dv = omega / 36 ** 3
    assert np.abs(got["rho"] - ref["rho"]).sum() * dv < 1e-6
    assert np.abs(got["energies"] - ref["energies"]).max() / n_at < 1e-7


@pytest.mark.parametrize("npa", [0, 4])
def test_scf_mixed_precision_matches_fp64(ctx, npa):
    """complex64 filter with FP64 Rayleigh-Ritz, back to complex128 once
    |dE|/atom < max(2 etol, 3e-7) (at least fp64_final = 2 FP64 iterations),
    at etol = 1e-6: the energy within 1e-7 Ha/atom of the plain FP64 SCF at the
    same tolerance (which the tests above pin to the oracle), and eigenvalues
    and density no further from the tightly converged FP64 solution
    (etol = 1e-8) than the plain FP64 run at etol = 1e-6 is."""
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=100,
              n_proj_per_atom=npa, q0=0.25 if npa else 0.0)
    pos = O.positions(1)
    r64 = chs.run_scf(ctx, chs.scf_params(**kw), pos, want_rho=True)
    r32 = chs.run_scf(ctx, chs.scf_params(**kw, precision=chs.C64), pos, want_rho=True)
    rt = chs.run_scf(ctx, chs.scf_params(**kw, etol=1e-8), pos, want_rho=True)
    assert r64["converged"] and r32["converged"] and rt["converged"]
    assert r64["n_iter_fp32"] == 0
    assert r32["n_iter_fp32"] >= 3 and r32["n_iter"] - r32["n_iter_fp32"] >= 2
    dv = 10.26 ** 3 / 36 ** 3

    def err(r):
        return (abs(r["energy"] - rt["energy"]) / 8, np.abs(r["eps"] - rt["eps"]).max(),
                np.abs(r["rho"] - rt["rho"]).sum() * dv)

    e64, e32 = err(r64), err(r32)
    print(f"npa={npa}: iterations {r32['n_iter']} ({r32['n_iter_fp32']} FP32) vs "
          f"{r64['n_iter']} (tight {rt['n_iter']}); distance to the tight solution "
          f"(E/atom, eps, rho L1): FP64 {e64[0]:.2e} {e64[1]:.2e} {e64[2]:.2e}, "
          f"mixed {e32[0]:.2e} {e32[1]:.2e} {e32[2]:.2e}")
    assert r32["n_iter"] <= r64["n_iter"] + 3
    assert abs(r32["energy"] - r64["energy"]) / 8 < 1e-7
    assert e32[1] <= 3 * e64[1] + 1e-7
    assert e32[2] <= 3 * e64[2] + 1e-6
