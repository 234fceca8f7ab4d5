"""GPU SCF parity with the FP64 oracle at the larger cells, against golden
fixtures the oracle wrote (tests/golden/make_scf_fixtures.py):

  64-atom cell (72^3, 160 bands), full SCF;
  216-atom USPP cell (108^3, 512 bands), the first SCF iteration;
  8-atom USPP cell on the Gamma-centred {0, 1/2}^3 k-point set, full SCF;
  64-atom USPP cell on that set (8 k-points x 160 bands), the first
  iteration.

Tolerances of the complex128 form (north star): every iteration's energy
within 1e-7 Ha/atom, eigenvalues within 1e-8 Ha, density L1 within 1e-6
(estimated from the fixture's sub-grid), identical iteration counts. The
complex64-filter form's error against the same fixture is printed and bounded
by the looser tolerance its FP32 arithmetic allows."""
import ast
import os

import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
A0 = 10.26


def load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    f = np.load(path, allow_pickle=False)
    kw = dict(ast.literal_eval(str(f["params"])))
    return f, kw


def rho_l1(f, rho, kw):
    n, s = kw["n_grid"], int(f["stride"])
    sub = rho.reshape(n, n, n)[::s, ::s, ::s]
    dv = (kw["n_rep"] * A0) ** 3 / n ** 3
    return float(np.abs(sub - f["rho_sub"]).sum() * dv * s ** 3)


def check(f, kw, got, n_at, e_tol=1e-7, eps_tol=1e-8, rho_tol=1e-6, same_iters=True):
    ref_e = f["energies"]
    if same_iters:
        assert got["n_iter"] == int(f["n_iter"])
        assert np.abs(got["energies"] - ref_e).max() / n_at < e_tol
    de = abs(got["energy"] - ref_e[-1]) / n_at
    deps = float(np.abs(np.asarray(got["eps"]).reshape(-1) - f["eps"].reshape(-1)).max())
    drho = rho_l1(f, got["rho"], kw)
    assert de < e_tol and deps < eps_tol and drho < rho_tol, (de, deps, drho)
    return de, deps, drho


def test_scf_64atom_full_vs_oracle(ctx):
    f, kw = load("scf_64atom_full.npz")
    got = chs.run_scf(ctx, chs.scf_params(**kw), O.positions(2), want_rho=True)
    assert bool(f["converged"]) and got["converged"]
    de, deps, drho = check(f, kw, got, 64)
    print(f"64 atoms: {got['n_iter']} iterations; E {de:.1e} Ha/atom, eps {deps:.1e} Ha, "
          f"rho L1 {drho:.1e}")


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_scf_216atom_first_iteration_vs_oracle(ctx, precision):
    f, kw = load("scf_216atom_first.npz")
    prec = chs.C64 if precision == "c64" else chs.C128
    got = chs.run_scf(ctx, chs.scf_params(**kw, precision=prec), O.positions(3), want_rho=True)
    if prec == chs.C128:
        de, deps, drho = check(f, kw, got, 216)
    else:
        # complex64 filter, FP64 Rayleigh-Ritz: FP32 rounding of the filtered
        # block reaches the eigenvalues and density of the first iteration,
        # whose subspace is far from converged; the summation order of the
        # tuned projector GEMMs alone moves it by up to 8e-6 Ha/atom in the
        # energy and 1.2e-5 Ha in the eigenvalues (measured over kernel
        # choices), so the energy bound matches the eigenvalue bound
        de, deps, drho = check(f, kw, got, 216, e_tol=2e-5, eps_tol=2e-5, rho_tol=1e-3)
    print(f"216 atoms USPP, first iteration, {precision}: E {de:.1e} Ha/atom, eps {deps:.1e} Ha, "
          f"rho L1 {drho:.1e}")


def test_kpoint_scf_8atom_vs_oracle(ctx):
    f, kw = load("scf_8atom_kpts.npz")
    kp = f["kpts"]
    got = chs.run_scf_k(ctx, chs.scf_params(**kw), O.positions(1), kp,
                        np.full(len(kp), 1.0 / len(kp)), want_rho=True)
    assert bool(f["converged"]) and got["converged"]
    assert len(set(f["n_pw"].tolist())) > 1  # uneven plane-wave counts
    de, deps, drho = check(f, kw, got, 8)
    assert abs(got["mu"] - float(f["mu"])) < 1e-8
    assert np.abs(got["drho_trace"] - f["drho"]).max() < 1e-6
    print(f"8 atoms x {len(kp)} k: {got['n_iter']} iterations; E {de:.1e}, eps {deps:.1e}, "
          f"rho L1 {drho:.1e}")


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_kpoint_64atom_first_iteration_vs_oracle(ctx, precision):
    """Config D's shape: 64-atom USPP cell, 8 k-points x 160 bands, every
    k-point's filter in one graph launch, per-k Rayleigh-Ritz, one Fermi
    level."""
    f, kw = load("scf_64atom_kpts_first.npz")
    kp = f["kpts"]
    prec = chs.C64 if precision == "c64" else chs.C128
    got = chs.run_scf_k(ctx, chs.scf_params(**kw, precision=prec), O.positions(2), kp,
                        np.full(len(kp), 1.0 / len(kp)), want_rho=True)
    if prec == chs.C128:
        de, deps, drho = check(f, kw, got, 64)
    else:
        # complex64 first iteration: bounds as for the 216-atom cell above
        de, deps, drho = check(f, kw, got, 64, e_tol=2e-5, eps_tol=2e-5, rho_tol=1e-3)
    print(f"64 atoms USPP x {len(kp)} k, first iteration, {precision}: E {de:.1e} Ha/atom, "
          f"eps {deps:.1e} Ha, rho L1 {drho:.1e}")
