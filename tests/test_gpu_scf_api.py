"""The caller-supplied SCF system (chs_scf_context_create / _run / _destroy,
the reference's planned scf_context_create / scf_run / scf_context_destroy,
PROJECT_ROOT_PLAN.md:880-1029) and the driver's exit paths, through the
C-ABI.

The 8-atom cell of the oracle (36^3 grid, 20 bands), local-only and with
USPP projectors, is handed in as plain arrays (lattice, positions, V_ion,
Miller indices, beta, D0, start wavefunctions): the SCF must match the
oracle's run of the same system at the north-star tolerances (E within
1e-7 Ha/atom, eigenvalues within 1e-8 Ha, density L1 within 1e-6)."""
import ctypes as C

import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu

A0 = 10.26


def oracle_system(npa):
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=40,
              n_proj_per_atom=npa, q0=0.25 if npa else 0.0)
    p = O.default_params(**kw)
    pos = O.positions(1)
    m = O.Model(p, pos)
    mil, idx, kin = m.pw()
    sysd = dict(grid=(36, 36, 36), lattice=A0 * np.eye(3), positions=pos.reshape(-1, 3),
                n_electrons=4.0 * 8, v_ion=m.vion(), n_bands=20, kpts=[[0.0, 0.0, 0.0]],
                weights=[1.0], miller=[mil], proj_per_atom=npa,
                beta=[m.beta()] if npa else None, d0=m.d() if npa else None,
                q0=0.25 if npa else 0.0, q_rc=p.q_rc, psi0=[m.initial_psi(20)])
    return kw, p, pos, sysd


@pytest.mark.parametrize("npa", [0, 4])
def test_context_api_matches_oracle(ctx, npa):
    kw, p, pos, sysd = oracle_system(npa)
    ref = O.scf_run(p, pos, with_rho=True)
    s = chs.ScfContext(ctx, solver=chs.scf_params(**kw), **sysd)
    got = s.run(want_rho=True)
    assert ref["converged"] and got["converged"]
    assert got["n_iter"] == ref["n_iter"]
    assert abs(got["energy"] - ref["energies"][-1]) / 8 < 1e-7
    assert np.abs(got["eps"][0] - ref["eps"]).max() < 1e-8
    dv = A0 ** 3 / 36 ** 3
    assert np.abs(got["rho"] - ref["rho"]).sum() * dv < 1e-6
    assert np.abs(got["energies"] - ref["energies"]).max() / 8 < 1e-7
    # the density criterion beside |dE| (src/scf.rs:357): the oracle's
    # per-iteration integral of |rho_out - rho_in|
    assert np.abs(got["drho_trace"] - ref["drho"]).max() < 1e-6
    # the run leaves its final state (wavefunctions, last output density): a
    # second run starts there and converges within a few iterations to the
    # same fixed point (|dE| < etol = 1e-6 Ha/atom per step bounds how far
    # either run's last energy sits from it)
    again = s.run()
    assert again["converged"] and again["n_iter"] <= 5
    assert abs(again["energy"] - got["energy"]) / 8 < 1e-5
    psi = s.psi(0)
    assert psi.shape == (20, len(sysd["miller"][0]))
    s.close()


def test_context_api_seeded_start_equals_synthetic_run(ctx):
    """Without psi0 the seeded start of seed_psi is generated: the same
    system then gives chs_scf_run's result exactly (same plane waves, same
    start, same arithmetic)."""
    kw, p, pos, sysd = oracle_system(0)
    sysd["psi0"] = None
    s = chs.ScfContext(ctx, solver=chs.scf_params(**kw), **sysd)
    got = s.run()
    ref = chs.run_scf(ctx, chs.scf_params(**kw), pos)
    s.close()
    assert got["n_iter"] == ref["n_iter"]
    assert abs(got["energy"] - ref["energy"]) < 1e-10


def test_context_api_rejects_non_cubic(ctx):
    kw, p, pos, sysd = oracle_system(0)
    sysd["grid"] = (36, 36, 40)
    with pytest.raises(RuntimeError, match="cubic"):
        chs.ScfContext(ctx, solver=chs.scf_params(**kw), **sysd)
    _, _, _, sysd = oracle_system(0)
    lat = A0 * np.eye(3)
    lat[0, 1] = 0.5
    sysd["lattice"] = lat
    with pytest.raises(RuntimeError, match="cubic"):
        chs.ScfContext(ctx, solver=chs.scf_params(**kw), **sysd)


def test_context_outlives_nothing():
    """Destroying the device context first frees the system's device arrays:
    run then reports an invalid handle instead of touching freed memory, and
    destroy frees the host side only."""
    c = chs.Context(0)
    kw, p, pos, sysd = oracle_system(0)
    s = chs.ScfContext(c, solver=chs.scf_params(**kw), **sysd)
    c.close()
    res = chs.ScfResult()
    e = np.zeros(40)
    rc = chs.lib().chs_scf_context_run(s.h, C.byref(res), e.ctypes.data_as(C.POINTER(C.c_double)),
                                       None, None, None)
    assert rc == -9
    assert chs.lib().chs_scf_context_destroy(s.h) == 0
    s.h = None


def test_failed_scf_leaves_nothing_bound():
    """An SCF that stops on an error after binding its arrays (here a
    Lanczos step count outside 1..16, rejected inside the loop) unbinds the
    Hamiltonian before freeing them: a later operator call reports that
    nothing is bound instead of reading freed memory."""
    c = chs.Context(0)
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=5, lanczos_k=17)
    with pytest.raises(RuntimeError, match="Lanczos"):
        chs.run_scf(c, chs.scf_params(**kw), O.positions(1))
    psi = np.zeros((2, 2945), complex)
    buf = chs.Gpu.from_host(c, psi)
    out = chs.Gpu(c, (2, 2945), np.complex128)
    rc = chs.lib().chs_apply_h(c.h, chs.C128, buf.ptr, 2, out.ptr, None)
    assert rc == -9
    assert b"no Hamiltonian bound" in chs.lib().chs_last_error(c.h)
    del buf, out
    c.close()
