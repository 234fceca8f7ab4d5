"""GPU parity of the batched band-CG refinement against the oracle's
band_cg_minimize (src/eigensolver/band_cg.rs:157-437) run band by band with
the same converged sets, for the TPA and the TPA + USPP preconditioners."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu


def start_block(m, nb, passes=2, degree=6):
    psi = m.initial_psi(nb)
    b4 = O.spectral_bounds(None, m.lanczos_bound(10, 4)[0])
    for _ in range(passes):
        _, psi, _ = m.chebyshev_filter(psi, degree, b4)
        psi, eps = m.rayleigh_ritz(psi)
        b4 = O.spectral_bounds(eps, b4[0])
    return psi


# precond + 256: lower-band orthogonalisation with the complex overlap
@pytest.mark.parametrize("uspp,precond", [(False, 0), (True, 0), (True, 1), (False, 256),
                                          (True, 257)])
def test_band_cg_vs_oracle(ctx, uspp, precond):
    kw = dict(n_rep=1, n_grid=36, n_bands=12)
    if uspp:
        kw.update(n_proj_per_atom=4, q0=0.25)
    m = O.Model(O.default_params(**kw), O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    if uspp:
        H.bind_projectors(m.beta(), m.d(), np.full(m.n_proj, 0.25))
    psi0 = start_block(m, 12)
    rpsi, reps, rsteps, rconv = m.band_cg(psi0, 6, 1e-9, precond)
    g = chs.Gpu.from_host(ctx, psi0)
    eps, steps, conv = chs.band_cg_minimize(ctx, g, 6, 1e-9, precond & 255,
                                            full_overlap=bool(precond & 256))
    assert np.array_equal(steps, rsteps)
    assert np.array_equal(conv, rconv)
    assert np.abs(eps - reps).max() < 1e-10
    gp = g.sync_to_host()
    for b in range(12):
        assert abs(abs(np.vdot(gp[b], rpsi[b])) / np.vdot(rpsi[b], rpsi[b]).real - 1) < 1e-8
