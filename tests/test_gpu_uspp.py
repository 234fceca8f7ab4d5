"""GPU parity of the USPP projector path (H = T + V + beta D beta^H,
S = 1 + beta Q beta^H, S^{-1} by Woodbury) against the FP64 oracle: operator
applications, the generalized Chebyshev filter (S^{-1} H), the Lanczos bound
in the S inner product, Rayleigh-Ritz with S, and a full USPP SCF on the
8-atom cell with 4 projectors per atom."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu

Q0 = 0.25


def model(n_bands=20):
    p = O.default_params(n_rep=1, n_grid=36, n_bands=n_bands, n_proj_per_atom=4, q0=Q0)
    return O.Model(p, O.positions(1))


def bind(ctx, m):
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    H.bind_projectors(m.beta(), m.d(), np.full(m.n_proj, Q0))
    return H


def test_apply_h_and_s(ctx):
    m = model()
    H = bind(ctx, m)
    psi = m.initial_psi(12)
    rh, rs = m.apply_h(psi, want_s=True)
    gh, gs = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi), want_s=True)
    assert np.abs(gh.sync_to_host() - rh).max() / np.abs(rh).max() < 1e-12
    assert np.abs(gs.sync_to_host() - rs).max() / np.abs(rs).max() < 1e-12


def test_apply_h_fp32_projectors(ctx):
    m = model()
    H = bind(ctx, m)
    psi = m.initial_psi(12)
    rh = m.apply_h(psi)
    gh = chs.apply_full_hamiltonian(ctx, chs.Gpu.from_host(ctx, psi.astype(np.complex64)))
    assert np.abs(gh.sync_to_host() - rh).max() / np.abs(rh).max() < 3e-5


def test_generalized_filter_lanczos_rr(ctx):
    m = model()
    H = bind(ctx, m)
    psi = m.initial_psi(20)
    up, lo = chs.lanczos_bound(ctx, 10, 4)
    oup, olo = m.lanczos_bound(10, 4)
    assert abs(up - oup) < 1e-9 * abs(oup)
    b4 = O.spectral_bounds(None, oup)
    rc, ref, _ = m.chebyshev_filter(psi, 8, b4)
    assert rc == 0
    got = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 8, b4)
    g = got.sync_to_host()
    assert np.abs(g - ref).max() / np.abs(ref).max() < 1e-10
    ry, reps = m.rayleigh_ritz(ref)
    eps, _ = chs.rayleigh_ritz(ctx, got)
    assert np.abs(eps - reps).max() < 1e-9 * max(1, np.abs(reps).max())
    # S-orthonormal Ritz vectors
    Y = got.sync_to_host()
    _, SY = m.apply_h(Y, want_s=True)
    assert np.abs(Y.conj() @ SY.T - np.eye(20)).max() < 1e-9


def test_scf_uspp_config_a_vs_oracle(ctx):
    kw = dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10, max_iter=40, n_proj_per_atom=4,
              q0=Q0)
    pos = O.positions(1)
    ref = O.scf_run(O.default_params(**kw), pos, with_rho=True)
    got = chs.run_scf(ctx, chs.scf_params(**kw), pos, want_rho=True)
    assert ref["converged"] and got["converged"]
    assert got["n_iter"] == ref["n_iter"]
    assert abs(got["energy"] - ref["energies"][-1]) / 8 < 1e-7
    assert np.abs(got["eps"] - ref["eps"]).max() < 1e-8
    dv = 10.26 ** 3 / 36 ** 3
    assert np.abs(got["rho"] - ref["rho"]).sum() * dv < 1e-6
    # augmented density integrates to N_e
    assert abs(got["rho"].sum() * dv - 32.0) < 1e-8


def test_filter_fp32_projectors_vs_oracle(ctx):
    """complex64 filter with projectors (coefficient-space recurrences in FP32)
    against the FP64 oracle filter: FP32 accuracy of the degree-8 iterate."""
    m = model()
    H = bind(ctx, m)
    psi = m.initial_psi(20)
    oup, _ = m.lanczos_bound(10, 4)
    b4 = O.spectral_bounds(None, oup)
    rc, ref, _ = m.chebyshev_filter(psi, 8, b4)
    got = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi.astype(np.complex64)), 8, b4)
    g = got.sync_to_host().astype(np.complex128)
    assert np.abs(g - ref).max() / np.abs(ref).max() < 2e-5


def test_filter_projectors_without_s(ctx):
    """Projectors with Q = 0 (S = 1): the coefficient recurrence without the
    Woodbury term, against the oracle."""
    p = O.default_params(n_rep=1, n_grid=36, n_bands=16, n_proj_per_atom=4, q0=0.0)
    m = O.Model(p, O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    H.bind_projectors(m.beta(), m.d(), np.zeros(m.n_proj))
    psi = m.initial_psi(16)
    oup, _ = m.lanczos_bound(10, 4)
    b4 = O.spectral_bounds(None, oup)
    rc, ref, _ = m.chebyshev_filter(psi, 9, b4)
    assert rc == 0
    got = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 9, b4).sync_to_host()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10


def test_filter_fp16x3_projectors_vs_oracle(monkeypatch):
    """The opt-in fp16x3 projector products (CHS_GEMM_FP16X3=1: split fp16
    operands on the matrix cores, 22-bit) keep the complex64 filter within the
    FP32 tolerance of the FP64 oracle, and agree with the FP32 form."""
    m = model()
    psi = m.initial_psi(20)
    oup, _ = m.lanczos_bound(10, 4)
    b4 = O.spectral_bounds(None, oup)
    rc, ref, _ = m.chebyshev_filter(psi, 8, b4)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("CHS_GEMM_FP16X3", flag)
        c = chs.Context(0)  # the switch is read when a context is created
        H = bind(c, m)  # keeps the bound device arrays alive
        x = chs.Gpu.from_host(c, psi.astype(np.complex64))
        got = chs.chebyshev_filter(c, x, 8, b4)
        out[flag] = got.sync_to_host().astype(np.complex128)
        c.synchronize()
        del got, x, H
        c.close()
    scale = np.abs(ref).max()
    assert np.abs(out["1"] - ref).max() / scale < 2e-5
    assert np.abs(out["1"] - out["0"]).max() / scale < 2e-5
