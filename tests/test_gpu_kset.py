"""k-point sets: the graph-captured filter of several k-points (uneven
plane-wave counts, USPP projectors per k) against the FP64 oracle's filter
of each k-point, across replays with new windows and eager work between
them; the autotuner keeps results unchanged."""
import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu

KF = [(0.0, 0.0, 0.0), (0.25, 0.25, 0.25), (0.5, 0.0, 0.25)]


def models(q0=0.25, npa=4):
    out = []
    for kf in KF:
        p = O.default_params(n_rep=1, n_grid=36, n_bands=12, n_proj_per_atom=npa, q0=q0,
                             kfrac=kf)
        out.append(O.Model(p, O.positions(1)))
    return out


def bind(ctx, ms, npa):
    mils, kins = [], []
    for m in ms:
        mil, idx, kin = m.pw()
        mils.append(chs.Gpu.from_host(ctx, mil))
        kins.append(chs.Gpu.from_host(ctx, kin))
    veff = chs.Gpu.from_host(ctx, ms[0].vion())
    if npa:
        betas = [chs.Gpu.from_host(ctx, m.beta()) for m in ms]
        d = chs.Gpu.from_host(ctx, ms[0].d())
        q = chs.Gpu.from_host(ctx, np.full(ms[0].n_proj, 0.25))
        return chs.KSet(ctx, 36, mils, kins, veff, betas, d, q)
    return chs.KSet(ctx, 36, mils, kins, veff)


@pytest.mark.parametrize("npa", [0, 4])
def test_kset_filter_vs_oracle_replays(ctx, npa):
    ms = models(0.25 if npa else 0.0, npa)
    assert len({m.n_pw for m in ms}) > 1  # uneven plane-wave counts
    ks = bind(ctx, ms, npa)
    psis0 = [m.initial_psi(12) for m in ms]
    ups = [m.lanczos_bound(10, 4)[0] for m in ms]
    for rep, scale in enumerate([1.0, 1.3, 0.9]):
        b4 = [O.spectral_bounds(None, u * scale) for u in ups]
        refs = [m.chebyshev_filter(p, 7, b)[1] for m, p, b in zip(ms, psis0, b4)]
        g = [chs.Gpu.from_host(ctx, p) for p in psis0]
        ks.filter(g, 7, np.array(b4))
        for k in range(len(ms)):
            got = g[k].sync_to_host()
            assert np.abs(got - refs[k]).max() / np.abs(refs[k]).max() < 1e-10, (rep, k)
        # eager work between replays: a single-k filter on the same context
        ks.select(1)
        chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psis0[1]), 3, b4[1])
    ks.select(-1)


def test_kset_filter_c64(ctx):
    ms = models()
    ks = bind(ctx, ms, 4)
    psis0 = [m.initial_psi(12) for m in ms]
    b4 = [O.spectral_bounds(None, m.lanczos_bound(10, 4)[0]) for m in ms]
    refs = [m.chebyshev_filter(p, 7, b)[1] for m, p, b in zip(ms, psis0, b4)]
    for _ in range(2):
        g = [chs.Gpu.from_host(ctx, p.astype(np.complex64)) for p in psis0]
        ks.filter(g, 7, np.array(b4))
        for k in range(len(ms)):
            got = g[k].sync_to_host().astype(np.complex128)
            assert np.abs(got - refs[k]).max() / np.abs(refs[k]).max() < 2e-5


def test_autotune_keeps_results(ctx):
    m = O.Model(O.default_params(), O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    psi = m.initial_psi(20)
    b4 = O.spectral_bounds(None, m.lanczos_bound(10, 4)[0])
    before = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 8, b4).sync_to_host()
    sb, rd, th, ms = chs.autotune_hpsi(ctx, 20)
    assert sb in (16, 32, 64) and rd in (1, 2, 4) and th in (256, 512) and ms > 0
    after = chs.chebyshev_filter(ctx, chs.Gpu.from_host(ctx, psi), 8, b4).sync_to_host()
    assert np.abs(after - before).max() / np.abs(before).max() < 1e-13
