"""GPU tests of the learned band-CG preconditioner: the device forward /
backward / Adam kernels against the numpy oracle (oracle/mlp.py, itself
pinned by finite differences in test_oracle_mlp.py), the scale table, the
training-pair builder's invariants, and what the trained network gains over
plain TPA in band-CG on held-out samples."""
import numpy as np
import pytest

import chs
import mlp as M
import pyoracle as O
from chs import learned

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_rep,nb,seed", [(1, 20, 1000), (2, 3, 2021)])
def test_initial_psi_matches_oracle(ctx, n_rep, nb, seed):
    """chs_initial_psi (host threads with LCG jump-ahead) gives the oracle's
    sequential start block bit for bit."""
    p = O.default_params(n_rep=n_rep, n_grid=36 * n_rep, n_bands=nb)
    m = O.Model(p, O.positions(n_rep))
    mil, idx, kin = m.pw()
    got = chs.initial_psi(ctx, chs.Gpu.from_host(ctx, kin), nb, seed).sync_to_host()
    assert np.array_equal(got, m.initial_psi(nb, seed))


def _theta(seed=5):
    rng = np.random.default_rng(seed)
    th = M.init_params(11).astype(np.float64) + rng.standard_normal(M.N_PARAMS) * 0.05
    return M.to_bf16(th).astype(np.float32)  # master == bf16 copy: same weights as the oracle


def _pairs(ctx, nb, band_len, seed=3):
    rng = np.random.default_rng(seed)
    n = nb * band_len
    f = np.column_stack([rng.uniform(0, 8, n), rng.uniform(-1, 1, n), rng.uniform(-0.5, 0, n),
                         rng.uniform(0.1, 0.6, n)]).astype(np.float32)
    r = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    e = (-0.3 * r + 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    w = np.repeat(rng.uniform(0.5, 1.5, nb), band_len).astype(np.float32)
    if nb > 2:
        w[band_len:2 * band_len] = 0.0  # an unused band
    p = {"feat": chs.Gpu.from_host(ctx, f), "r": chs.Gpu.from_host(ctx, r),
         "e": chs.Gpu.from_host(ctx, e), "w": chs.Gpu.from_host(ctx, w), "n": n,
         "band_len": band_len}
    return p, (f.astype(np.float64), r.astype(np.complex128), e.astype(np.complex128),
               w.astype(np.float64))


@pytest.mark.parametrize("nb,band_len", [(1, 5), (3, 37), (40, 129)])
def test_loss_grad_vs_oracle(ctx, nb, band_len):
    """Band tiles that end mid-tile, an unused band, one-sample bands."""
    net = chs.Mlp(ctx)
    th = _theta()
    net.params = th
    p, (f, r, e, w) = _pairs(ctx, nb, band_len)
    L, g = net.loss_grad(p)
    Lo, go = M.loss_grad(th.astype(np.float64), f, r, e, w, band_len)
    assert abs(L - Lo) <= 2e-5 * abs(Lo)
    scale = np.abs(go).max()
    assert np.abs(g - go).max() <= 2e-4 * scale, np.abs(g - go).max() / scale


def test_adam_step_vs_oracle(ctx):
    net = chs.Mlp(ctx)
    th = _theta(7)
    net.params = th
    p, _ = _pairs(ctx, 32, 128, 9)
    _, g = net.loss_grad(p)
    net.train(p, 1, 32, 1e-3)  # full batch: the gradient above
    got = net.params
    ref, _, _ = M.adam_step(th.astype(np.float64), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS),
                            g.astype(np.float64), 1, 1e-3)
    assert np.abs(got - ref).max() < 1e-6


def test_untrained_table_is_per_band_tpa(ctx):
    m = O.Model(O.default_params(), O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(ctx, 36, mil, kin, m.vion())
    psi = m.initial_psi(8)
    g = chs.Gpu.from_host(ctx, psi)
    net = chs.Mlp(ctx)  # zero output layer
    t = net.scale_table(g).sync_to_host()
    eb = (np.abs(psi) ** 2 @ kin) / (np.abs(psi) ** 2).sum(1)
    ref = M.tpa(kin[None, :] / eb[:, None]) / eb[:, None]
    assert np.abs(t - ref).max() <= 2e-6 * np.abs(ref).max()


def test_scale_table_vs_oracle(ctx):
    m = O.Model(O.default_params(), O.positions(1))
    mil, idx, kin = m.pw()
    v = m.vion()
    H = chs.Hamiltonian(ctx, 36, mil, kin, v)
    psi = m.initial_psi(6)
    net = chs.Mlp(ctx)
    th = _theta(13)
    net.params = th
    t = net.scale_table(chs.Gpu.from_host(ctx, psi)).sync_to_host()
    eb = (np.abs(psi) ** 2 @ kin) / (np.abs(psi) ** 2).sum(1)
    f = M.features(kin, eb, v.mean(), v.std()).reshape(-1, 4)
    ref = M.scale(th.astype(np.float64), f).reshape(t.shape)
    assert np.abs(t / ref - 1).max() < 2e-4


def test_pairs_invariants(ctx):
    s = learned.Sample(ctx, seed=21, n_bands=16)
    ref, eps_ref = s.converged()
    blk, _ = s.rough(2)
    pairs = chs.mlp_pair_buffers(ctx, 16 * s.H.n_pw)
    used = chs.mlp_pairs(ctx, blk, ref, eps_ref, pairs)
    assert 0 < used <= 16
    w = pairs["w"].sync_to_host().reshape(16, -1)
    r = pairs["r"].sync_to_host().reshape(16, -1).astype(np.complex128)
    e = pairs["e"].sync_to_host().reshape(16, -1).astype(np.complex128)
    f = pairs["feat"].sync_to_host().reshape(16, -1, 4)
    psi = blk.sync_to_host()
    psi = psi / np.linalg.norm(psi, axis=1, keepdims=True)  # S = I here
    # r is orthogonal to psi (eps is the Rayleigh quotient); e has no psi component
    assert np.abs(np.einsum("bg,bg->b", psi.conj(), r)).max() < 1e-5 * np.abs(r).max()
    assert np.abs(np.einsum("bg,bg->b", psi.conj(), e)).max() < 1e-5 * np.abs(e).max()
    # used bands carry weights 1 / (used |e_b|^2): the zero-scale loss is 1
    wb = w[:, 0]
    assert np.allclose(w, wb[:, None])
    assert abs((wb * (np.abs(e) ** 2).sum(1)).sum() - 1.0) < 1e-4
    # features: x = kin / E_b
    kin = s.kin
    assert np.allclose(f[:, :, 0] * np.exp(f[:, :, 1]), kin[None, :], rtol=1e-5, atol=1e-6)
    # a converged block gives (almost) no residual
    pz = chs.mlp_pair_buffers(ctx, 16 * s.H.n_pw)
    chs.mlp_pairs(ctx, ref, ref, eps_ref, pz)
    assert np.abs(pz["r"].sync_to_host()).max() < 1e-3 * np.abs(r).max()


def _host_pairs(pairs):
    n, bl = pairs["n"], pairs["band_len"]
    f = pairs["feat"].sync_to_host()[:n].astype(np.float64)
    r = pairs["r"].sync_to_host()[:n].astype(np.complex128)
    e = pairs["e"].sync_to_host()[:n].astype(np.complex128)
    w = pairs["w"].sync_to_host()[:n].astype(np.float64)
    return f, r, e, w, bl


def test_training_gain_over_tpa(ctx):
    """Train on plain-run pairs of 12 seeded samples; on the pairs of 3
    held-out samples the trained network leaves a smaller error after one
    optimally scaled preconditioned step (the band loss: relative error left
    by a step along s r) than the method without it: per-band TPA (the
    untrained network) and the reference TPA (x = kin / mean PW kinetic
    energy, uspp_preconditioner.rs:148-180). And band CG with the lower-band
    orthogonalisation on the complex overlap (the reference's real part lets
    bands drift), from a common start block of every held-out sample to
    |d eps| < 1e-6 within 200 steps per band, takes fewer steps with the
    trained network than with TPA and converges at least as many bands."""
    train = [learned.Sample(ctx, seed=100 + i, n_bands=24) for i in range(12)]
    held = [learned.Sample(ctx, seed=900 + i, n_bands=24) for i in range(3)]
    npw = train[0].H.n_pw
    pairs = chs.mlp_pair_buffers(ctx, 12 * 2 * 3 * 24 * npw)
    vpairs = chs.mlp_pair_buffers(ctx, 3 * 2 * 3 * 24 * npw)
    assert learned.collect_pairs(ctx, train, pairs) > 0
    assert learned.collect_pairs(ctx, held, vpairs) > 0
    net = chs.Mlp(ctx, seed=11)
    L0, _ = net.loss_grad(pairs)
    V0, _ = net.loss_grad(vpairs)
    net.train(pairs, 2000, 64, 1e-3)
    L1, _ = net.loss_grad(pairs)
    V1, _ = net.loss_grad(vpairs)
    # the reference TPA on the held-out pairs, evaluated with the oracle's loss
    f, r, e, w, bl = _host_pairs(vpairs)
    kin = f[:, 0] * np.exp(f[:, 1])
    s_ref = M.tpa(kin / kin[:bl].mean())
    Vt, _ = M.band_loss(s_ref, r, e, w, bl)
    cg = [learned.evaluate(ctx, s, net, tol=1e-6, max_steps=200, full_overlap=True) for s in held]
    tot = {k: (sum(c[k]["steps"] for c in cg), sum(c[k]["converged"] for c in cg))
           for k in ("tpa", "mlp")}
    print(f"band loss train {L0:.5f} -> {L1:.5f}; held-out: reference TPA {Vt:.5f}, "
          f"per-band TPA {V0:.5f}, trained {V1:.5f}; band CG (steps, converged of 72): {tot}")
    assert L1 < L0
    assert V1 < V0 and V1 < Vt
    assert tot["mlp"][0] < tot["tpa"][0] and tot["mlp"][1] >= tot["tpa"][1]
