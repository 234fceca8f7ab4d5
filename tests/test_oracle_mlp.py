"""Pins the numpy oracle of the learned preconditioner (oracle/mlp.py) by
what the mathematics fixes: bfloat16 rounding known answers, the TPA
polynomial's limits, the zero-output network reducing to per-band TPA, the
loss gradient against central finite differences, and Adam's first step."""
import numpy as np
import pytest

import mlp as M


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.140625, 1e-3],
                    np.float32)
    got = M.to_bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0);
    # 1 + 3 * 2^-9 is above the midpoint: rounds up to 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -7
    assert got[3] == -2.5 and got[4] == 3.140625
    assert abs(got[5] - 1e-3) < 1e-3 * 2.0 ** -8


def test_tpa_limits():
    assert M.tpa(np.array([0.0]))[0] == 1.0
    x = np.array([1e3])
    # t -> 27 + ... ~ 8x^3 / 16x^4 = 1 / (2x) for large x
    assert abs(M.tpa(x)[0] * 2 * x[0] - 1.0) < 2e-3
    xs = np.linspace(0, 10, 101)
    assert np.all(np.diff(M.tpa(xs)) <= 0)


def test_zero_network_is_per_band_tpa():
    th = M.init_params(11)
    kin = np.linspace(0, 15, 50)
    f = M.features(kin, np.array([0.7, 2.0]), -0.3, 0.4).reshape(-1, 4)
    s = M.scale(th, f)
    eb = np.repeat([0.7, 2.0], 50)
    ref = M.tpa(np.tile(kin, 2) / eb) / eb
    assert np.allclose(s, ref, rtol=1e-15, atol=0)


def _random_problem(seed, nb=6, band_len=50):
    rng = np.random.default_rng(seed)
    n = nb * band_len
    th = M.init_params(5).astype(np.float64)
    # give the output layer non-zero weights so every layer's gradient is exercised
    th += rng.standard_normal(th.shape) * 0.05
    th = M.to_bf16(th)  # exactly representable: finite differences see no rounding
    f = np.column_stack([rng.uniform(0, 8, n), rng.uniform(-1, 1, n),
                         rng.uniform(-0.5, 0, n), rng.uniform(0.1, 0.6, n)])
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    e = -0.3 * r + 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    w = np.repeat(rng.uniform(0.5, 1.5, nb), band_len)
    w[band_len:2 * band_len] = 0.0  # one unused band
    return th, f, r, e, w, band_len


def test_band_loss_closed_forms():
    """e = -lambda s r exactly: zero loss; e orthogonal to s r: loss 1 per band."""
    rng = np.random.default_rng(1)
    L_ = 40
    s = rng.uniform(0.1, 2.0, 2 * L_)
    r = rng.standard_normal(2 * L_) + 1j * rng.standard_normal(2 * L_)
    w = np.ones(2 * L_)
    L0, _ = M.band_loss(s, r, -0.7 * s * r, w, L_)
    assert abs(L0) < 1e-14
    L1, _ = M.band_loss(s, r, 1j * s * r, w, L_)  # Re(conj(r) i s r) = 0
    assert abs(L1 - 1.0) < 1e-14


def test_gradient_vs_finite_differences():
    th, f, r, e, w, bl = _random_problem(3)
    L, g = M.loss_grad(th, f, r, e, w, bl)
    assert np.isfinite(L) and 0 < L < 1

    def loss_fp64(theta):
        # same loss without the bf16 rounding (theta moves off the bf16 grid)
        p = M.unpack(theta)
        z1 = f @ p["W1"].T + p["b1"]
        h1 = M.silu(z1)
        h2 = M.silu(h1 @ p["W2"].T + p["b2"])
        h3 = M.silu(h2 @ p["W3"].T + p["b3"])
        o = h3 @ p["w4"] + p["b4"][0]
        s = M.tpa(f[:, 0]) * np.exp(o) / np.exp(f[:, 1])
        nb = len(s) // bl
        a = -np.real(np.conj(r) * e)
        A = (s * a).reshape(nb, bl).sum(1)
        B = (s * s * np.abs(r) ** 2).reshape(nb, bl).sum(1)
        E = (np.abs(e) ** 2).reshape(nb, bl).sum(1)
        used = w.reshape(nb, bl)[:, 0] > 0
        return float(np.sum((1 - A * A / (B * E))[used]) / nb)

    assert abs(loss_fp64(th) - L) < 1e-12 * abs(L)
    rng = np.random.default_rng(0)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [M.N_PARAMS - 1, 0, 256, 8640]])  # b4, W1, b1, w4
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (loss_fp64(tp) - loss_fp64(tm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_adam_first_step():
    th = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 0.0])
    t1, m, v = M.adam_step(th, np.zeros(3), np.zeros(3), g, 1, 1e-2)
    # first step: m_hat = g, v_hat = g^2 -> update lr * g / (|g| + eps)
    assert np.allclose(t1, th - 1e-2 * g / (np.abs(g) + 1e-8))
