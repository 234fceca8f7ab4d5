"""Pins the oracle's USPP and SCF-glue functions (CPU, no GPU) against
routes that share no code or formula with them:

- projectors beta_lm(k+G) (oracle/system.cpp build_projectors) against a
  radial quadrature of r^l exp(-r^2/s^2) Y_lm with spherical Bessel
  functions, and against the same functions sampled in real space (periodic
  images summed) and Fourier transformed with numpy;
- the nonlocal H and the overlap S (oracle/hamiltonian.cpp apply_h) against
  per-ion loops in the form of the reference's gate test
  (tests/phase0_gate1_consistency.rs:320-385);
- S^-1 by the Woodbury identity (apply_sinv): S (S^-1 x) = x;
- screened D (oracle/scf.cpp screen_d) against a real-space sum of V_eff
  times the periodic Gaussian augmentation function;
- the augmentation charge in the density against real-space Gaussians with
  weights sum_b f_b q |<beta|psi_b>|^2, and the integral of rho against
  sum_b f_b <psi_b|S|psi_b>;
- the SCF energy expression (Harris form sum f eps - int V_Hxc rho + E_H +
  E_xc - TS) at self-consistency against the direct Kohn-Sham functional
  T_s + E_NL[D0] + int V_ion rho + E_H + E_xc - TS;
- Pulay mixing against the Lagrange closed form of the constrained
  least-squares DIIS problem, Kerker filtering on single Fourier modes;
- Fermi occupations with k weights against the unweighted form and the
  electron count.
"""
import numpy as np
import pytest
from scipy import integrate, special

import pyoracle as O

A0 = 10.26


def model(**kw):
    p = O.default_params(**kw)
    pos = O.positions(p.n_rep)
    return O.Model(p, pos), p, pos.reshape(-1, 3)


def gvecs(m, p):
    mil, idx, kin = m.pw()
    L = p.n_rep * A0
    g = 2 * np.pi / L * (mil + np.array(p.kfrac[:]))
    return g, idx, kin


# ---------------------------------------------------------------------------
# projectors
# ---------------------------------------------------------------------------
def radial_beta(l, s, G):
    """4 pi (-i)^l int r^2 r^l exp(-r^2/s^2) j_l(G r) dr by adaptive quadrature."""
    f = lambda r: r ** (2 + l) * np.exp(-(r / s) ** 2) * special.spherical_jn(l, G * r)
    v, err = integrate.quad(f, 0, 12 * s, limit=400, epsabs=1e-14, epsrel=1e-13)
    return 4 * np.pi * (-1j) ** l * v


@pytest.mark.parametrize("kfrac", [(0, 0, 0), (0.5, 0.25, 0)])
def test_beta_matches_radial_quadrature(kfrac):
    m, p, pos = model(n_proj_per_atom=4, q0=0.25, kfrac=kfrac)
    g, _, _ = gvecs(m, p)
    beta = m.beta()
    L = p.n_rep * A0
    omega = L ** 3
    s = p.beta_rc
    rng = np.random.default_rng(5)
    sel = rng.choice(m.n_pw, 40, replace=False)
    y00 = 1 / np.sqrt(4 * np.pi)
    y1 = np.sqrt(3 / (4 * np.pi))
    for a in (0, 5):
        for gi in sel:
            G = g[gi]
            Gn = np.linalg.norm(G)
            ph = np.exp(-1j * G @ pos[a])
            r0 = radial_beta(0, s, Gn)
            want0 = r0 * y00 * ph / np.sqrt(omega)
            assert abs(beta[4 * a, gi] - want0) < 1e-11 * max(1.0, abs(want0))
            r1 = radial_beta(1, s, Gn) if Gn > 0 else 0.0
            for c in range(3):
                want = r1 * y1 * (G[c] / Gn if Gn > 0 else 0.0) * ph / np.sqrt(omega)
                assert abs(beta[4 * a + 1 + c, gi] - want) < 1e-11 * max(1.0, abs(want))


def periodic_gaussian(n, L, center, rc, power=None, axis=None):
    """sum over periodic images of r^l exp(-r^2/rc^2) Y (l = 0: 1, l = 1: x_axis)
    sampled on the n^3 grid, array [iz, iy, ix]."""
    x = np.arange(n) * L / n
    out = np.zeros((n, n, n))
    for ia in (-1, 0, 1):
        for ja in (-1, 0, 1):
            for ka in (-1, 0, 1):
                d = [x - center[0] + ia * L, x - center[1] + ja * L, x - center[2] + ka * L]
                r2 = d[2][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[0][None, None, :] ** 2
                f = np.exp(-r2 / rc ** 2)
                if axis is not None:
                    comp = [d[2][:, None, None], d[1][None, :, None], d[0][None, None, :]][2 - axis]
                    f = f * comp
                out += np.broadcast_to(f, (n, n, n))
    return out


def test_beta_matches_real_space_projector():
    """beta_{Ip}(G) = Omega^-1/2 int beta_p(r - R_I) e^{-iG.r} d^3r from the
    real-space functions (Y00 e^{-r^2/s^2}, sqrt(3/4pi) x e^{-r^2/s^2}) on the
    grid, Fourier transformed with numpy: pins normalisation, the e^{-iG.R}
    phase, the -i of l = 1 and the x, y, z order."""
    m, p, pos = model(n_proj_per_atom=4, q0=0.25)
    mil, idx, _ = m.pw()
    n = p.n_grid
    L = A0
    omega = L ** 3
    beta = m.beta()
    s = p.beta_rc
    for a in (1, 6):
        funcs = [periodic_gaussian(n, L, pos[a], s) / np.sqrt(4 * np.pi)]
        funcs += [np.sqrt(3 / (4 * np.pi)) * periodic_gaussian(n, L, pos[a], s, axis=c)
                  for c in range(3)]
        for j, f in enumerate(funcs):
            # forward transform e^{-iG.r}: numpy fftn; integral = Omega / N^3 sum
            coef = np.fft.fftn(f).reshape(-1) * (omega / n ** 3) / np.sqrt(omega)
            got = beta[4 * a + j]
            want = coef[idx]
            assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


# ---------------------------------------------------------------------------
# nonlocal H and S per ion
# ---------------------------------------------------------------------------
def test_nonlocal_h_and_s_per_ion_loops():
    m, p, pos = model(n_proj_per_atom=4, q0=0.25)
    beta = m.beta()
    n_at, npa = 8, 4
    rng = np.random.default_rng(7)
    d = rng.uniform(0.5, 2.0, n_at * npa)
    m.set_d(d)
    psi = m.initial_psi(3, seed=11)
    h_full, s_full = m.apply_h(psi, want_s=True)
    m.set_d(np.zeros_like(d))
    h_loc = m.apply_h(psi)
    nl = np.zeros_like(psi)
    sv = psi.copy()
    for I in range(n_at):
        for i in range(npa):
            bi = beta[I * npa + i]
            for j in range(npa):
                bj = beta[I * npa + j]
                dij = d[I * npa + i] if i == j else 0.0
                qij = p.q0 if i == j else 0.0
                for b in range(psi.shape[0]):
                    c = np.sum(np.conj(bj) * psi[b])
                    nl[b] += bi * dij * c
                    sv[b] += bi * qij * c
    assert np.abs((h_full - h_loc) - nl).max() < 1e-12 * np.abs(nl).max()
    assert np.abs(s_full - sv).max() < 1e-13 * np.abs(sv).max()


@pytest.mark.parametrize("kfrac", [(0, 0, 0), (0.5, 0.5, 0.5)])
def test_sinv_inverts_s(kfrac):
    m, p, pos = model(n_proj_per_atom=4, q0=0.25, kfrac=kfrac)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, m.n_pw)) + 1j * rng.standard_normal((4, m.n_pw))
    y = m.apply_sinv(x)
    _, sy = m.apply_h(y, want_s=True)
    assert np.abs(sy - x).max() < 1e-12 * np.abs(x).max()
    # and S^-1 differs from the identity (the check is not vacuous)
    assert np.abs(y - x).max() > 1e-3 * np.abs(x).max()


# ---------------------------------------------------------------------------
# screened D and augmentation charge
# ---------------------------------------------------------------------------
def gauss_norm(n, L, center, rq):
    return periodic_gaussian(n, L, center, rq) / (np.pi * rq ** 2) ** 1.5


def test_screened_d_real_space_integral():
    m, p, pos = model(n_proj_per_atom=4, q0=0.25)
    n = p.n_grid
    L = A0
    rng = np.random.default_rng(2)
    veff = m.vion() + 0.05 * np.cos(2 * np.pi * 2 * np.arange(n ** 3) / n) \
        + 0.01 * rng.standard_normal(n ** 3)
    m.set_veff(veff)
    got = m.screen_d()
    d0 = m.d()
    v3 = veff.reshape(n, n, n)
    dv = L ** 3 / n ** 3
    for a in range(8):
        integral = np.sum(v3 * gauss_norm(n, L, pos[a], p.q_rc)) * dv
        for j in range(4):
            assert abs(got[4 * a + j] - (d0[4 * a + j] + p.q0 * integral)) < 1e-12


def test_augmentation_charge_real_space():
    m, p, pos = model(n_proj_per_atom=4, q0=0.25)
    n = p.n_grid
    L = A0
    omega = L ** 3
    _, idx, _ = m.pw()
    psi = m.initial_psi(5, seed=9)
    occ = np.array([2.0, 2.0, 1.5, 0.5, 0.0])
    rho = m.density(psi, occ)
    # plain part: inverse transform e^{+iG.r} unnormalised
    plain = np.zeros(n ** 3)
    for b in range(5):
        g = np.zeros(n ** 3, np.complex128)
        g[idx] = psi[b]
        pr = np.fft.ifftn(g.reshape(n, n, n)).reshape(-1) * n ** 3
        plain += occ[b] * np.abs(pr) ** 2 / omega
    beta = m.beta()
    w = np.conj(beta) @ psi.T  # P x n_b
    aug = np.zeros((n, n, n))
    for a in range(8):
        wa = p.q0 * np.sum(occ[None, :] * np.abs(w[4 * a:4 * a + 4]) ** 2)
        aug += wa * gauss_norm(n, L, pos[a], p.q_rc)
    assert np.abs(rho - plain - aug.reshape(-1)).max() < 1e-12 * np.abs(rho).max()
    # charge: int rho = sum_b f_b <psi_b|S|psi_b>
    _, spsi = m.apply_h(psi, want_s=True)
    q = np.sum(occ * np.real(np.sum(np.conj(psi) * spsi, axis=1)))
    assert abs(rho.sum() * omega / n ** 3 - q) < 1e-11 * q


# ---------------------------------------------------------------------------
# total energy at self-consistency
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("npa", [0, 4])
def test_total_energy_equals_direct_functional(npa):
    """On a small cell (24^3 grid, Ecut 6 Ha) converged to 1e-10 Ha/atom: the
    oracle's energy sum f eps - int V_Hxc[rho_in] rho_out + E_H + E_xc - TS
    equals the direct Kohn-Sham functional of the final block."""
    kw = dict(n_grid=24, ecut=6.0, n_bands=20, cheb_degree=8, etol=1e-10, max_iter=80,
              n_proj_per_atom=npa, q0=0.25 if npa else 0.0)
    p = O.default_params(**kw)
    pos = O.positions(1)
    r = O.scf_run_ex(p, pos)
    assert r["converged"]
    m = r["model"]
    psi, occ, rho = r["psi"], r["occ"], r["rho"]
    n = p.n_grid
    omega = A0 ** 3
    dv = omega / n ** 3
    _, _, kin = m.pw()
    ts_kin = np.sum(occ * np.sum(kin[None, :] * np.abs(psi) ** 2, axis=1))
    e_nl = 0.0
    if npa:
        w = np.conj(m.beta()) @ psi.T
        e_nl = np.sum(occ[None, :] * m.d()[:, None] * np.abs(w) ** 2)
    e_ion = np.sum(m.vion() * rho) * dv
    _, eh, exc = m.vhxc(rho)
    f = occ / 2
    ok = (f > 1e-300) & (f < 1 - 1e-16)
    ts = 2 * p.smear_kt * np.sum(f[ok] * np.log(f[ok]) + (1 - f[ok]) * np.log(1 - f[ok]))
    direct = ts_kin + e_nl + e_ion + eh + exc + ts
    assert abs(direct - r["energies"][-1]) < 1e-7, (direct, r["energies"][-1])
    # the electron count of the final density
    assert abs(rho.sum() * dv - 32.0) < 1e-8


# ---------------------------------------------------------------------------
# mixing
# ---------------------------------------------------------------------------
def kerker_np(f, n, L, q):
    g = np.fft.fftn(f.reshape(n, n, n))
    fr = np.fft.fftfreq(n, 1.0 / n)
    nyq = np.abs(fr) == n // 2
    b = 2 * np.pi / L
    g2 = b * b * (fr[:, None, None] ** 2 + fr[None, :, None] ** 2 + fr[None, None, :] ** 2)
    k = np.where(g2 > 0, g2 / (g2 + q * q), 0.0)
    k[nyq[:, None, None] | nyq[None, :, None] | nyq[None, None, :]] = 0.0
    return np.real(np.fft.ifftn(g * k)).reshape(-1)


def test_kerker_single_modes():
    n, L, q, alpha = 8, 5.0, 0.8, 0.5
    x = np.arange(n ** 3)
    ix, iy, iz = x % n, (x // n) % n, x // (n * n)
    for (h, k, l) in [(1, 0, 0), (1, 2, 0), (3, 1, 2)]:
        mode = np.cos(2 * np.pi * (h * ix + k * iy + l * iz) / n)
        mix = O.Pulay(n, L, depth=1, alpha=alpha, kerker_q=q)
        rin = np.full(n ** 3, 0.3)
        out = mix.mix(rin, rin + mode + 0.7)  # constant part of the residual is removed
        g2 = (2 * np.pi / L) ** 2 * (h * h + k * k + l * l)
        assert np.abs(out - (rin + alpha * g2 / (g2 + q * q) * mode)).max() < 1e-13


def test_pulay_matches_lagrange_closed_form():
    """Eight steps of a linear fixed-point map rho_out = M rho_in + c through
    the oracle's mixer (depth 6) against the closed form: coefficients
    c = G^-1 1 / (1^T G^-1 1), G_ij = <R_i|R_j>, over the last six pairs;
    rho_next = sum c rho_in + alpha K[sum c R]."""
    n, L, depth, alpha, q = 6, 4.0, 6, 0.5, 0.8
    n3 = n ** 3
    rng = np.random.default_rng(4)
    M = 0.4 * np.eye(n3) + 0.02 * rng.standard_normal((n3, n3))
    c0 = rng.standard_normal(n3)
    mix = O.Pulay(n, L, depth=depth, alpha=alpha, kerker_q=q)
    hin, hres = [], []
    rin = rng.standard_normal(n3)
    for step in range(8):
        rout = M @ rin + c0
        got = mix.mix(rin, rout)
        hin.append(rin.copy())
        hres.append(rout - rin)
        hin, hres = hin[-depth:], hres[-depth:]
        R = np.array(hres)
        G = R @ R.T
        gi = np.linalg.solve(G, np.ones(len(hres)))
        c = gi / gi.sum()
        ropt = c @ np.array(hin)
        rres = c @ R
        want = ropt + alpha * kerker_np(rres, n, L, q)
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max(), step
        # optimality: the combined residual is no longer than any single one
        assert np.linalg.norm(rres) <= min(np.linalg.norm(r) for r in hres) + 1e-12
        rin = got


# ---------------------------------------------------------------------------
# occupations with k weights
# ---------------------------------------------------------------------------
def test_weighted_occupations():
    rng = np.random.default_rng(8)
    eps = np.sort(rng.uniform(-0.5, 0.5, 40))
    occ1, mu1, ts1 = O.fermi_occupations(eps, 32.0, 0.01)
    occw, muw, tsw = O.fermi_occupations_w(eps, np.ones(40), 32.0, 0.01)
    assert np.array_equal(occ1, occw) and mu1 == muw and ts1 == tsw
    # four k-points of weight 1/4 with shifted spectra: the weighted count
    e4 = np.concatenate([eps + 0.01 * k for k in range(4)])
    w4 = np.full(160, 0.25)
    occ, mu, ts = O.fermi_occupations_w(e4, w4, 32.0, 0.01)
    assert abs(np.sum(w4 * occ) - 32.0) < 1e-10
    f = occ / 2
    ok = (f > 1e-300) & (f < 1 - 1e-16)
    assert abs(ts - 2 * 0.01 * np.sum(w4[ok] * (f[ok] * np.log(f[ok])
                                                 + (1 - f[ok]) * np.log(1 - f[ok])))) < 1e-14
    # the same spectrum at every k equals the single-k result
    occ_s, mu_s, ts_s = O.fermi_occupations_w(np.tile(eps, 4), w4, 32.0, 0.01)
    assert abs(mu_s - mu1) < 1e-12 and abs(ts_s - ts1) < 1e-12
