"""Training data, training and evaluation of the learned band-CG
preconditioner (chs_mlp_* in libchs_gpu.so).

Plain runs: for a seeded displacement sample of the synthetic cell, with
the potential fixed at V_ion (+ USPP projectors when asked), the converged
block psi* comes from plain Chebyshev-filtered subspace iteration with
Rayleigh-Ritz; the unconverged blocks are the early subspace-iteration
passes and a few band-CG (TPA) steps started from them. Each unconverged
block gives residual / correction pairs (chs_mlp_pairs); the network is
trained on the pairs of many samples and judged on held-out samples by
band-CG steps to convergence against plain TPA, with psi*'s eigenvalues as
the reference.
"""
import numpy as np

import chs


class Sample:
    """One seeded system bound on a context."""

    def __init__(self, ctx, seed, n_rep=1, n_grid=36, n_bands=24, uspp=False):
        self.ctx = ctx
        self.params = chs.scf_params(n_rep=n_rep, n_grid=n_grid, n_bands=n_bands,
                                     n_proj_per_atom=4 if uspp else 0,
                                     q0=0.25 if uspp else 0.0)
        self.pos = chs.synth_positions(n_rep, seed=seed)
        self.sys = chs.build_system(ctx, self.params, self.pos)
        self.H = chs.Hamiltonian(ctx, n_grid, self.sys["miller"], self.sys["kin"],
                                 self.sys["vion"])
        if uspp:
            self.H.bind_projectors(self.sys["beta"], self.sys["d0"], self.sys["q"])
        self.kin = self.sys["kin"].sync_to_host()
        self.n_b = n_bands
        self.seed = seed

    def bind(self):
        """Re-bind this sample's arrays (another sample may have been bound)."""
        self.H = chs.Hamiltonian(self.ctx, self.params.n_grid, self.sys["miller"],
                                 self.sys["kin"], self.sys["vion"])
        if self.params.n_proj_per_atom:
            self.H.bind_projectors(self.sys["beta"], self.sys["d0"], self.sys["q"])

    def chfsi(self, psi, passes, degree, eps=None):
        """`passes` filter + Rayleigh-Ritz passes of a device c128 block (in place)."""
        lmax, _ = chs.lanczos_bound(self.ctx, 10, 4)
        for _ in range(passes):
            b4 = chs.compute_spectral_bounds(eps, lmax)
            chs.chebyshev_filter(self.ctx, psi, degree, b4, out=psi)
            eps, _ = chs.rayleigh_ritz(self.ctx, psi)
        return eps

    def converged(self, degree=16, max_passes=60, tol=1e-12):
        """psi* and eps* by plain subspace iteration to eigenvalue changes < tol."""
        psi = chs.initial_psi(self.ctx, self.sys["kin"], self.n_b, 1000 + self.seed)
        eps = self.chfsi(psi, 2, degree)
        for _ in range(max_passes):
            new = self.chfsi(psi, 1, degree, eps)
            done = np.abs(new - eps).max() < tol
            eps = new
            if done:
                break
        return psi, eps

    def rough(self, passes=3, degree=8):
        """An unconverged block: `passes` low-degree passes from a fresh start."""
        psi = chs.initial_psi(self.ctx, self.sys["kin"], self.n_b, 2000 + self.seed)
        eps = self.chfsi(psi, passes, degree)
        return psi, eps


def copy_block(ctx, psi):
    out = chs.Gpu(ctx, psi.shape, psi.dtype)
    ctx.check(chs.lib().chs_d2d(ctx.h, out.ptr, psi.ptr, psi.nbytes))
    return out


def collect_pairs(ctx, samples, pairs, passes=(2, 3), cg_steps=(2, 5), gap_min=1e-3, degree=8):
    """Training pairs of every sample: its early subspace-iteration blocks
    (`passes` passes of degree `degree`) and band-CG (TPA) from them after
    each of cg_steps steps."""
    used = 0
    for s in samples:
        s.bind()
        ref, eps_ref = s.converged()
        for p in passes:
            blk, _ = s.rough(p, degree)
            used += chs.mlp_pairs(ctx, blk, ref, eps_ref, pairs, gap_min)
            for k in cg_steps:
                b2 = copy_block(ctx, blk)
                chs.band_cg_minimize(ctx, b2, k, 1e-12, precond=0)
                used += chs.mlp_pairs(ctx, b2, ref, eps_ref, pairs, gap_min)
    return used


def collect_pairs_seeded(ctx, seeds, make, pairs, **kw):
    """collect_pairs over samples built one at a time by make(seed) and
    dropped after use (large cells: one sample's projectors on the device at
    a time)."""
    used = 0
    for sd in seeds:
        s = make(sd)
        used += collect_pairs(ctx, [s], pairs, **kw)
        del s
    return used


def evaluate(ctx, sample, net, tol=1e-9, max_steps=200, uspp_term=False, passes=3, degree=8,
             full_overlap=False):
    """Band-CG from the same ChFSI block (`passes` passes of degree `degree`)
    with TPA and with the network (full_overlap: the lower-band
    S-orthogonalisation removes the complex overlap, see band_cg_minimize).
    Returns dict with total steps, converged bands, max |eps - eps*| over all
    bands and over the converged ones, for both."""
    return evaluate_settings(ctx, sample, net, [(tol, max_steps, full_overlap)], uspp_term,
                             passes, degree)[0]


def evaluate_settings(ctx, sample, net, settings, uspp_term=False, passes=3, degree=8):
    """evaluate() for several (tol, max_steps, full_overlap) settings from
    one converged reference and one start block of the sample."""
    sample.bind()
    ref, eps_ref = sample.converged()
    blk, _ = sample.rough(passes, degree)
    return [_cg_pair(ctx, net, blk, eps_ref, tol, max_steps, full_overlap, uspp_term)
            for tol, max_steps, full_overlap in settings]


def _cg_pair(ctx, net, blk, eps_ref, tol, max_steps, full_overlap, uspp_term):
    out = {}
    for name in ("tpa", "mlp"):
        b = copy_block(ctx, blk)
        if name == "tpa":
            eps, steps, conv = chs.band_cg_minimize(ctx, b, max_steps, tol,
                                                    precond=1 if uspp_term else 0,
                                                    full_overlap=full_overlap)
        else:
            table = net.scale_table(b)
            eps, steps, conv = chs.band_cg_minimize(ctx, b, max_steps, tol, precond=2,
                                                    mlp_scale=table, full_overlap=full_overlap)
        err = np.abs(eps - eps_ref)
        done = np.asarray(conv) != 0
        out[name] = {"steps": int(steps.sum()), "max_steps": int(steps.max()),
                     "converged": int(done.sum()), "eps_err": float(err.max()),
                     "eps_err_converged": float(err[done].max()) if done.any() else None}
    return out
