#!/usr/bin/env python3
"""Writes the SCF golden fixtures of the 64-atom and 216-atom cells from the
FP64 oracle (oracle/liboracle.so), for the GPU parity tests in
tests/test_gpu_golden.py.

  scf_64atom_full.npz      64-atom cell (2x2x2 diamond cubic, local V_loc,
                           72^3, 160 bands, Chebyshev degree 12), full SCF to
                           |dE|/atom < 1e-6 Ha twice in a row;
  scf_216atom_first.npz    216-atom USPP cell (3x3x3, 4 projectors/atom,
                           Q = 0.25 Gaussian augmentation, 108^3, 512 bands,
                           degree 12), the first SCF iteration only (4 filter +
                           Rayleigh-Ritz passes from the seeded start).
  scf_8atom_kpts.npz       8-atom USPP cell, 2x2x2 Gamma-centred k-point set
                           {0, 1/2}^3 (8 k-points, uneven plane-wave counts),
                           full k-point SCF.
  scf_64atom_kpts_first.npz  64-atom USPP cell (72^3, 160 bands, degree 12)
                           on the same {0, 1/2}^3 set (n_pw 23,847-23,984),
                           the first SCF iteration with one filter +
                           Rayleigh-Ritz pass per k-point.

Each file holds the energy of every iteration, the eigenvalues of the last
one, the per-iteration density change, and the last output density on a
coarse sub-grid (every `stride`-th point per axis) plus its integral, so the
fixtures stay small. Run from the repository root after `make -C oracle`:

  OMP_NUM_THREADS=8 python tests/golden/make_scf_fixtures.py [b] [c] [k] [d]

The 64-atom run takes about ten minutes on 8 cores, the 216-atom first
iteration about forty.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

CASES = {
    "b": ("scf_64atom_full.npz", dict(n_rep=2, n_grid=72, n_bands=160, cheb_degree=12,
                                      max_iter=60), 3),
    "c": ("scf_216atom_first.npz", dict(n_rep=3, n_grid=108, n_bands=512, cheb_degree=12,
                                        n_proj_per_atom=4, q0=0.25, max_iter=1), 4),
}
# Gamma-centred 2x2x2 set {0, 1/2}^3, equal weights
KPTS = np.array([[i, j, k] for i in (0.0, 0.5) for j in (0.0, 0.5) for k in (0.0, 0.5)])
KCASES = {
    "k": ("scf_8atom_kpts.npz", dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10,
                                     n_proj_per_atom=4, q0=0.25, max_iter=60), 1),
    "d": ("scf_64atom_kpts_first.npz", dict(n_rep=2, n_grid=72, n_bands=160, cheb_degree=12,
                                            n_proj_per_atom=4, q0=0.25, max_iter=1,
                                            first_passes=1), 3),
}


def sub(rho, n, stride):
    return np.ascontiguousarray(rho.reshape(n, n, n)[::stride, ::stride, ::stride])


def main():
    which = sys.argv[1:] or ["b", "c", "k", "d"]
    for w in which:
        t0 = time.time()
        if w in KCASES:
            name, kw, stride = KCASES[w]
            p = O.default_params(**kw)
            pos = O.positions(kw["n_rep"])
            r = O.scf_run_k(p, pos, KPTS, np.full(len(KPTS), 1.0 / len(KPTS)))
            extra = dict(kpts=KPTS, n_pw=r["n_pw"], mu=r["mu"])
        else:
            name, kw, stride = CASES[w]
            p = O.default_params(**kw)
            pos = O.positions(kw["n_rep"])
            r = O.scf_run(p, pos, with_rho=True)
            extra = {}
        n = kw["n_grid"]
        L = kw["n_rep"] * 10.26
        dv = L ** 3 / n ** 3
        np.savez_compressed(
            os.path.join(HERE, name), params=np.array(repr(sorted(kw.items()))),
            n_iter=r["n_iter"], converged=r["converged"], energies=r["energies"],
            drho=r["drho"], eps=r["eps"], rho_sub=sub(r["rho"], n, stride), stride=stride,
            rho_integral=r["rho"].sum() * dv, **extra)
        print(f"{name}: {r['n_iter']} iterations (converged={r['converged']}), "
              f"E = {r['energies'][-1]:.10f}, {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
