#!/usr/bin/env python3
"""SCF time-to-1e-6 Ha/atom on one MI355X (libchs_gpu.so).

Workload (default): the 64-atom diamond-cubic cell (2x2x2 supercell of
a = 10.26 bohr, atoms displaced by seeded N(0, 0.05 bohr)), V_loc(r) =
-4 erf(r)/r, Gamma point, 0.5|G|^2 <= 15 Ha (23,847 plane waves), 72^3
grid, 160 bands, Chebyshev degree 12, complex128 throughout with FP64
Cholesky-QR + Rayleigh-Ritz, Pulay(6)/Kerker mixing, Fermi-Dirac kT = 0.01
Ha; converged when |dE|/N_atoms < 1e-6 Ha on two consecutive iterations.

One "step" is one full SCF run from the seeded start to convergence with
the start wavefunctions already resident in HBM; `value` is the mean
SCF-loop time of the K timed runs (lower is better).

The JSON line also carries
  roofline     - the dominant kernel of the Chebyshev H|psi> step, timed
                 with HIP events on the context stream at the workload's
                 shape, against MI355X HBM peak (8 TB/s), with the
                 algorithmic bytes of the stick/plane design;
  cpu_baseline - the FP64 CPU oracle (oracle/) on this host: one SCF
                 iteration of the same workload timed, extrapolated to the
                 GPU run's iteration count.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chemrust-scf_amd"))

METRIC = "SCF time-to-1e-6 Ha/atom (64/216 atoms); FP32+learned vs plain FP64 on GPU"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "A": dict(n_rep=1, n_grid=36, n_bands=20, cheb_degree=10),
    "B": dict(n_rep=2, n_grid=72, n_bands=160, cheb_degree=12),
}
WORKLOAD = {
    "A": "8-atom diamond-cubic cell, Gamma, Ecut 15 Ha (2,945 PW), 36^3, 20 bands, "
         "Chebyshev degree 10, complex128",
    "B": "64-atom diamond-cubic 2x2x2 cell, Gamma, Ecut 15 Ha (23,847 PW), 72^3, 160 bands, "
         "Chebyshev degree 12, complex128 filter + FP64 Cholesky-QR/Rayleigh-Ritz",
}


def roofline(chs, ctx, cfg):
    """Time the three H|psi> kernels at the workload's shape; the dominant one
    (largest time) is reported against HBM peak."""
    import numpy as np
    L = cfg["n_rep"] * 10.26
    n = cfg["n_grid"]
    mil, idx, kin = chs.gsphere(ctx, L, 15.0, n)
    rng = np.random.default_rng(0)
    veff = chs.Gpu.from_host(ctx, rng.standard_normal(n ** 3) * 0.1)
    chs.Hamiltonian(ctx, n, mil, kin, veff)
    prof = chs.profile_hpsi(ctx, cfg["n_bands"], chs.C128, reps=20)
    names = ["k_stick_inv", "k_plane", "k_stick_fwd"]
    k = int(np.argmax(prof["ms"]))
    achieved = prof["bytes"][k] / (prof["ms"][k] * 1e-3) / 1e9
    nb, s = cfg["n_bands"], 16
    canonical = nb * (4 * s * n ** 3 + 3 * s * prof["n_pw"]) + 8 * n ** 3
    step_ms = float(prof["ms"].sum())
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{names[k]}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": names[k], "kernel_ms": round(float(prof["ms"][k]), 4),
        "algorithmic_bytes": float(prof["bytes"][k]),
        "step_kernels_ms": [round(float(x), 4) for x in prof["ms"]],
        "step_canonical_bytes": float(canonical),
        "step_canonical_gbs": round(canonical / (step_ms * 1e-3) / 1e9, 1),
        "sticks": prof["sticks"], "active_x": prof["active_x"],
    }


def cpu_baseline(cfg, n_iter_gpu):
    """FP64 oracle on this host: one SCF iteration (Lanczos k=10, one filter +
    Rayleigh-Ritz pass, density, V_Hxc), extrapolated to the GPU's iteration
    count (+3 extra first-iteration passes)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O
    p = O.default_params(**cfg)
    pos = O.positions(cfg["n_rep"])
    m = O.Model(p, pos)
    n3 = m.n3
    omega = (cfg["n_rep"] * 10.26) ** 3
    rho = np.full(n3, 4.0 * 8 * cfg["n_rep"] ** 3 / omega)
    v, _, _ = m.vhxc(rho)
    m.set_veff(m.vion() + v)
    psi = m.initial_psi(cfg["n_bands"])
    t0 = time.perf_counter()
    lmax, _ = m.lanczos_bound(10, 4)
    t1 = time.perf_counter()
    _, y, _ = m.chebyshev_filter(psi, cfg["cheb_degree"], O.spectral_bounds(None, lmax))
    y, eps = m.rayleigh_ritz(y)
    t2 = time.perf_counter()
    occ, _, _ = O.fermi_occupations(eps, 4.0 * 8 * cfg["n_rep"] ** 3, 0.01)
    rho_out = m.density(y, occ)
    m.vhxc(rho_out)
    t3 = time.perf_counter()
    t_pass = t2 - t1
    t_iter = (t1 - t0) + t_pass + (t3 - t2)
    total = n_iter_gpu * t_iter + 3 * t_pass
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": round(total, 3), "unit": "s", "cores": cores, "kind": "port",
        "sample": f"one SCF iteration of the same workload with the FP64 oracle "
                  f"({t_iter:.2f} s: Lanczos {t1 - t0:.2f} s, filter+RR {t_pass:.2f} s, "
                  f"density+V_Hxc {t3 - t2:.2f} s) x {n_iter_gpu} GPU iterations "
                  f"+ 3 first-iteration passes (extrapolated)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="B", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--precision", default="c128", choices=["c128", "c64"],
                    help="filter precision: c128 throughout, or complex64 filter with FP64 "
                         "Rayleigh-Ritz and the last SCF iterations in c128")
    ap.add_argument("--hpsi-only", action="store_true",
                    help="only time the H|psi> kernels (for counter runs) and print the roofline")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="with --hpsi-only: also run 4 device-to-device copies of this many "
                         "bytes (the counter scale reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by k-points/bands, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import chs
    ctx = chs.Context(0)
    cfg = CONFIGS[args.config]
    if args.hpsi_only:
        if args.copy_bytes:
            import numpy as np
            a = chs.Gpu(ctx, (args.copy_bytes,), np.uint8).zero()
            b = chs.Gpu(ctx, (args.copy_bytes,), np.uint8)
            for _ in range(4):
                ctx.check(chs.lib().chs_d2d(ctx.h, b.ptr, a.ptr, args.copy_bytes))
            ctx.synchronize()
        print(json.dumps(roofline(chs, ctx, cfg)))
        return
    pos = chs.synth_positions(cfg["n_rep"])
    prec = chs.C128 if args.precision == "c128" else chs.C64
    params = chs.scf_params(**cfg, max_iter=100, precision=prec)
    for _ in range(args.warmup):
        r = chs.run_scf(ctx, params, pos)
        print(f"warmup: {r['n_iter']} iterations, {r['seconds']:.3f} s, "
              f"converged={r['converged']}", file=sys.stderr, flush=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    runs = []
    for k in range(args.steps):
        r = chs.run_scf(ctx, params, pos)
        runs.append(r)
        print(f"step {k}: {r['n_iter']} iterations, {r['seconds']:.3f} s, E = {r['energy']:.10f}",
              file=sys.stderr, flush=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    if not all(r["converged"] for r in runs):
        print("bench.py: SCF did not converge", file=sys.stderr)
        sys.exit(1)
    value = sum(r["seconds"] for r in runs) / len(runs)
    rl = roofline(chs, ctx, cfg)
    last = runs[-1]
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) / args.steps * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": WORKLOAD[args.config], "atoms": 8 * cfg["n_rep"] ** 3,
                   "n_pw": last["n_pw"], "grid": cfg["n_grid"], "bands": cfg["n_bands"],
                   "scf_iterations": last["n_iter"], "energy_Ha": last["energy"],
                   "phase_s": {k: round(last[k], 4) for k in
                               ("t_filter", "t_rr", "t_density", "t_veff", "t_lanczos")}},
        "roofline": rl,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, last["n_iter"])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
