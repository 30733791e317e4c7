"""Twist-angle sweep: TBG cells m = 8..14 (n = 868 .. 2524), 3 x 3 k grid each,
fillings seeded uniform in [-2, 2], FP64, Fermi smearing 1 meV, optimal-damping
mixing.  Times SCF iterations of all seven systems advanced together (one
uneven-size eigensolver batch of 63 matrices per iteration, forge_sweep_run)
against the same iterations run system by system (forge_scf_run), and prints
one JSON line.

    python -m forge_amd.sweep_bench [--iters 3] [--warmup 1] [--mmin 8] [--mmax 14] [--precision 0|1] [--graph]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .hf import Engine, Setup, Sweep
from .seed import seeded_rho


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mmin", type=int, default=8)
    ap.add_argument("--mmax", type=int, default=14)
    ap.add_argument("--numk", type=int, default=3)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--precision", type=int, default=0,
                    help="1: mixed precision (FP32-copy hemv, FP32 back-transform) for every system")
    ap.add_argument("--graph", action="store_true",
                    help="also time the batched iterations with the reduction replayed from a HIP graph")
    a = ap.parse_args()
    rng = np.random.default_rng(a.seed)
    ms = list(range(a.mmin, a.mmax + 1))
    nus = [float(x) for x in rng.uniform(-2.0, 2.0, len(ms))]
    engs = [Engine(Setup(ntheta=m, numk=a.numk, nfilling=nu, xi_nm=40.0)) for m, nu in zip(ms, nus)]
    starts = [seeded_rho(e.n, e.nnc, nu) for e, nu in zip(engs, nus)]

    def reset(graph=0):
        for e, s in zip(engs, starts):
            e.set_options(occ_mode=2, mix_mode=0, kT=0.001, graph=graph, precision=a.precision,
                          switch_de=1e-12)
            e.scf_init(s)

    # system by system
    reset()
    for e in engs:
        e.scf_run(a.warmup)
    t0 = time.perf_counter()
    seq = [e.scf_run(a.iters) for e in engs]
    t_seq = (time.perf_counter() - t0) / a.iters
    print(f"[sweep_bench] sequential: {t_seq * 1e3:.1f} ms per iteration of all systems", flush=True)
    # batched
    reset()
    sw = Sweep(engs)
    sw.run(a.warmup)
    t0 = time.perf_counter()
    bat = sw.run(a.iters)
    t_bat = (time.perf_counter() - t0) / a.iters
    sw.close()
    print(f"[sweep_bench] batched:    {t_bat * 1e3:.1f} ms per iteration of all systems", flush=True)
    t_gr = None
    if a.graph:
        reset(graph=1)
        sw = Sweep(engs)
        sw.run(a.warmup)  # captures the reduction
        t0 = time.perf_counter()
        gr = sw.run(a.iters)
        t_gr = (time.perf_counter() - t0) / a.iters
        sw.close()
        assert all(np.array_equal(x, y) for x, y in zip(gr, bat))
        print(f"[sweep_bench] batched, reduction from a HIP graph: {t_gr * 1e3:.1f} ms per iteration",
              flush=True)
    dE = max(float(np.abs(x[:, 1] - y[:, 1]).max()) for x, y in zip(seq, bat))
    out = {
        "metric": "SCF iteration of a twist-angle sweep (all systems)",
        "value": round(t_bat * 1e3, 2), "unit": "ms", "higher_is_better": False,
        "sequential_ms": round(t_seq * 1e3, 2), "speedup_vs_sequential": round(t_seq / t_bat, 3),
        "systems": [{"m": m, "n": e.n, "nu": round(nu, 4), "k_points": e.Nk}
                    for m, e, nu in zip(ms, engs, nus)],
        "eigensolves_per_iteration": int(sum(e.Nk for e in engs)),
        "max_energy_difference_eV": dE,
        "dtype": "f64" if a.precision == 0 else "f64 (FP32-copy hemv, FP32 back-transform)",
    }
    if t_gr is not None:
        out["graph_ms"] = round(t_gr * 1e3, 2)
    for e in engs:
        e.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
