"""A fixed amount of SCF work for profilers: context creation, the seeded
start and N SCF iterations of one system, with no timing of its own.

    python -m forge_amd.profile_step [--ntheta 31] [--nlayers 2] [--numk 3]
        [--nfilling 1] [--iters 1] [--precision 0|1] [--stages] [--graph] [--graph-mode N] [--c3]
        [--reduction 0|1|2]

Run under rocprofv3 (kernel trace, counters) to attribute device time and
HBM traffic to kernels for exactly the work of N iterations.
"""
from __future__ import annotations

import argparse
import sys
import time

from .hf import Engine, Setup
from .seed import seeded_rho


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ntheta", type=int, default=31)
    ap.add_argument("--nlayers", type=int, default=2)
    ap.add_argument("--numk", type=int, default=3)
    ap.add_argument("--nfilling", type=float, default=1.0)
    ap.add_argument("--iters", type=int, default=1)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--stages", action="store_true",
                    help="time the stages of each iteration (device events)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the Householder reduction from a HIP graph (graph mode 1)")
    ap.add_argument("--graph-mode", type=int, default=None,
                    help="forge_scf_options.graph: 1 reduction graph, 2 device iteration, "
                         "3 iteration graph")
    ap.add_argument("--reduction", type=int, default=0,
                    help="forge_scf_options.reduction: 0 by order, 1 one-stage, 2 two-stage")
    ap.add_argument("--c3", action="store_true",
                    help="C3-reduced k-set (nenforceC3 = 1): one eigensolve per C3 orbit")
    a = ap.parse_args()
    t0 = time.perf_counter()
    e = Engine(Setup(ntheta=a.ntheta, nlayers=a.nlayers, numk=a.numk, nfilling=a.nfilling,
                     xi_nm=40.0))
    gm = a.graph_mode if a.graph_mode is not None else (1 if a.graph else 0)
    e.set_options(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3, precision=a.precision,
                  switch_de=1e-12, timing=1 if a.stages else 0, graph=gm, reduction=a.reduction)
    if a.c3:
        e.set_c3(True)
    e.scf_init(seeded_rho(e.n, e.nnc, a.nfilling))
    print(f"[profile_step] n={e.n} Nk={e.Nk} c3={int(a.c3)} precision={a.precision} "
          f"set-up {time.perf_counter() - t0:.1f}s", flush=True)
    for it in range(a.iters):
        t1 = time.perf_counter()
        r = e.scf_run(1)
        print(f"[profile_step] iteration {it + 1}: E_tot {r[0, 1]:.9f} dE {r[0, 8]:+.3e} "
              f"({time.perf_counter() - t1:.2f}s)", flush=True)
        if a.stages:
            t = e.timings()
            print("[profile_step]   stage ms: " + " ".join(
                f"{name} {t[q]:.1f}" for q, name in enumerate(
                    ("assembly", "reduction", "tridiag_eig", "back_transform", "density",
                     "energies_mixing"))) + f" (back-transform block width {t[13]:.0f}, "
                f"hemv variants {t[14]:04.0f} by band)",
                  flush=True)
    e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
