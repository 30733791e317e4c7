#!/usr/bin/env python3
"""Benchmark: wall-time to a converged SCF (|E_in - E_new| < 1e-6 eV per moire
cell) for 1.05 deg twisted bilayer graphene (cell index 31, 11908 pz orbitals,
3 x 3 k grid, filling +1, seeded start) on one MI355X.

A step is one full SCF solve from the seeded start to convergence (assembly,
Hermitian eigensolves of all k, density products, energies, mixing, every
iteration).  Inputs are synthetic (geometry from the cell index, seeded
Gaussian start density); nothing is read from disk.

Prints one JSON line (rank 0) with the roofline of the dominant kernel (the
Householder hemv, HBM-bound) measured with HIP events inside the timed region,
and a CPU baseline: the oracle (a restatement of the reference's OpenMP/MKL
CPU path) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (ntheta, nlayers, numk, nfilling)
    "tbg_1.05deg_n11908_k3": (31, 2, 3, 1.0),
    "tbg_1.61deg_n5044_k4": (20, 2, 4, 1.0),
    "tbg_6.01deg_n364_k6": (5, 2, 6, 1.0),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="tbg_1.05deg_n11908_k3", choices=sorted(WORKLOADS))
    ap.add_argument("--etol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=80)
    ap.add_argument("--beta", type=float, default=0.3)
    ap.add_argument("--kT", type=float, default=0.001)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-ntheta", type=int, default=14)
    return ap.parse_args()


def hemv_algorithmic_bytes(n: int, K: int) -> float:
    """Lower triangle of every trailing matrix read once per column:
    sum_j 8 (n - j - 1)^2 bytes per matrix (complex128, half of m^2)."""
    m = n - 1
    return 8.0 * (m * (m + 1) * (2 * m + 1) / 6.0) * K


def cpu_baseline(args, iters_gpu: int, n_full: int):
    """Oracle (reference CPU path restated: OpenMP over k, MKL zheevr range 'I',
    rank-1 zgemm bulk density, ODA-free linear mixing here) for ONE SCF
    iteration at a smaller cell of the same family, scaled by (n_full/n)^3 and
    by the GPU's iteration count."""
    import numpy as np

    from forge_amd.seed import seeded_rho
    from oracle.pyoracle import System

    nt = args.cpu_sample_ntheta
    ntheta, nlayers, numk, nfill = WORKLOADS[args.workload]
    nthreads = min(os.cpu_count() or 1, numk * numk)
    s = System(nt, numk, nfill, nlayers=nlayers, xi_nm=40.0)
    start = seeded_rho(s.n, s.nnc, nfill)
    s.longrange()  # setup cost excluded, as for the GPU
    t0 = time.perf_counter()
    its, traj, _, timing = s.scf(1, rho_start=start, occ_mode=2, kT=args.kT, mix_mode=1,
                                 beta=args.beta, nthreads=nthreads)
    dt = time.perf_counter() - t0
    scale = (n_full / s.n) ** 3
    per_iter = dt * scale
    return {
        "value": per_iter * max(iters_gpu, 1),
        "unit": "s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"one SCF iteration of the oracle at cell index {nt} (n={s.n}, {numk}x{numk} k, "
                   f"{nthreads} OpenMP threads, MKL sequential per k) took {dt:.2f} s; scaled by "
                   f"(n/{s.n})^3={scale:.1f} to n={n_full} and by the GPU run's {iters_gpu} "
                   f"iterations (extrapolated)"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the k-point "
              "sharding across GPUs is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from forge_amd.hf import Engine, Setup
    from forge_amd.lib import device_count
    from forge_amd.seed import seeded_rho

    if device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(3)
    ntheta, nlayers, numk, nfill = WORKLOADS[args.workload]
    setup = Setup(ntheta=ntheta, nlayers=nlayers, numk=numk, nfilling=nfill, xi_nm=40.0)
    eng = Engine(setup)
    n, K = eng.n, eng.Nk
    start = seeded_rho(n, eng.nnc, nfill)
    eng.set_options(occ_mode=2, mix_mode=1, kT=args.kT, beta=args.beta, etol=args.etol, timing=1)

    def one_solve():
        eng.scf_init(start)
        traj = eng.scf_run(args.max_iter)
        return traj

    for _ in range(args.warmup):
        one_solve()
    walls, iters, hemv_ms, hemv_launches, stage = [], [], 0.0, 0, np.zeros(16)
    last = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        traj = one_solve()
        walls.append(time.perf_counter() - t0)
        tm = eng.timings()
        stage += tm
        hemv_ms += tm[10]
        hemv_launches += int(tm[9])
        iters.append(len(traj))
        last = traj
        print(f"[bench] solve {len(walls)}: {len(traj)} iterations, {walls[-1]:.2f} s, "
              f"E_tot {traj[-1, 1]:.9f} eV, last dE {traj[-1, 8]:.2e}", file=sys.stderr, flush=True)
    wall = float(np.mean(walls))
    converged = bool(abs(last[-1, 8]) < args.etol)
    per_iter_ms = wall * 1e3 / max(iters[-1], 1)
    hemv_bytes_total = hemv_algorithmic_bytes(n, K) * sum(iters)
    avg_launch_ms = hemv_ms / max(hemv_launches, 1)
    bytes_per_launch = hemv_bytes_total / max(hemv_launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if hemv_launches else 0.0
    out = {
        "metric": "wall-time to SCF converged 1e-6 eV, 1.05deg TBG; mixed+learned vs plain FP64 GPU",
        "value": round(wall, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (geometry from the cell index; seeded Gaussian start density)",
        "config": {
            "workload": args.workload,
            "ndim": n, "k_grid": f"{numk}x{numk}", "nfilling": nfill, "layers": nlayers,
            "path": "plain FP64 GPU (Householder + divide-and-conquer), random (seeded) start",
            "mixing": f"linear {args.beta}", "smearing_eV": args.kT, "etol_eV": args.etol,
            "iterations": iters, "converged": converged, "ms_per_iteration": round(per_iter_ms, 1),
            "stage_ms_per_iteration": {
                "assembly": round(stage[0] / sum(iters), 1),
                "tridiagonal_reduction": round(stage[1] / sum(iters), 1),
                "tridiagonal_eigensolver": round(stage[2] / sum(iters), 1),
                "back_transformation": round(stage[3] / sum(iters), 1),
                "density_products": round(stage[4] / sum(iters), 1),
                "energies_mixing": round(stage[5] / sum(iters), 1),
            },
        },
        "roofline": {
            "kernel": "k_trd_hemv",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "launches": hemv_launches,
            "avg_launch_us": round(avg_launch_ms * 1e3, 2),
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
        },
    }
    eng.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, iters[-1], n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
