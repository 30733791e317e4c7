"""Reduces a rocprofv3 kernel trace of the default bench command to the
summary kept under profiles/: the top kernels by total time, and for the
chase's back-transformation (k_ts_q2, FP64 and FP32 eigenvectors) the
trace's durations and matrix-core rate next to the bench's own HIP-event
figure for the same launches.  Deletes the raw trace directory afterwards
(it is too large to keep).

    python -m forge_amd.trace_summary <trace dir> <bench json line file> <summary out> [n] [K]

Every two-stage back-transformation is G = ceil((n - 1) / 32) consecutive
k_ts_q2 launches, group g = G - 1 first; the bench samples the launches with
g % 4 == 0.  Algorithmic flops of a launch: 16 per reflector element per
eigenvector column, sweeps 32 g .. 32 g + 31 holding n - 1 - s elements each,
m = nUpper eigenvectors of each of the K matrices.
"""
from __future__ import annotations

import csv
import json
import os
import shutil
import sys


def main():
    d, bj, out = sys.argv[1], sys.argv[2], sys.argv[3]
    n = int(sys.argv[4]) if len(sys.argv) > 4 else 11908
    K = int(sys.argv[5]) if len(sys.argv) > 5 else 9
    m = (n + 2 * (min(20, n) // 2)) // 2  # nUpper (Setup.f90:68)
    G = (n - 1 + 31) // 32
    b = json.load(open(bj))
    stats = list(csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv"))))
    tot = sum(float(r["TotalDurationNs"]) for r in stats)
    L = [f"{'kernel':48s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'pct':>6s}"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:20]:
        name = r["Name"].split("(")[0].replace("void ", "").replace("HIP_vector_type<float, 2u>", "float2")
        L.append(f"{name[:48]:48s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.1f} "
                 f"{float(r['AverageNs']) / 1e3:9.2f} {100 * float(r['TotalDurationNs']) / tot:6.2f}")
    L.append(f"all kernels: {tot / 1e6:.1f} ms")
    q2 = {"z2": [], "f32": []}
    with open(os.path.join(d, "run_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            nm = r["Kernel_Name"]
            if "k_ts_q2<double" in nm:
                key = "z2"
            elif "k_ts_q2<float" in nm:
                key = "f32"
            else:
                continue
            q2[key].append((int(r["Dispatch_Id"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))

    def flops(g):
        elems = sum(n - 1 - s for s in range(32 * g, min(32 * g + 32, n - 1)))
        return 16.0 * elems * m * K

    L.append("")
    L.append(f"k_ts_q2: trace vs the bench's HIP-event sample (n={n}, m={m} eigenvectors, K={K}; "
             f"{G} launches per back-transformation, sampled groups g % 4 == 0)")
    rfs = {"f32": b.get("roofline"), "z2": b["config"].get("roofline_fp64_iterations")}
    for key, lab, peak in (("z2", "FP64", 78.6), ("f32", "FP32", 157.3)):
        v = sorted(q2[key])
        if not v:
            continue
        runs = len(v) // G
        fl_all = fl_s = t_all = t_s = 0.0
        ns = 0
        for q in range(runs):
            for j, (_, dur) in enumerate(v[q * G:(q + 1) * G]):
                g = G - 1 - j
                fl_all += flops(g)
                t_all += dur
                if g % 4 == 0:
                    fl_s += flops(g)
                    t_s += dur
                    ns += 1
        L.append(f"{lab} launches in the run: {len(v)} ({runs} back-transformations of {G})")
        L.append(f"{lab} trace, all launches: avg {t_all / (runs * G) / 1e3:.2f} us, "
                 f"{fl_all / (t_all * 1e-9) / 1e12:.2f} TFLOP/s")
        L.append(f"{lab} trace, the {ns} sampled launches: avg {t_s / ns / 1e3:.2f} us, "
                 f"{fl_s / (t_s * 1e-9) / 1e12:.2f} TFLOP/s = {fl_s / (t_s * 1e-9) / 1e12 / peak:.4f} "
                 f"of {peak} TFLOP/s")
        rf = rfs[key]
        if rf:
            L.append(f"{lab} bench HIP events (timed solve only): {rf['timed_launches']} launches, avg "
                     f"{rf['avg_launch_us']:.2f} us, {rf['achieved']:.2f} TFLOP/s, frac {rf['frac']:.4f}")
    c = b["config"]
    L.append(f"bench line (under the profiler): value {b['value']} s, {b['steps']} iterations "
             f"{c['iterations_by_phase']}, converged {c['converged']}")
    open(out, "w").write("\n".join(L) + "\n")
    print("\n".join(L))
    shutil.rmtree(d)
    return 0


if __name__ == "__main__":
    sys.exit(main())
