"""Reduces a rocprofv3 kernel trace of the default bench command to the
summary kept under profiles/: the top kernels by total time, and for the
Householder hemv (FP64 matrix and FP32 copy) the trace's bandwidth over the
timed pass next to the bench's own HIP-event figure for the same launches.
Deletes the raw trace directory afterwards (it is too large to keep).

    python -m forge_amd.trace_summary <trace dir> <bench json line file> <summary out> [n] [K]

The bench's timed plain-FP64 iteration is the second FP64 reduction of the
run and its timed mixed-phase iteration the second FP32 one (each follows
one warm-up iteration); every reduction is n - 1 hemv launches, and the
bench samples the launches with column index j % 17 == 0.
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
    b = json.load(open(bj))
    stats = list(csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv"))))
    tot = sum(float(r["TotalDurationNs"]) for r in stats)
    L = [f"{'kernel':48s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'pct':>6s}"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:20]:
        name = r["Name"].split("(")[0].replace("void ", "").replace("HIP_vector_type<float, 2u>", "float2")
        if name.startswith("Cijk_"):
            name = "rocblas:" + name[:36]
        L.append(f"{name[:48]:48s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.1f} "
                 f"{float(r['AverageNs']) / 1e3:9.2f} {100 * float(r['TotalDurationNs']) / tot:6.2f}")
    L.append(f"all kernels: {tot / 1e6:.1f} ms")
    hem = {"z2": [], "f32": []}
    with open(os.path.join(d, "run_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            nm = r["Kernel_Name"]
            if "k_trd_hemv<forge::z2" in nm:
                key = "z2"
            elif "k_trd_hemv<HIP_vector_type<float" in nm:
                key = "f32"
            else:
                continue
            hem[key].append((int(r["Dispatch_Id"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    L.append("")
    L.append("k_trd_hemv: trace vs the bench's HIP-event sample (same run; algorithmic bytes 8 m^2 (FP64) "
             "or 4 m^2 (FP32 copy) per matrix x K, m = n-1-j)")
    for key, bpe, rf, lab in (("z2", 8.0, b["roofline"], "FP64     "),
                              ("f32", 4.0, b["config"]["roofline_mixed_phase"], "FP32 copy")):
        v = sorted(hem[key])
        npass = n - 1
        L.append(f"{lab} launches in the run: {len(v)} ({len(v) // npass} passes of {npass})")
        p = v[npass:2 * npass]
        byts = [bpe * (n - 1 - j) ** 2 * K for j in range(npass)]
        dur = [x[1] for x in p]
        alls = sum(byts) / (sum(dur) * 1e-9) / 1e9
        sj = [j for j in range(npass) if j % 17 == 0]
        samp = sum(byts[j] for j in sj) / (sum(dur[j] for j in sj) * 1e-9) / 1e9
        L.append(f"{lab} trace, timed pass, all {npass} launches: avg {sum(dur) / len(dur) / 1e3:.2f} us, "
                 f"{alls:.1f} GB/s")
        L.append(f"{lab} trace, timed pass, the {len(sj)} sampled launches (j % 17 == 0): avg "
                 f"{sum(dur[j] for j in sj) / len(sj) / 1e3:.2f} us, {samp:.1f} GB/s")
        L.append(f"{lab} bench HIP events, same launches: avg {rf['avg_launch_us']:.2f} us, "
                 f"{rf['achieved']:.1f} GB/s, frac {rf['frac']:.4f} of 8000 GB/s")
    c = b["config"]
    L.append(f"bench line (under the profiler): value {b['value']} s projected, ms_per_iteration "
             f"{c['ms_per_iteration']}, speedup_vs_plain {c['speedup_vs_plain']}")
    open(out, "w").write("\n".join(L) + "\n")
    print("\n".join(L))
    shutil.rmtree(d)
    return 0


if __name__ == "__main__":
    sys.exit(main())
