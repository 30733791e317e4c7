"""Reduces rocprofv3 output of `python -m forge_amd.profile_step` runs to the
tables kept under profiles/.

    python -m forge_amd.counter_summary kernels <run_results.db | run_kernel_stats.csv> [top]
        per-kernel calls, summed and mean duration of a kernel trace
    python -m forge_amd.counter_summary traffic <fetch csv> <write csv> <n> <K> <elem bytes>
        HBM bytes per kernel from a FETCH_SIZE run and a WRITE_SIZE run
        (run_counter_collection.csv, KiB per dispatch), calibrated on
        k_bt_promote, a streaming copy of known size (it reads the n x m
        real eigenvectors of K matrices, 8 B each, and writes them as
        complex, elem bytes each; m = nUpper)
    python -m forge_amd.counter_summary busy <csv>
        matrix-core busy fraction SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE
        / 8 x 1024 SIMDs) (GRBM_GUI_ACTIVE summed over the 8 XCDs) and LDS
        bank conflicts per LDS instruction, per kernel

Counter runs serialise the kernels, so their durations are those of
kernels running alone.
"""
from __future__ import annotations

import collections
import csv
import sqlite3
import sys


def _name(raw: str) -> str:
    return (raw.split("(")[0].replace("void ", "").replace("forge::", "")
            .replace("HIP_vector_type<float, 2u>", "float2"))


def kernels(path: str, top: int = 30) -> list[str]:
    agg = collections.defaultdict(lambda: [0, 0.0])
    if path.endswith(".db"):
        con = sqlite3.connect(path)
        for name, s, e in con.execute("select name, start, end from kernels"):
            a = agg[_name(name)]
            a[0] += 1
            a[1] += e - s
    else:
        for r in csv.DictReader(open(path)):
            a = agg[_name(r["Name"])]
            a[0] += int(r["Calls"])
            a[1] += float(r["TotalDurationNs"])
    tot = sum(v[1] for v in agg.values())
    out = [f"total kernel time {tot / 1e6:.1f} ms, {sum(v[0] for v in agg.values())} dispatches"]
    for nm, (cnt, t) in sorted(agg.items(), key=lambda x: -x[1][1])[:top]:
        out.append(f"{t / 1e6:10.1f} ms {cnt:7d} {t / cnt / 1e3:10.1f} us {100 * t / tot:6.1f}%  {nm[:80]}")
    return out


def _counter(path: str):
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(dict)
    for r in csv.DictReader(open(path)):
        k = _name(r["Kernel_Name"])
        agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k][r["Dispatch_Id"]] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
    return agg, disp


def traffic(fetch: str, write: str, n: int, K: int, eb: int) -> list[str]:
    f, fd = _counter(fetch)
    w, _ = _counter(write)
    m = (n + 2 * (min(20, n) // 2)) // 2
    prom = next(k for k in f if k.startswith("k_bt_promote"))
    calls = len(fd[prom])
    sf = n * m * 8 * K * calls / (f[prom]["FETCH_SIZE"] * 1024)
    sw = n * m * eb * K * calls / (w[prom]["WRITE_SIZE"] * 1024)
    out = [f"calibration on {prom} ({calls} calls): reads = FETCH_SIZE x {sf:.4f}, "
           f"writes = WRITE_SIZE x {sw:.4f}",
           f"{'kernel':58s} {'calls':>5s} {'read GB':>9s} {'write GB':>9s}"]
    for k in sorted(f, key=lambda k: -f[k]["FETCH_SIZE"]):
        rd = f[k]["FETCH_SIZE"] * 1024 * sf / 1e9
        wr = w[k]["WRITE_SIZE"] * 1024 * sw / 1e9 if k in w else 0.0
        out.append(f"{k[:58]:58s} {len(fd[k]):5d} {rd:9.1f} {wr:9.1f}")
    return out


def busy(path: str) -> list[str]:
    agg, disp = _counter(path)
    out = [f"{'kernel':58s} {'calls':>5s} {'ms':>8s} {'busy':>6s} {'LDS conflicts/instr':>20s}"]
    for k in sorted(agg, key=lambda k: -sum(disp[k].values())):
        c = agg[k]
        b = c["SQ_VALU_MFMA_BUSY_CYCLES"] * 8 / (c["GRBM_GUI_ACTIVE"] * 1024) if c["GRBM_GUI_ACTIVE"] else 0
        bc = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_INSTS_LDS"] if c.get("SQ_INSTS_LDS") else 0.0
        out.append(f"{k[:58]:58s} {len(disp[k]):5d} {sum(disp[k].values()) / 1e6:8.1f} {b:6.3f} {bc:20.2f}")
    return out


def main(argv):
    cmd = argv[1]
    if cmd == "kernels":
        lines = kernels(argv[2], int(argv[3]) if len(argv) > 3 else 30)
    elif cmd == "traffic":
        lines = traffic(argv[2], argv[3], int(argv[4]), int(argv[5]), int(argv[6]))
    elif cmd == "busy":
        lines = busy(argv[2])
    else:
        raise SystemExit(__doc__)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
