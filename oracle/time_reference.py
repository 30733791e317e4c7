"""Timing check of the CPU restatement against the reference program built by
ref_build.sh (case t5k6: the 6.01 deg cell, 364 orbitals, 6 x 6 k grid,
filling +1, gate distance 40 nm, 2 SCF iterations from the reference's own
first-diagonalisation start, optimal-damping steps), with equal OpenMP thread
counts.  Test infrastructure: it runs oracle/_ref/t5k6/forge.x in a scratch
directory and the restatement through oracle/pyoracle.py, and prints both
wall times and both energy sequences.

    python3 oracle/time_reference.py [threads ...]      (default 8 4 1)
"""
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle.pyoracle import System  # noqa: E402

REF = os.path.join(HERE, "_ref", "t5k6", "forge.x")


def main():
    threads = [int(a) for a in sys.argv[1:]] or [8, 4, 1]
    if not os.path.exists(REF):
        raise SystemExit(f"{REF} is not built (oracle/ref_build.sh t5k6 5 6 1.0 2 0 40.0 2)")
    for t in threads:
        rd = tempfile.mkdtemp(prefix="forge_ref_")
        try:
            os.makedirs(os.path.join(rd, "output"))
            os.makedirs(os.path.join(rd, "dataFock"))
            env = dict(os.environ, OMP_NUM_THREADS=str(t), OMP_STACKSIZE="512m", MKL_NUM_THREADS="1")
            t0 = time.perf_counter()
            subprocess.run(["bash", "-c", f"ulimit -s unlimited; {REF} > forge.log 2>&1"], cwd=rd,
                           check=True, env=env)
            t_ref = time.perf_counter() - t0
            eng = [f for f in os.listdir(os.path.join(rd, "output")) if f.startswith("Energy")][0]
            e_ref = [float(l.split()[1]) for l in open(os.path.join(rd, "output", eng)) if l.strip()]
        finally:
            shutil.rmtree(rd, ignore_errors=True)
        t0 = time.perf_counter()
        s = System(5, 6, 1.0, nlayers=2, xi_nm=40.0)
        _, traj, _, _ = s.scf(2, rho_start=None, occ_mode=0, mix_mode=0, nthreads=t)
        t_port = time.perf_counter() - t0
        print(f"threads {t}: reference {t_ref:.2f} s, restatement {t_port:.2f} s "
              f"(restatement / reference {t_port / t_ref:.2f}); E_tot reference {e_ref}, "
              f"restatement {[round(float(x), 9) for x in traj[:, 1]]}", flush=True)


if __name__ == "__main__":
    main()
