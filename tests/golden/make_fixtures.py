#!/usr/bin/env python3
"""Generate the committed golden fixtures from the compiled reference.

Runs only in a container where /root/reference and amdflang + MKL exist:
  1. oracle/ref_build.sh compiles the reference program and the project's
     harness (oracle/ref_harness.f90) for each case into oracle/_ref/<case>/;
  2. the harness dumps every stage of the hot path (geometry, LongRange,
     H_TB(k), first diagonalisation, energies, potential, H_HF(k), eigenvalues,
     one ODA iteration) as raw arrays;
  3. the unmodified reference program runs the SCF loop and writes its
     Energy/Convergence files; seeded cases start from a checkpoint written by
     forge_amd.checkpoint (the reference's nRead=1 path, Main.f90:187-235).
The outputs become tests/golden/<case>.npz (arrays) and <case>.traj.txt (the
reference's Energy and Convergence rows).  Nothing here is imported by tests.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from forge_amd.checkpoint import checkpoint_name, write_checkpoint  # noqa: E402
from forge_amd.seed import seeded_rho, seeded_rho_2spins  # noqa: E402

REF_BUILD = os.path.join(REPO, "oracle", "ref_build.sh")
REF_OUT = os.path.join(REPO, "oracle", "_ref")

# name, ntheta, numk, nfilling, itmax, seeded, xi_nm, nlayers, large arrays kept
# (True: all, False: none, or the set of names to keep)
CASES = [
    ("t1k3", 1, 3, 1.0, 12, False, 40.0, 2, True),
    ("t1k3s", 1, 3, 1.0, 12, True, 40.0, 2, False),
    ("t2k3", 2, 3, 1.0, 12, False, 40.0, 2, True),
    ("t2k3s", 2, 3, 1.0, 12, True, 40.0, 2, False),
    ("t2k3xi10", 2, 3, -1.0, 6, False, 10.0, 2, True),
    ("t3k6", 3, 6, 1.0, 6, False, 40.0, 2, {"rho0", "rhoNew1", "zH_HF_k2"}),
    ("tri1k3", 1, 3, 2.0, 6, False, 40.0, 3, True),
    ("t5k6s", 5, 6, 1.0, 8, True, 40.0, 2, False),
]
LARGE = {"zFockBulk0", "zFockWin0", "rho0", "zFockBulk1", "zFockWin1", "rhoNew1", "rho1",
         "zH_TB_k2", "zH_TB_kN", "zH_HF_k2", "zH_HF_kN"}
DT = {"f8": "<f8", "c16": "<c16", "i8": "<i8"}


def env():
    e = dict(os.environ)
    e["OMP_STACKSIZE"] = "512m"
    e["OMP_NUM_THREADS"] = "4"
    return e


def run(cmd, cwd):
    full = "ulimit -s unlimited; " + cmd
    subprocess.run(["bash", "-c", full], cwd=cwd, check=True, env=env())


def load_dump(d):
    out = {}
    with open(os.path.join(d, "manifest.txt")) as f:
        for line in f:
            parts = line.split()
            name, dt, dims = parts[0], parts[1], [int(x) for x in parts[2:]]
            a = np.fromfile(os.path.join(d, name + ".bin"), DT[dt])
            out[name] = a.reshape(dims, order="F")
    return out


def main(selected):
    for (name, ntheta, numk, nfill, itmax, seeded, xinm, nlayers, keep) in CASES:
        if selected and name not in selected:
            continue
        print("case", name, flush=True)
        subprocess.run([REF_BUILD, name, str(ntheta), str(numk), f"{nfill:.1f}", str(itmax),
                        "1" if seeded else "0", f"{xinm:.1f}", str(nlayers)], check=True,
                       stdout=subprocess.DEVNULL)
        cdir = os.path.join(REF_OUT, name)
        arrays = {}
        if not seeded:
            ddir = os.path.join(cdir, "dump")
            shutil.rmtree(ddir, ignore_errors=True)
            os.makedirs(ddir)
            run("../harness.x > harness.log 2>&1", ddir)
            dump = load_dump(ddir)
            for k, v in dump.items():
                if k in LARGE and (keep is False or (isinstance(keep, set) and k not in keep)):
                    continue
                arrays[k] = v
        rdir = os.path.join(cdir, "run")
        shutil.rmtree(rdir, ignore_errors=True)
        os.makedirs(os.path.join(rdir, "output"))
        os.makedirs(os.path.join(rdir, "dataFock"))
        ndim = nlayers * 2 * (3 * ntheta * ntheta + 3 * ntheta + 1)
        if seeded:
            rho = seeded_rho(ndim, 4, nfill)
            fname = checkpoint_name(1, statename="SSS", nfilling=nfill, numS=1, ntheta=ntheta,
                                    nlayers=nlayers, nrelax=0, epsilon=10.0, U=4.0, nscreen=2,
                                    xi_nm=xinm, delta=0.0, numI=1, numC=3, numk=numk)
            write_checkpoint(os.path.join(rdir, "dataFock", fname), rho)
        run("../forge.x > forge.log 2>&1", rdir)
        write_traj(os.path.join(HERE, name + ".traj.txt"), rdir,
                   f"ntheta={ntheta} nlayers={nlayers} numk={numk} nfilling={nfill} xi={xinm}nm "
                   f"itmax={itmax} start={'seeded' if seeded else 'reference'}")
        arrays["case"] = np.array([ntheta, numk, nfill, itmax, int(seeded), xinm, nlayers])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, flush=True)


def write_traj(path, rdir, header):
    """The reference's Energy and Convergence rows of one run, under a header
    naming the configuration."""
    outs = sorted(os.listdir(os.path.join(rdir, "output")))
    eng = [f for f in outs if f.startswith("Energy")][0]
    conv = [f for f in outs if f.startswith("Convergence")][0]
    with open(path, "w") as f:
        f.write(f"# reference FORGE run: {header}\n")
        f.write(f"# file: {eng}\n")
        f.write("# Energy: it Etot 2Ekin 2EF EH EU\n")
        f.write(open(os.path.join(rdir, "output", eng)).read())
        f.write("# Convergence: it maxdDensity Step dE\n")
        f.write(open(os.path.join(rdir, "output", conv)).read())


# C3-reduced k-set (nenforceC3 = 1): name, ntheta, numk, nfilling, itmax, nlayers
C3_CASES = [
    ("t1k3", 1, 3, 1.0, 12, 2),
    ("tri1k3", 1, 3, 2.0, 6, 3),
    ("t2k3", 2, 3, 1.0, 6, 2),
    ("t1k6", 1, 6, 1.0, 4, 2),
]


def c3_fixtures():
    """Trajectories of the reference program built with nenforceC3 = 1
    (Solve_C3 / GetFock_C3, HartreeFock.f90:131-398, 564-881) from its own
    first diagonalisation: tests/golden/<case>_c3.traj.txt."""
    e = dict(os.environ, C3="1")
    for (name, ntheta, numk, nfill, itmax, nlayers) in C3_CASES:
        cname = name + "_c3"
        subprocess.run([REF_BUILD, cname, str(ntheta), str(numk), f"{nfill:.1f}", str(itmax), "0",
                        "40.0", str(nlayers)], check=True, stdout=subprocess.DEVNULL, env=e)
        rdir = os.path.join(REF_OUT, cname, "run")
        shutil.rmtree(rdir, ignore_errors=True)
        os.makedirs(os.path.join(rdir, "output"))
        os.makedirs(os.path.join(rdir, "dataFock"))
        run("../forge.x > forge.log 2>&1", rdir)
        if "ERROR C3" in open(os.path.join(rdir, "forge.log")).read():
            raise RuntimeError(f"{cname}: the reference reports a cell that is not C3-symmetric")
        write_traj(os.path.join(HERE, cname + ".traj.txt"), rdir,
                   f"ntheta={ntheta} nlayers={nlayers} numk={numk} nfilling={nfill} xi=40.0nm "
                   f"itmax={itmax} start=reference nenforceC3=1")
        print("wrote", cname, flush=True)


# Two spins (numS = 2): name, ntheta, numk, nfilling, itmax, split start
SPIN_CASES = [
    ("t1k3", 1, 3, 1.0, 8, False),
    ("t1k3s", 1, 3, 1.0, 8, True),
    ("t2k3s", 2, 3, 1.0, 6, True),
]


def spin_fixtures():
    """Trajectories of the reference program built with numS = numSIn = 2
    (Main.f90:364-679 with two spins): from its own first diagonalisation, or
    from two checkpoint files with spin-split densities (Main.f90:201-235):
    tests/golden/<case>_s2.traj.txt."""
    e = dict(os.environ, NUMS="2")
    for (name, ntheta, numk, nfill, itmax, split) in SPIN_CASES:
        cname = name + "_s2"
        subprocess.run([REF_BUILD, cname, str(ntheta), str(numk), f"{nfill:.1f}", str(itmax),
                        "1" if split else "0", "40.0", "2"], check=True, stdout=subprocess.DEVNULL,
                       env=e)
        rdir = os.path.join(REF_OUT, cname, "run")
        shutil.rmtree(rdir, ignore_errors=True)
        os.makedirs(os.path.join(rdir, "output"))
        os.makedirs(os.path.join(rdir, "dataFock"))
        if split:
            ndim = 2 * 2 * (3 * ntheta * ntheta + 3 * ntheta + 1)
            for nspin, rho in enumerate(seeded_rho_2spins(ndim, 4, nfill), start=1):
                fname = checkpoint_name(nspin, statename="SSS", nfilling=nfill, numS=2,
                                        ntheta=ntheta, nlayers=2, nrelax=0, epsilon=10.0, U=4.0,
                                        nscreen=2, xi_nm=40.0, delta=0.0, numI=1, numC=3, numk=numk)
                write_checkpoint(os.path.join(rdir, "dataFock", fname), rho)
        run("../forge.x > forge.log 2>&1", rdir)
        write_traj(os.path.join(HERE, cname + ".traj.txt"), rdir,
                   f"ntheta={ntheta} nlayers=2 numk={numk} nfilling={nfill} xi=40.0nm "
                   f"itmax={itmax} start={'split' if split else 'reference'} numS=2")
        print("wrote", cname, flush=True)


# Symmetry enforcement (Main.f90:514-534): name, ntheta, numk, nfilling, itmax,
# seeded start, flags (T, C2, C2T)
SYM_CASES = [
    ("t1k3_T", 1, 3, 1.0, 8, False, (1, 0, 0)),
    ("t1k3s_C2", 1, 3, 1.0, 8, True, (0, 1, 0)),
    ("t1k3s_C2T", 1, 3, 1.0, 8, True, (0, 0, 1)),
    ("t2k3s_TC2C2T", 2, 3, 1.0, 6, True, (1, 1, 1)),
]


def sym_fixtures():
    """Trajectories of the reference program built with nenforceT /
    nenforceC2 / nenforceC2T set: tests/golden/<case>.traj.txt."""
    for (name, ntheta, numk, nfill, itmax, seeded, (ft, fc2, fc2t)) in SYM_CASES:
        e = dict(os.environ, ENFT=str(ft), ENFC2=str(fc2), ENFC2T=str(fc2t))
        subprocess.run([REF_BUILD, name, str(ntheta), str(numk), f"{nfill:.1f}", str(itmax),
                        "1" if seeded else "0", "40.0", "2"], check=True, stdout=subprocess.DEVNULL,
                       env=e)
        rdir = os.path.join(REF_OUT, name, "run")
        shutil.rmtree(rdir, ignore_errors=True)
        os.makedirs(os.path.join(rdir, "output"))
        os.makedirs(os.path.join(rdir, "dataFock"))
        if seeded:
            ndim = 2 * 2 * (3 * ntheta * ntheta + 3 * ntheta + 1)
            fname = checkpoint_name(1, statename="SSS", nfilling=nfill, numS=1, ntheta=ntheta,
                                    nlayers=2, nrelax=0, epsilon=10.0, U=4.0, nscreen=2,
                                    xi_nm=40.0, delta=0.0, numI=1, numC=3, numk=numk)
            write_checkpoint(os.path.join(rdir, "dataFock", fname), seeded_rho(ndim, 4, nfill))
        run("../forge.x > forge.log 2>&1", rdir)
        if "ERROR C2" in open(os.path.join(rdir, "forge.log")).read():
            raise RuntimeError(f"{name}: the reference reports a cell that is not C2-symmetric")
        write_traj(os.path.join(HERE, name + ".traj.txt"), rdir,
                   f"ntheta={ntheta} nlayers=2 numk={numk} nfilling={nfill} xi=40.0nm "
                   f"itmax={itmax} start={'seeded' if seeded else 'reference'} "
                   f"nenforceT={ft} nenforceC2={fc2} nenforceC2T={fc2t}")
        print("wrote", name, flush=True)


# Combinations: name, ntheta, numk, nfilling, itmax, start ("reference",
# "split": one seeded checkpoint per spin), settings
COMBO_CASES = [
    ("t1k3_c3T", 1, 3, 1.0, 8, "reference", dict(numS=1, nenforceC3=1, nenforceT=1)),
    ("t1k3s_s2C2", 1, 3, 1.0, 8, "split", dict(numS=2, nenforceC2=1)),
    ("t1k3s_s2c3", 1, 3, 1.0, 8, "split", dict(numS=2, nenforceC3=1)),
    ("t1k3_V", 1, 3, 1.0, 6, "reference", dict(numS=1, nenforceValley=1)),
    ("t1k3s_V", 1, 3, 1.0, 8, "split", dict(numS=1, nenforceValley=1)),
    ("t1k3_c3V", 1, 3, 1.0, 6, "reference", dict(numS=1, nenforceC3=1, nenforceValley=1)),
    ("t1k3s_c3V", 1, 3, 1.0, 8, "split", dict(numS=1, nenforceC3=1, nenforceValley=1)),
]


def combo_fixtures(only=None):
    """Trajectories of the reference program built with several of the
    settings at once (C3 with T; two spins with C2 or with C3), and with the
    valley projection, alone and with C3: tests/golden/<case>.traj.txt.  The
    header lists every setting.  only: the case names to (re)write."""
    keys = ("numS", "nenforceC3", "nenforceT", "nenforceC2", "nenforceC2T", "nenforceValley")
    env_of = dict(numS="NUMS", nenforceC3="C3", nenforceT="ENFT", nenforceC2="ENFC2",
                  nenforceC2T="ENFC2T", nenforceValley="ENFV")
    for (name, ntheta, numk, nfill, itmax, start, st) in COMBO_CASES:
        if only and name not in only:
            continue
        full = {k: st.get(k, 1 if k == "numS" else 0) for k in keys}
        e = dict(os.environ, **{env_of[k]: str(v) for k, v in full.items()})
        subprocess.run([REF_BUILD, name, str(ntheta), str(numk), f"{nfill:.1f}", str(itmax),
                        "0" if start == "reference" else "1", "40.0", "2"], check=True,
                       stdout=subprocess.DEVNULL, env=e)
        rdir = os.path.join(REF_OUT, name, "run")
        shutil.rmtree(rdir, ignore_errors=True)
        os.makedirs(os.path.join(rdir, "output"))
        os.makedirs(os.path.join(rdir, "dataFock"))
        if start == "split":
            ndim = 2 * 2 * (3 * ntheta * ntheta + 3 * ntheta + 1)
            rhos = seeded_rho_2spins(ndim, 4, nfill) if full["numS"] == 2 else [seeded_rho(ndim, 4, nfill)]
            for nspin, rho in enumerate(rhos, start=1):
                fname = checkpoint_name(nspin, statename="SSS", nfilling=nfill, numS=full["numS"],
                                        ntheta=ntheta, nlayers=2, nrelax=0, epsilon=10.0, U=4.0,
                                        nscreen=2, xi_nm=40.0, delta=0.0, numI=1, numC=3, numk=numk)
                write_checkpoint(os.path.join(rdir, "dataFock", fname), rho)
        run("../forge.x > forge.log 2>&1", rdir)
        log = open(os.path.join(rdir, "forge.log")).read()
        if "ERROR C2" in log or "ERROR C3" in log:
            raise RuntimeError(f"{name}: the reference reports a cell that is not symmetric")
        write_traj(os.path.join(HERE, name + ".traj.txt"), rdir,
                   f"ntheta={ntheta} nlayers=2 numk={numk} nfilling={nfill} xi=40.0nm "
                   f"itmax={itmax} start={start} " + " ".join(f"{k}={v}" for k, v in full.items()))
        print("wrote", name, flush=True)


RELAX_CELLS = [2, 3, 5]


def relax_fixtures():
    """The bilayer cell before and after the reference's LatticeRelaxation
    (Geometry.f90:279-1135, applied as Main.f90:156-160), from the
    project-written oracle/ref_relax.f90: tests/golden/relax_t<ntheta>.npz."""
    for t in RELAX_CELLS:
        name = f"relax_t{t}"
        subprocess.run([REF_BUILD, name, str(t), "3", "1.0", "1", "0", "40.0", "2"], check=True,
                       stdout=subprocess.DEVNULL)
        d = os.path.join(REF_OUT, name, "relax")
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        r = subprocess.run(["../relax.x"], cwd=d, capture_output=True, text=True, check=True)
        if "ERROR" in r.stdout:
            raise RuntimeError(f"{name}: {r.stdout}")
        ndim = 2 * 2 * (3 * t * t + 3 * t + 1)
        arrays = {k: np.fromfile(os.path.join(d, k + ".bin"), "<f8").reshape((ndim, 3), order="F")
                  for k in ("Coords", "CoordsRelaxed")}
        np.savez_compressed(os.path.join(HERE, name + ".npz"), ntheta=np.array(t), **arrays)
        print("wrote", name, flush=True)


def checkpoint_fixture():
    """A density checkpoint written by the reference itself (nWriteFock=1,
    Main.f90:647-677) after one iteration of case t1k3."""
    e = dict(os.environ, WRITEFOCK="1")
    subprocess.run([REF_BUILD, "t1k3w", "1", "3", "1.0", "1", "0", "40.0", "2"], check=True,
                   stdout=subprocess.DEVNULL, env=e)
    rdir = os.path.join(REF_OUT, "t1k3w", "run")
    shutil.rmtree(rdir, ignore_errors=True)
    os.makedirs(os.path.join(rdir, "output"))
    os.makedirs(os.path.join(rdir, "dataFock"))
    run("../forge.x > forge.log 2>&1", rdir)
    (fname,) = os.listdir(os.path.join(rdir, "dataFock"))
    shutil.copy(os.path.join(rdir, "dataFock", fname), os.path.join(HERE, "t1k3_it1.ckpt"))


def bands_fixture():
    """The Bands file the reference program writes with nPrintOutputs = 1
    (PlotBands, TightBinding.f90:275-369, after the last iteration) for the
    28-orbital cell on a 6 x 6 k grid, filling +1, 3 iterations from the
    reference's own start: tests/golden/t1k6.bands.txt."""
    e = dict(os.environ, PRINTOUT="1")
    subprocess.run([REF_BUILD, "t1k6p", "1", "6", "1.0", "3", "0", "40.0", "2"], check=True,
                   stdout=subprocess.DEVNULL, env=e)
    rdir = os.path.join(REF_OUT, "t1k6p", "run")
    shutil.rmtree(rdir, ignore_errors=True)
    os.makedirs(os.path.join(rdir, "output"))
    os.makedirs(os.path.join(rdir, "dataFock"))
    run("../forge.x > forge.log 2>&1", rdir)
    (fname,) = [f for f in os.listdir(os.path.join(rdir, "output")) if f.startswith("Bands")]
    shutil.copy(os.path.join(rdir, "output", fname), os.path.join(HERE, "t1k6.bands.txt"))
    print("wrote t1k6.bands.txt from", fname, flush=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    combo_only = [a.split("=", 1)[1] for a in args if a.startswith("combo=")]
    args = ["combo" if a.startswith("combo=") else a for a in args]
    cases = [a for a in args if a not in ("ckpt", "bands", "c3", "spin", "sym", "combo", "relax")]
    if not args or cases:
        main(cases)
    if not args or "ckpt" in args:
        checkpoint_fixture()
    if not args or "bands" in args:
        bands_fixture()
    if not args or "c3" in args:
        c3_fixtures()
    if not args or "spin" in args:
        spin_fixtures()
    if not args or "sym" in args:
        sym_fixtures()
    if not args or "combo" in args:
        combo_fixtures(combo_only or None)
    if not args or "relax" in args:
        relax_fixtures()
