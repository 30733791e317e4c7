"""Readers for the committed golden fixtures (reference outputs)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def case_params(f):
    ntheta, numk, nfill, itmax, seeded, xinm, nl = f["case"]
    return dict(ntheta=int(ntheta), numk=int(numk), nfilling=float(nfill), itmax=int(itmax),
                seeded=bool(seeded), xi_nm=float(xinm), nlayers=int(nl))


def load_traj(name):
    """(energy_rows, convergence_rows) of the reference's Energy/Convergence files:
    energy rows = it, E_tot, 2E_kin, 2E_F, E_H, E_U; convergence rows =
    it, max|dDensity|, Step, E_in - E_new (Main.f90:616-639)."""
    eng, conv, cur = [], [], None
    with open(os.path.join(GOLDEN, name + ".traj.txt")) as f:
        for line in f:
            if line.startswith("# Energy"):
                cur = eng
            elif line.startswith("# Convergence"):
                cur = conv
            elif line.startswith("#") or not line.strip():
                continue
            else:
                cur.append([float(x) for x in line.split()])
    return np.array(eng), np.array(conv)


def traj_meta(name):
    """The key=value pairs of a trajectory fixture's first header line
    (ntheta, nlayers, numk, nfilling, xi, itmax, start, ...)."""
    with open(os.path.join(GOLDEN, name + ".traj.txt")) as f:
        head = f.readline()
    return dict(kv.split("=") for kv in head.split(": ", 1)[1].split())
