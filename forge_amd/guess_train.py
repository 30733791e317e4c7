"""Training of the learned-guess MLP on SCF solutions of this engine (GPU).

    python -m forge_amd.guess_train [--cells 6-15] [--tri-cells 6-10]
        [--fillings=-2,-1,0,1,2] [--data PATH] [--out forge_amd/data/guess_mlp.npz]
        [--steps 4000]

1. Data: for every cell index m and filling nu -- twisted bilayers (--cells)
   and alternating-twist trilayers (--tri-cells, so that the middle layer's
   descriptors are seen in training) -- the first diagonalisation (rho_TB)
   and a plain FP64 SCF from the seeded start (linear mixing 0.3, Fermi
   smearing 1 meV, |dE| < 1e-6 eV); the converged site densities and the
   nearest-neighbour bond values of both densities are written, and the
   descriptors and targets of every site (forge_amd.learned) are formed from
   them when training.  The largest cell index of each layer count is held
   out for validation.
2. Training: Adam on the library's MLP kernels (forward, backward and update
   on the GPU), standardised inputs and targets.
3. Report: validation loss and the SCF iterations from the learned start
   against the seeded start on the held-out cell.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import learned as Lr
from .geometry import build_geometry
from .hf import Engine, Setup
from .mlp import Mlp, loss_grad
from .seed import seeded_rho

HERE = os.path.dirname(os.path.abspath(__file__))
SCF = dict(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3, etol=1e-6)


def say(*a):
    print(*a, flush=True)


def solve_system(m: int, nu: float, nlayers: int = 2, max_iter: int = 250):
    """SCF of cell m (nlayers layers) at filling nu; returns the raw fields
    the targets are formed from (converged diagonal, converged and
    non-interacting bond values in nn_bonds order) and a summary row."""
    e = Engine(Setup(ntheta=m, nlayers=nlayers, numk=3, nfilling=nu, xi_nm=40.0))
    e.set_options(**SCF)
    t0 = time.time()
    bonds = Lr.nn_bonds(e.geom)
    e.scf_init(None)
    b_tb = Lr.bond_values(e.density(), bonds)
    e.scf_init(seeded_rho(e.n, e.nnc, nu))
    tr = e.scf_run(max_iter)
    rho = e.density()
    raw = (np.real(np.diagonal(rho[:, :, 0])).copy(), Lr.bond_values(rho, bonds), b_tb)
    info = (m, nu, e.n, len(tr), tr[-1, 1] - tr[-1, 8], tr[-1, 8], time.time() - t0, nlayers)
    e.close()
    return raw, info


def generate(systems, fillings, path):
    """Solves every (layers, m) of systems at every filling nu and writes the
    raw fields (npz: info rows, and per system its diagonal and bond values,
    concatenated)."""
    diag, bconv, btb, infos = [], [], [], []
    for nl, m in systems:
        for nu in fillings:
            (d, bc, bt), info = solve_system(m, nu, nl)
            say("  m=%2d nu=%+.0f n=%5d: %3d iterations, E=%.6f eV, last dE %+.1e (%.1fs), %d layers" % info)
            diag.append(d); bconv.append(bc); btb.append(bt)
            infos.append(info)
    np.savez(path, info=np.array(infos), diag=np.concatenate(diag), bconv=np.concatenate(bconv),
             btb=np.concatenate(btb))
    return path


def load_data(paths):
    """Descriptors and targets of every system in the raw data files:
    (X, Y, M, system id per site, info rows)."""
    Xs, Ys, Ms, sid, infos = [], [], [], [], []
    for p in paths:
        f = np.load(p)
        od = ob = 0
        for row in f["info"]:
            m, nu, n = int(row[0]), float(row[1]), int(row[2])
            nl = int(row[7]) if len(row) > 7 else 2
            g = build_geometry(m, nl, 3)
            bonds = Lr.nn_bonds(g)
            nb = len(bonds.i)
            X = Lr.descriptors(g, nu, bonds)
            Y, M = Lr.targets_from_values(g, f["diag"][od:od + n], f["bconv"][ob:ob + nb],
                                          f["btb"][ob:ob + nb], bonds)
            od += n
            ob += nb
            Xs.append(X); Ys.append(Y); Ms.append(M)
            sid.append(np.full(n, len(infos)))
            infos.append(row)
    return (np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms), np.concatenate(sid),
            np.array(infos))


def iterations_from(e, start_fn, max_iter=250):
    t0 = time.time()
    start_fn()
    tr = e.scf_run(max_iter)
    return len(tr), tr[-1, 1] - tr[-1, 8], time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", default="6-15", help="bilayer cell indices")
    ap.add_argument("--tri-cells", default="", help="trilayer cell indices (e.g. 6-10)")
    ap.add_argument("--fillings", default="-2,-1,0,1,2")
    ap.add_argument("--data", default="guess_data.npz")
    ap.add_argument("--out", default=os.path.join(HERE, "data", "guess_mlp.npz"))
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--reuse-data", action="store_true",
                    help="train on existing data files (comma-separated --data)")
    ap.add_argument("--data-only", action="store_true", help="generate the data and stop")
    a = ap.parse_args()
    def span(text):
        if not text:
            return []
        lo, hi = (int(x) for x in text.split("-"))
        return list(range(lo, hi + 1))

    cells, tri = span(a.cells), span(a.tri_cells)
    fillings = [float(x) for x in a.fillings.split(",")]
    paths = a.data.split(",")
    if not a.reuse_data:
        os.makedirs(os.path.dirname(os.path.abspath(paths[0])), exist_ok=True)
        say(f"[data] {len(cells)} bilayer + {len(tri)} trilayer cells x {len(fillings)} fillings")
        generate([(2, m) for m in cells] + [(3, m) for m in tri], fillings, paths[0])
        if a.data_only:
            return 0
    X, Y, M, sid, info = load_data(paths)
    nls = info[:, 7].astype(int) if info.shape[1] > 7 else np.full(len(info), 2)
    held = np.zeros(len(info), bool)
    for nl in np.unique(nls):
        sel = nls == nl
        held |= sel & (info[:, 0] == info[sel, 0].max())
    val = np.isin(sid, np.nonzero(held)[0])
    say(f"[train] {int((~val).sum())} training sites, {int(val.sum())} validation sites "
        f"(held out: " + ", ".join(f"m = {int(info[held & (nls == nl), 0].max())}, {nl} layers"
                                    for nl in np.unique(nls)) + ")")
    t0 = time.time()
    model, hist = Mlp.fit(X[~val], Y[~val], M[~val], steps=a.steps, batch=a.batch, lr=a.lr)
    say(f"[train] {a.steps} Adam steps in {time.time() - t0:.1f}s; batch loss "
        f"{hist[:100].mean():.4f} (first 100) -> {hist[-100:].mean():.4f} (last 100)")
    for tag, sel in (("train", ~val), ("validation", val)):
        Xn = (X[sel] - model.xm) / model.xs
        Yn = (Y[sel] - model.ym) / model.ys
        lval, _ = loss_grad(model.params, Xn, Yn, M[sel])
        yp = model.predict(X[sel])
        r2 = 1 - ((yp - Y[sel]) ** 2).sum(0) / np.maximum(((Y[sel] - Y[sel].mean(0)) ** 2).sum(0), 1e-30)
        say(f"[train] {tag}: standardised loss {lval:.4f}; R^2 per output {np.round(r2, 3)}")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    say(f"[train] model written to {a.out}")
    # each held-out cell (bilayer at nu = +1, trilayer at nu = +2): seeded
    # start against the learned start, with linear mixing 0.3 and with
    # optimal damping
    checks = [(2, cells[-1], 1.0)] if cells else []
    if tri:
        checks.append((3, tri[-1], 2.0))
    for nl, m, nu in checks:
        e = Engine(Setup(ntheta=m, nlayers=nl, numk=3, nfilling=nu, xi_nm=40.0))
        for mixer, opts in (("linear 0.3", SCF), ("optimal damping", dict(SCF, mix_mode=0))):
            e.set_options(**opts)
            its_s = iterations_from(e, lambda: e.scf_init(seeded_rho(e.n, e.nnc, nu)))
            its_tb = iterations_from(e, lambda: e.scf_init(None))
            its_l = iterations_from(e, lambda: Lr.learned_start(e, model, nu))
            for tag, r in (("seeded", its_s), ("first diagonalisation", its_tb), ("learned", its_l)):
                say(f"[check] {nl} layers m={m} nu={nu:+.0f} {mixer}, start {tag:22s}: {r[0]:3d} "
                    f"iterations, E={r[1]:.6f} eV ({r[2]:.1f}s)")
        e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
