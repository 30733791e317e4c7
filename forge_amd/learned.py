"""Learned initial guess: descriptors, targets and the guess density.

A small MLP maps per-site stacking descriptors to the converged SCF fields of
that site: its charge (the on-site Hartree shift follows from it) and the
three nearest-neighbour bond values of the density matrix (the bond Fock
fields follow from them).  The network only proposes the initial state; the
plain SCF converges the result and judges the proposal's error.

Descriptors of site i (all angle-independent in form, so one network serves
every cell index m):
  * sublattice (2, one-hot) and layer (2, one-hot; bilayer);
  * moire registry: the site's fractional coordinates (u, v) in the moire
    lattice t1, t2 enter through cos/sin of 2 pi h (u, v, u + v) for the
    harmonics h = 1, 2 (12); AA stacking sits at u = v = 0;
  * filling nu / 4 and 10 / m.
Targets of site i (from an FP64 SCF converged by this engine):
  * n * (rho_ii - 1/2), the site charge in units of the mean charge per site;
  * Re, Im of rho*_ij - rho_TB_ij for the three intralayer nearest neighbours
    j of i, ordered by bond angle in the layer's own frame (rho_TB: the
    non-interacting ground state of the first diagonalisation).
The guess density is rho_TB with its diagonal set to 1/2 + (predicted
charge) / n and its nearest-neighbour bonds shifted by the predicted
differences (both orientations of every bond, Hermitian).

Bond bookkeeping follows the engine's storage: rho_c(i, j) pairs r_i with
r_j - T_c, T_c = n1_c t1 + n2_c t2 (HartreeFock.f90:76-98); a pair with
r_j + T_c is stored as conj(rho_c(j, i)).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

from .geometry import Geometry

A0 = 1.0 / np.sqrt(3.0)  # carbon-carbon distance / a
D_IN = 18
D_OUT = 7


def cell_shifts(g: Geometry) -> np.ndarray:
    """T_c (nnc x 2) of the stored cells."""
    t1, t2 = g.tn[:, 0], g.tn[:, 1]
    return np.array([n1 * t1 + n2 * t2 for n1, n2 in zip(g.nUnitCell_1, g.nUnitCell_2)])


def site_labels(g: Geometry):
    """(layer, sublattice) of every site: per layer the construction emits the
    A sublattice and one boundary A site, then the B sublattice and one
    boundary B site, each half of the layer's sites."""
    n = g.ndim
    per = n // g.nlayers
    half = per // 2
    layer = np.repeat(np.arange(g.nlayers), per)
    sub = np.tile(np.r_[np.zeros(half, int), np.ones(half, int)], g.nlayers)
    return layer, sub


@dataclass
class Bonds:
    """Directed nearest-neighbour bonds i -> j: stored as rho[row, col, cell],
    conjugated when conj is set.  slot: 0..2, the bond's rank by angle at i."""
    i: np.ndarray
    row: np.ndarray
    col: np.ndarray
    cell: np.ndarray
    conj: np.ndarray
    slot: np.ndarray


def nn_bonds(g: Geometry) -> Bonds:
    C = np.asarray(g.Coords)
    T = cell_shifts(g)
    n = g.ndim
    layer, _ = site_labels(g)
    out = {k: [] for k in ("i", "j", "cell", "sign", "dx", "dy")}
    for c in range(len(T)):
        for sgn in ((0,) if c == 0 else (-1, 1)):
            P = C[:, :2] + sgn * T[c]  # r_j - T_c (sgn -1) or r_j + T_c (sgn +1)
            tree = cKDTree(P)
            pairs = tree.query_ball_point(C[:, :2], A0 * 1.01)
            for i, js in enumerate(pairs):
                for j in js:
                    if layer[i] != layer[j]:
                        continue
                    d = P[j] - C[i, :2]
                    if abs(np.hypot(*d) - A0) > 1e-6:
                        continue
                    out["i"].append(i); out["j"].append(j); out["cell"].append(c)
                    out["sign"].append(sgn); out["dx"].append(d[0]); out["dy"].append(d[1])
    i = np.array(out["i"]); j = np.array(out["j"]); c = np.array(out["cell"])
    sgn = np.array(out["sign"])
    cnt = np.bincount(i, minlength=n)
    if not np.all(cnt == 3):
        raise RuntimeError(f"nearest-neighbour count per site is not 3: {np.unique(cnt)}")
    # +T_c pairs are the conjugate of the transposed entry of cell c
    conj = sgn == 1
    row = np.where(conj, j, i)
    col = np.where(conj, i, j)
    # rank the three bonds of each site by angle: the layers are rotated by
    # +-theta/2 only, so the order is the same for every site of a sublattice
    ang = np.mod(np.arctan2(np.array(out["dy"]), np.array(out["dx"])), 2 * np.pi)
    order = np.lexsort((ang, i))
    slot = np.empty_like(order)
    slot[order] = np.tile(np.arange(3), n)
    return Bonds(i, row, col, c, conj, slot)


def descriptors(g: Geometry, nfilling: float) -> np.ndarray:
    """(n, 18) float32 site descriptors."""
    n = g.ndim
    layer, sub = site_labels(g)
    C = np.asarray(g.Coords)
    t1, t2 = g.tn[:, 0], g.tn[:, 1]
    M = np.array([t1, t2]).T  # columns t1, t2
    uv = np.linalg.solve(M, C[:, :2].T).T  # fractional moire coordinates
    u, v = uv[:, 0], uv[:, 1]
    X = np.zeros((n, D_IN))
    X[np.arange(n), sub] = 1.0
    X[np.arange(n), 2 + np.minimum(layer, 1)] = 1.0
    col = 4
    for h in (1, 2):
        for w in (u, v, u + v):
            X[:, col] = np.cos(2 * np.pi * h * w)
            X[:, col + 1] = np.sin(2 * np.pi * h * w)
            col += 2
    X[:, 16] = nfilling / 4.0
    X[:, 17] = 10.0 / g.ntheta
    return X.astype(np.float32)


def targets(g: Geometry, rho_conv: np.ndarray, rho_tb: np.ndarray, bonds: Bonds | None = None):
    """(Y (n, 7) float32, mask (n, 7) float32) from a converged and a
    non-interacting density (n x n x nnc, Fortran order)."""
    bonds = bonds or nn_bonds(g)
    n = g.ndim
    Y = np.zeros((n, D_OUT))
    Y[:, 0] = n * (np.real(np.diagonal(rho_conv[:, :, 0])) - 0.5)
    d = rho_conv[bonds.row, bonds.col, bonds.cell] - rho_tb[bonds.row, bonds.col, bonds.cell]
    d = np.where(bonds.conj, np.conj(d), d)
    Y[bonds.i, 1 + 2 * bonds.slot] = d.real
    Y[bonds.i, 2 + 2 * bonds.slot] = d.imag
    mask = np.ones((n, D_OUT))
    return Y.astype(np.float32), mask.astype(np.float32)


def guess_updates(g: Geometry, Yp: np.ndarray, bonds: Bonds | None = None):
    """The guess as changes to rho_TB: (diag (n), row, col, cell, delta) with
    every stored entry once.  Each bond is predicted from both its ends; the
    two predictions of an entry are averaged, and a home-cell entry also
    receives the conjugate of its transpose's prediction (Hermiticity)."""
    bonds = bonds or nn_bonds(g)
    n = g.ndim
    diag = 0.5 + Yp[:, 0].astype(np.float64) / n
    d = Yp[bonds.i, 1 + 2 * bonds.slot].astype(np.float64) + 1j * Yp[bonds.i, 2 + 2 * bonds.slot]
    # d shifts the entry "i -> j"; in storage it is conjugated when flagged
    ds = np.where(bonds.conj, np.conj(d), d)
    home = bonds.cell == 0
    row = np.r_[bonds.row, bonds.col[home]]
    col = np.r_[bonds.col, bonds.row[home]]
    cell = np.r_[bonds.cell, np.zeros(home.sum(), int)]
    val = np.r_[ds, np.conj(ds[home])]
    key = (cell.astype(np.int64) * n + col) * n + row
    uk, inv = np.unique(key, return_inverse=True)
    acc = np.zeros(uk.size, np.complex128)
    np.add.at(acc, inv, val)
    cnt = np.bincount(inv, minlength=uk.size)
    return (diag, (uk % n).astype(np.int64), ((uk // n) % n).astype(np.int64),
            (uk // (n * n)).astype(np.int64), acc / cnt)


def learned_start(engine, model, nfilling: float, bf16: bool = True, diagonalise: bool = True):
    """Start state of the engine's SCF from the network: the first
    diagonalisation (rho_TB, on the device), the network's site charges and
    bond shifts (bf16 matrix-core inference), then either one diagonalisation
    of H[rho_TB + proposal] (diagonalise) or the proposal itself as the start
    state, whose first SCF iteration solves H[proposal] and mixes.  Returns
    the network outputs."""
    engine.scf_init(None)
    Yp = model.predict(descriptors(engine.geom, nfilling), bf16=bf16)
    if diagonalise:
        engine.scf_init_guess_update(*guess_updates(engine.geom, Yp))
    else:
        engine.scf_init_update(*guess_updates(engine.geom, Yp))
    return Yp


def guess_density(g: Geometry, rho_tb: np.ndarray, Yp: np.ndarray,
                  bonds: Bonds | None = None) -> np.ndarray:
    """rho_TB with the predicted site charges and bond shifts applied (host
    arrays; the engine applies the same updates on the device)."""
    diag, row, col, cell, delta = guess_updates(g, Yp, bonds)
    rho = np.array(rho_tb, dtype=np.complex128, order="F", copy=True)
    idx = np.arange(g.ndim)
    rho[idx, idx, 0] = diag
    rho[row, col, cell] += delta
    return rho
