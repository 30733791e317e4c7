"""Learned initial guess: descriptors, targets and the guess density.

A small MLP maps per-site local stacking descriptors to the converged SCF
fields of that site: its charge (the on-site Hartree shift follows from it)
and the three nearest-neighbour bond values of the density matrix (the bond
Fock fields follow from them).  The network only proposes the initial state;
the plain SCF converges the result and judges the proposal's error.

Descriptors of site i (local, so one network serves every cell index m and
layer count):
  * sublattice (2, one-hot) and the layer's place in the stack: bottom,
    middle, top (3, one-hot);
  * the stacking registry with the layer below and with the layer above
    (6 each; zeros and a 0 flag where there is none): with G_1, G_2 the
    reciprocal vectors of that layer's lattice and o one of its A sites, the
    phases phi_k = G_k . (r_i - o) for k = 1, 2 and phi_3 = -phi_1 - phi_2,
    as cos and sin.  They are the local interlayer displacement modulo that
    layer's lattice: site i on top of an A site of the other layer has all
    phases 0 (AA), on top of a B site or of a hexagon centre the AB and BA
    values;
  * presence flags of the layers below and above (2);
  * filling nu / 4 and 10 / m.
Targets of site i (from an FP64 SCF converged by this engine), with
B_s = rho*_ij - rho_TB_ij the shift of the bond to the nearest neighbour j in
angle slot s = 0, 1, 2 of the layer's own frame (rho_TB: the
non-interacting ground state of the first diagonalisation):
  * n * (rho_ii - 1/2), the site charge in units of the mean charge per site;
  * the C3 harmonics of the three bond shifts, Re and Im of
    b_h = (1/3) sum_s w^(h s) B_s, w = exp(2 pi i / 3), h = 0, 1, 2:
    b_0 is the bond shift common to the three bonds (its real part the
    symmetric renormalisation of the hopping by the Fock field), b_1 and b_2
    the bond-direction-dependent (C3-breaking) parts.
The charge and Re b_0 follow the local stacking and carry over to cells
outside the training set; Im b_0, b_1 and b_2 are the spontaneously broken
(time-reversal and C3) parts of the converged states, which differ between
cells (held-out R^2 below 0).  They are still proposed: they seed a broken
symmetry that the plain SCF then settles.
The guess density is rho_TB with its diagonal set to 1/2 + (predicted
charge) / n and its nearest-neighbour bonds shifted by the predicted
B_s = sum_h w^(-h s) b_h (both orientations of every bond, Hermitian).

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
D_IN = 21
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


def bond_vectors(g: Geometry, bonds: Bonds) -> np.ndarray:
    """(nbonds, 2) in-plane vector r_j' - r_i of every directed bond i -> j
    (r_j' the image of r_j the stored entry pairs with)."""
    C = np.asarray(g.Coords)[:, :2]
    T = cell_shifts(g)
    j = np.where(bonds.conj, bonds.row, bonds.col)
    pj = np.where(bonds.conj[:, None], C[j] + T[bonds.cell], C[j] - T[bonds.cell])
    return pj - C[bonds.i]


def layer_lattices(g: Geometry, bonds: Bonds | None = None):
    """Per layer: reciprocal vectors G (2 x 2, rows G_1, G_2) and an A site o
    (2,).  The Bravais vectors a_1 = d_0 - d_1, a_2 = d_1 - d_2 come from the
    mean bond vectors d_s of the layer's A sites in angle slot s."""
    bonds = bonds or nn_bonds(g)
    C = np.asarray(g.Coords)[:, :2]
    layer, sub = site_labels(g)
    d = bond_vectors(g, bonds)
    out = []
    for l in range(g.nlayers):
        sel = (layer[bonds.i] == l) & (sub[bonds.i] == 0)
        ds = np.array([d[sel & (bonds.slot == q)].mean(0) for q in range(3)])
        a = np.array([ds[0] - ds[1], ds[1] - ds[2]])  # rows a_1, a_2
        G = 2.0 * np.pi * np.linalg.inv(a).T           # rows G_k with G_k . a_m = 2 pi delta
        o = C[np.nonzero((layer == l) & (sub == 0))[0][0]]
        out.append((G, o))
    return out


def descriptors(g: Geometry, nfilling: float, bonds: Bonds | None = None) -> np.ndarray:
    """(n, 21) float32 site descriptors."""
    bonds = bonds or nn_bonds(g)
    n = g.ndim
    L = g.nlayers
    layer, sub = site_labels(g)
    C = np.asarray(g.Coords)[:, :2]
    lat = layer_lattices(g, bonds)
    X = np.zeros((n, D_IN))
    X[np.arange(n), sub] = 1.0
    place = np.where(layer == 0, 0, np.where(layer == L - 1, 2, 1))
    X[np.arange(n), 2 + place] = 1.0
    for side, dl in ((0, -1), (1, 1)):  # the layer below, the layer above
        other = layer + dl
        has = (other >= 0) & (other < L)
        col = 5 + 6 * side
        for l in range(L):
            rows = np.nonzero((layer == l) & has)[0]
            if rows.size == 0:
                continue
            G, o = lat[l + dl]
            ph = (C[rows] - o) @ G.T  # phi_1, phi_2
            ph = np.concatenate([ph, -ph.sum(1, keepdims=True)], 1)
            X[rows, col:col + 6:2] = np.cos(ph)
            X[rows, col + 1:col + 6:2] = np.sin(ph)
        X[:, 17 + side] = has
    X[:, 19] = nfilling / 4.0
    X[:, 20] = 10.0 / g.ntheta
    return X.astype(np.float32)


def bond_values(rho: np.ndarray, bonds: Bonds) -> np.ndarray:
    """rho on every directed bond i -> j (the stored entry, conjugated where
    the bond is stored transposed)."""
    v = rho[bonds.row, bonds.col, bonds.cell]
    return np.where(bonds.conj, np.conj(v), v)


def targets_from_values(g: Geometry, diag_conv: np.ndarray, bond_conv: np.ndarray,
                        bond_tb: np.ndarray, bonds: Bonds | None = None):
    """Targets from the converged diagonal and the bond values (bond_values
    order) of the converged and the non-interacting density."""
    bonds = bonds or nn_bonds(g)
    n = g.ndim
    Y = np.zeros((n, D_OUT))
    Y[:, 0] = n * (np.real(diag_conv) - 0.5)
    B = np.zeros((n, 3), np.complex128)
    B[bonds.i, bonds.slot] = bond_conv - bond_tb
    Y[:, 1:] = to_harmonics(B)
    mask = np.ones((n, D_OUT))
    return Y.astype(np.float32), mask.astype(np.float32)


_W = np.exp(2j * np.pi / 3.0)


def to_harmonics(B: np.ndarray) -> np.ndarray:
    """(n, 3) complex slot shifts -> (n, 6) Re, Im of b_0, b_1, b_2."""
    s = np.arange(3)
    H = np.stack([(B * _W ** (h * s)).sum(1) / 3.0 for h in range(3)], 1)
    out = np.empty((B.shape[0], 6))
    out[:, 0::2] = H.real
    out[:, 1::2] = H.imag
    return out


def from_harmonics(Yh: np.ndarray) -> np.ndarray:
    """(n, 6) Re, Im of b_0, b_1, b_2 -> (n, 3) complex slot shifts."""
    H = Yh[:, 0::2].astype(np.float64) + 1j * Yh[:, 1::2].astype(np.float64)
    return np.stack([sum(H[:, h] * _W ** (-h * s) for h in range(3)) for s in range(3)], 1)


def targets(g: Geometry, rho_conv: np.ndarray, rho_tb: np.ndarray, bonds: Bonds | None = None):
    """(Y (n, 7) float32, mask (n, 7) float32) from a converged and a
    non-interacting density (n x n x nnc, Fortran order)."""
    bonds = bonds or nn_bonds(g)
    return targets_from_values(g, np.diagonal(rho_conv[:, :, 0]), bond_values(rho_conv, bonds),
                               bond_values(rho_tb, bonds), bonds)


def guess_updates(g: Geometry, Yp: np.ndarray, bonds: Bonds | None = None):
    """The guess as changes to rho_TB: (diag (n), row, col, cell, delta) with
    every stored entry once.  Each bond is predicted from both its ends; the
    two predictions of an entry are averaged, and a home-cell entry also
    receives the conjugate of its transpose's prediction (Hermiticity)."""
    bonds = bonds or nn_bonds(g)
    n = g.ndim
    diag = 0.5 + Yp[:, 0].astype(np.float64) / n
    d = from_harmonics(Yp[:, 1:])[bonds.i, bonds.slot]
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
    bonds = nn_bonds(engine.geom)
    Yp = model.predict(descriptors(engine.geom, nfilling, bonds), bf16=bf16)
    if diagonalise:
        engine.scf_init_guess_update(*guess_updates(engine.geom, Yp, bonds))
    else:
        engine.scf_init_update(*guess_updates(engine.geom, Yp, bonds))
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
