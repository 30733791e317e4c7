"""Host-side geometry of the commensurate twisted multilayer cell.

Follows the geometry block of the reference driver (Main.f90:123-154) and
Geometry.f90: NumberNeighborCells (:14-35), OrderNeighborCells (:40-91),
SampleBZ (:96-127) and WignerSeitzCell (:132-273).  The site order fixes the
matrix index order of every array the engine exchanges with a reference host,
so the loops keep the reference's order (n1 outer, n2 inner; per layer the A
sublattice, its zone-boundary site, the B sublattice, its boundary site).
Vectorised with numpy; the arithmetic per site is the reference's.
"""
from __future__ import annotations

import math
import os

from dataclasses import dataclass

import numpy as np

A1 = np.array([0.5, np.sqrt(3.0) * 0.5])
A2 = np.array([-0.5, np.sqrt(3.0) * 0.5])
EPS6 = float(np.float32(1e-6))  # default-real literal in WignerSeitzCell


def ndim_of(ntheta: int, nlayers: int) -> int:
    return nlayers * 2 * (ntheta * ntheta + (ntheta + 1) ** 2 + (ntheta + 1) * ntheta)


def number_neighbor_cells(numI: int) -> int:
    c = 0
    for n1 in range(-numI, numI + 1):
        for n2 in range(-numI, numI + 1):
            if -numI <= n1 + n2 <= numI and (n1 != 0 or n2 != 0):
                c += 1
    return c // 2 + 1


def order_neighbor_cells(numI: int, nnc: int):
    u1 = [0] * nnc
    u2 = [0] * nnc
    ic = 0
    if numI > 0:
        u1[1:4] = [-1, -1, 0]
        u2[1:4] = [0, 1, -1]
        ic = 3
    if numI > 1:
        for n1 in range(-numI, numI + 1):
            for n2 in range(-numI, numI + 1):
                if -numI <= n1 + n2 <= numI and ic + 1 < nnc and (n1 != 0 or n2 != 0):
                    ic += 1
                    u1[ic], u2[ic] = n1, n2
                    if (n1, n2) in ((-1, 0), (-1, 1), (0, -1)):
                        ic -= 1
    return np.array(u1, dtype=np.int64), np.array(u2, dtype=np.int64)


def _layer_sites(ntheta, rot, cs, sn, t1, t2):
    nrad = 3 * ntheta
    rmax = 3.0 * ntheta * ntheta + 3.0 * ntheta + 1.0
    A = 3.0 * ntheta + 1
    B = 3.0 * ntheta + 2
    sq3 = np.sqrt(3.0)
    n1, n2 = np.meshgrid(np.arange(-nrad, nrad + 1, dtype=np.float64),
                         np.arange(-nrad, nrad + 1, dtype=np.float64), indexing="ij")
    n1 = n1.ravel()
    n2 = n2.ravel()

    def inside(a, b):
        r1 = np.abs(a * A + b * B) + EPS6
        r2 = np.abs(a - b * A) + EPS6
        r3 = np.abs(a * B + b) + EPS6
        return (r1 < rmax) & (r2 < rmax) & (r3 < rmax)

    out = []
    if rot == -1:
        an1 = n1 + 1.0 / 3.0
        an2 = n2 - 2.0 / 3.0
        m = inside(an1, an2)
        out.append(np.stack([(n1[m] + 1.0 / 3.0) * A1[0] + (n2[m] - 2.0 / 3.0) * A2[0],
                             (n1[m] + 1.0 / 3.0) * A1[1] + (n2[m] - 2.0 / 3.0) * A2[1]], 1))
        out.append(np.array([[(-t2[0] - t1[0]) / 3.0, (-t2[1] - t1[1]) / 3.0]]))
        bn1 = n1 + 2.0 / 3.0
        bn2 = n2 - 1.0 / 3.0
        m = inside(bn1, bn2)
        out.append(np.stack([bn1[m] * A1[0] + bn2[m] * A2[0], bn1[m] * A1[1] + bn2[m] * A2[1]], 1))
        out.append(np.array([[(t2[0] + t1[0]) / 3.0, (t2[1] + t1[1]) / 3.0]]))
    else:
        an1 = (n1 + 1.0 / 3.0) * (cs - sn / sq3) - 2 * (n2 - 2.0 / 3.0) * sn / sq3
        an2 = (n2 - 2.0 / 3.0) * (cs + sn / sq3) + 2 * (n1 + 1.0 / 3.0) * sn / sq3
        m = inside(an1, an2)
        out.append(np.stack([an1[m] * A1[0] + an2[m] * A2[0], an1[m] * A1[1] + an2[m] * A2[1]], 1))
        out.append(np.array([[(t1[0] + t2[0]) / 3.0, (t1[1] + t2[1]) / 3.0]]))
        bn1 = (n1 + 2.0 / 3.0) * (cs - sn / sq3) - 2 * (n2 - 1.0 / 3.0) * sn / sq3
        bn2 = (n2 - 1.0 / 3.0) * (cs + sn / sq3) + 2 * (n1 + 2.0 / 3.0) * sn / sq3
        m = inside(bn1, bn2)
        out.append(np.stack([bn1[m] * A1[0] + bn2[m] * A2[0], bn1[m] * A1[1] + bn2[m] * A2[1]], 1))
        out.append(np.array([[(-t1[0] - t2[0]) / 3.0, (-t1[1] - t2[1]) / 3.0]]))
    return np.concatenate(out, 0)


@dataclass
class Geometry:
    ntheta: int
    nlayers: int
    numk: int
    numI: int
    Coords: np.ndarray            # (ndim, 3) Fortran order
    tn: np.ndarray                # (2, 3) = [t1 t2 t3]
    g1: np.ndarray
    g12: np.ndarray
    MomentaValues: np.ndarray     # (numk, numk, 2)
    nMomentaComponents: np.ndarray  # (Nk, 2)
    nUnitCell_1: np.ndarray
    nUnitCell_2: np.ndarray
    RotateLayers: np.ndarray = None  # (nlayers,) +-1 per layer (Setup.f90:8)

    @property
    def ndim(self):
        return self.Coords.shape[0]

    @property
    def nnc(self):
        return len(self.nUnitCell_1)

    def vk(self, icount: int) -> np.ndarray:
        i1, i2 = self.nMomentaComponents[icount]
        return np.array(self.MomentaValues[i1, i2, :])


def rot_apply(R, v):
    return np.array([R[0, 0] * v[0] + R[0, 1] * v[1], R[1, 0] * v[0] + R[1, 1] * v[1]])


_RELAX_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "relaxation_fourier.txt")


def relaxation_components(ntheta: int):
    """(uqx, uqy) as 5 x 5 arrays indexed (n-1, m-1): the Fourier components of
    the in-plane relaxation profile the reference holds for this cell
    (Geometry.f90:309-1056; zero for cells it has no table for)."""
    uqx = np.zeros((5, 5))
    uqy = np.zeros((5, 5))
    for line in open(_RELAX_TABLE):
        if line.startswith("#") or not line.strip():
            continue
        t, n, m, ux, uy = line.split()
        if int(t) == ntheta:
            uqx[int(n) - 1, int(m) - 1] = float(ux)
            uqy[int(n) - 1, int(m) - 1] = float(uy)
    return uqx, uqy


def lattice_relaxation(C: np.ndarray, g1: np.ndarray, g12: np.ndarray, ntheta: int) -> np.ndarray:
    """LatticeRelaxation (Geometry.f90:279-1135) of a bilayer cell: layer 1
    moves by half the Fourier displacement sum_{nn, n, m} -2 u sin(k.r) along
    the three reciprocal directions, layer 2 becomes layer 1 mirrored in x,
    re-ordered to the unrelaxed site order of each sublattice half (nearest
    site within 0.25; the half's corner takes (x, -y) of its own mirrored
    site).  Scalar math in the reference's order of operations."""
    n = C.shape[0]
    uqx, uqy = relaxation_components(ntheta)
    Gn = [np.asarray(g1, float), np.asarray(g12, float)]
    Gn.append(Gn[1] - Gn[0])
    Gn += [-Gn[0], -Gn[1], -Gn[2], Gn[0], Gn[1]]
    R = np.array(C, dtype=float, order="F", copy=True)
    for i in range(n // 2):
        rx, ry = R[i, 0], R[i, 1]
        d0 = d1 = 0.0
        for nn in (1, 3, 5):
            ca, sa = math.cos((nn - 1) * math.pi / 3.0), math.sin((nn - 1) * math.pi / 3.0)
            ga, gb = Gn[nn - 1], Gn[nn + 1]
            for N in range(2, 6):
                for M in range(1, N):
                    vx = (N - 1) * ga[0] + (M - 1) * gb[0]
                    vy = (N - 1) * ga[1] + (M - 1) * gb[1]
                    ux = ca * uqx[N - 1, M - 1] - sa * uqy[N - 1, M - 1]
                    uy = sa * uqx[N - 1, M - 1] + ca * uqy[N - 1, M - 1]
                    sk = math.sin(vx * rx + vy * ry)
                    d0 = d0 - 2 * ux * sk
                    d1 = d1 - 2 * uy * sk
        R[i, 0] = rx + d0 / 2
        R[i, 1] = ry + d1 / 2
        R[i + n // 2, 0] = -R[i, 0]
        R[i + n // 2, 1] = R[i, 1]
    out = R.copy()
    for lo, hi in ((n // 2, 3 * n // 4), (3 * n // 4, n)):
        matched = 0
        for i in range(lo, hi):
            for j in range(lo, hi):
                if math.sqrt((R[i, 0] - C[j, 0]) ** 2 + (R[i, 1] - C[j, 1]) ** 2) < 0.25:
                    matched += 1
                    out[j, :] = R[i, :]
                    break
        matched += 1
        out[hi - 1, 0] = R[hi - 1, 0]
        out[hi - 1, 1] = -R[hi - 1, 1]
        out[hi - 1, 2] = R[hi - 1, 2]
        if matched != n // 4:
            raise RuntimeError(f"lattice relaxation: {matched} of {n // 4} sites re-ordered")
    return np.asfortranarray(out)


def build_geometry(ntheta: int, nlayers: int = 2, numk: int = 3, numI: int = 1,
                   rotate=None, tz: float = 1.35772, nrelax: int = 0) -> Geometry:
    if rotate is None:
        rotate = [-1, 1] if nlayers == 2 else [(-1) ** (l + 1) for l in range(nlayers)]
    nnc = number_neighbor_cells(numI)
    u1, u2 = order_neighbor_cells(numI, nnc)
    aM = 3.0 * ntheta * ntheta + 3.0 * ntheta + 1.0
    g1 = (4.0 * np.pi / 3.0) / aM * ((3 * ntheta + 1) * A1 + A2)
    g12 = (4.0 * np.pi / 3.0) / aM * ((3 * ntheta + 2) * A2 - A1)
    cs = 1.0 - 1.0 / (2.0 * aM)
    sn = np.sqrt(1.0 - cs * cs)
    h = 0.5 * np.arccos(cs)
    R = np.array([[np.cos(h), np.sin(h)], [-np.sin(h), np.cos(h)]])
    g1 = rot_apply(R, g1)
    g12 = rot_apply(R, g12)
    Nk = numk * numk
    comp = np.array([[i1, i2] for i1 in range(numk) for i2 in range(numk)], dtype=np.int64)
    mom = np.zeros((numk, numk, 2), order="F")
    for i2 in range(numk):
        for i1 in range(numk):
            mom[i1, i2, :] = (i1 * g1 + i2 * g12) / numk
    t1 = ntheta * A1 + (ntheta + 1) * A2
    t2 = (-ntheta - 1) * A1 + (2 * ntheta + 1) * A2
    sites = []
    for l in range(nlayers):
        xy = _layer_sites(ntheta, rotate[l], cs, sn, t1, t2)
        z = float(np.float32(np.float32(l) - np.float32(nlayers - 1) / np.float32(2.0))) * tz
        sites.append(np.concatenate([xy, np.full((xy.shape[0], 1), z)], 1))
    C = np.concatenate(sites, 0)
    x = R[0, 0] * C[:, 0] + R[0, 1] * C[:, 1]
    y = R[1, 0] * C[:, 0] + R[1, 1] * C[:, 1]
    C[:, 0], C[:, 1] = x, y
    n = ndim_of(ntheta, nlayers)
    if C.shape[0] != n:
        raise RuntimeError(f"Wigner-Seitz cell holds {C.shape[0]} sites, expected {n}")
    if nrelax == 1 and nlayers == 2:
        # Main.f90:156-160: only bilayers are relaxed
        C = lattice_relaxation(C, g1, g12, ntheta)
    t1r = rot_apply(R, t1)
    t2r = rot_apply(R, t2)
    tn = np.asfortranarray(np.stack([t1r, t2r, t2r - t1r], 1))
    return Geometry(ntheta, nlayers, numk, numI, np.asfortranarray(C), tn, g1, g12, mom,
                    np.asfortranarray(comp), u1, u2, np.array(rotate, dtype=np.int64))
