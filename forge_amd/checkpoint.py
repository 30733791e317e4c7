"""Density-matrix checkpoint files in the reference's on-disk layout.

Layout (Main.f90:647-671 writes, Main.f90:202-233 reads): unformatted direct
access, one complex(8) per record, no record markers.  Cell 0 is stored as the
lower triangle row by row -- for i = 1..n the elements (i, 1..i-1) and then the
diagonal (i, i) -- and cells 2..numNeighborCells follow in full with j fastest
inside each row i.  The reader rebuilds the upper triangle of cell 0 as the
conjugate of the lower one, as Main.f90:217-218 does.  The record length is
16 bytes when the file is written by flang (byte units); ifort's default
4-byte-word unit gives 64-byte records whose first 16 bytes carry the value,
which `read_checkpoint(..., record_bytes=64)` accepts.

The file name encodes the run parameters exactly as Main.f90:180-183 formats
them; `checkpoint_name` reproduces that string.
"""
from __future__ import annotations

import numpy as np


def _fmt_f(x: float, d: int) -> str:
    """Fortran F0.d: minimal width; a zero integer part is dropped (flang)."""
    s = f"{abs(x):.{d}f}"
    if s.startswith("0."):
        s = s[1:]
    return ("-" if x < 0 else "") + s


def run_parameters(statename: str, nfilling: float, numS: int, ntheta: int,
                   nlayers: int, nrelax: int, epsilon: float, U: float,
                   nscreen: int, xi_nm: float, delta: float, numI: int,
                   numC: int, numk: int, dp: int = 8) -> str:
    """The `parameters` string of Main.f90:180-183."""
    dop = "-" if nfilling < 0 else "+"
    return (f"-{statename:3.3}-filling{dop}{abs(nfilling):3.1f}-numS{numS}-i{ntheta}"
            f"-nlayers{nlayers}-relax{nrelax}-eps{_fmt_f(epsilon, 1)}-U{U:4.2f}"
            f"-screen{nscreen}-xi{_fmt_f(xi_nm, 1)}-delta{_fmt_f(delta, 3)}"
            f"-numI{numI}-numC{numC}-numk{numk}-dp{dp}.dat")


def checkpoint_name(nspin: int, **params) -> str:
    return f"Fock-nspin{nspin}" + run_parameters(**params)


def pack(rho: np.ndarray) -> np.ndarray:
    """rho(n, n, ncells) -> record stream (complex128) in file order."""
    n, _, nc = rho.shape
    il, jl = np.tril_indices(n, -1)
    recs = []
    # cell 0: per row i, off-diagonal (i, j<i) then the diagonal (i, i)
    lower = rho[:, :, 0]
    row_parts = []
    for i in range(n):
        row_parts.append(lower[i, :i])
        row_parts.append(lower[i, i:i + 1])
    recs.append(np.concatenate(row_parts) if n else np.zeros(0, np.complex128))
    for c in range(1, nc):
        recs.append(np.ascontiguousarray(rho[:, :, c]).reshape(-1))  # row i, j fastest
    del il, jl
    return np.concatenate(recs).astype(np.complex128)


def unpack(stream: np.ndarray, n: int, ncells: int) -> np.ndarray:
    need = n * (n + 1) // 2 + (ncells - 1) * n * n
    if stream.size != need:
        raise ValueError(f"checkpoint holds {stream.size} records, expected {need} "
                         f"for ndim={n}, cells={ncells}")
    rho = np.zeros((n, n, ncells), dtype=np.complex128, order="F")
    pos = 0
    for i in range(n):
        row = stream[pos:pos + i]
        rho[i, :i, 0] = row
        rho[:i, i, 0] = np.conj(row)
        rho[i, i, 0] = stream[pos + i]
        pos += i + 1
    for c in range(1, ncells):
        rho[:, :, c] = stream[pos:pos + n * n].reshape(n, n)
        pos += n * n
    return rho


def write_checkpoint(path: str, rho: np.ndarray) -> None:
    pack(rho).tofile(path)


def read_checkpoint(path: str, n: int, ncells: int, record_bytes: int = 16) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.uint8)
    if record_bytes == 16:
        stream = raw.view(np.complex128)
    else:
        if raw.size % record_bytes:
            raise ValueError(f"{path}: size {raw.size} is not a multiple of {record_bytes}")
        stream = raw.reshape(-1, record_bytes)[:, :16].copy().view(np.complex128).reshape(-1)
    return unpack(stream, n, ncells)
