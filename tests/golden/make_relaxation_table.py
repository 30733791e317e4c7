#!/usr/bin/env python3
"""Writes forge_amd/data/relaxation_fourier.txt: the Fourier components
(uqx, uqy)(n, m) of the in-plane relaxation profile that the reference
assigns per cell (Geometry.f90:309-1056, one block per ntheta = 2..35).

They are the reference's physical data: a different table is a different
lattice, so the values are carried over as the reference's program holds
them.  Its literals without a kind suffix are default-real (single-precision)
constants, rounded to binary32 before they are stored in real(dp); that
rounding is reproduced here (correctly rounded from the decimal text), so
each line holds the double the reference computes with.

Runs only where the reference sources are mounted (/root/reference); it sits
beside the fixture generator because it reads the reference, which the
product package never does.
"""
from __future__ import annotations

import os
import re
import sys
from decimal import Decimal

import numpy as np

REF = "/root/reference/Geometry.f90"
OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                   "forge_amd", "data", "relaxation_fourier.txt")


def binary32(text: str) -> float:
    """The binary32 value nearest to the decimal text (ties to even)."""
    exact = Decimal(text)
    f = np.float32(float(text))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Decimal(float(c)) - exact),
                                     int(np.frombuffer(np.float32(c).tobytes(), np.uint32)[0]) & 1))
    return float(best)


def value(lit: str) -> float:
    lit = lit.strip()
    if lit.lower().endswith("_dp"):
        return float(lit[:-3])
    return binary32(lit)


def main() -> int:
    src = open(REF).read().splitlines()
    table: dict[int, dict[tuple[int, int], list[float]]] = {}
    cur = None
    for line in src:
        m = re.match(r"\s*if\s*\(\s*ntheta\s*\.EQ\.\s*(\d+)\s*\)\s*then", line, re.I)
        if m:
            cur = int(m.group(1))
            table[cur] = {}
            continue
        if cur is not None and re.match(r"\s*endif", line, re.I):
            cur = None
            continue
        m = re.match(r"\s*uq([xy])\((\d),(\d)\)\s*=\s*([^!\s]+)", line)
        if m and cur is not None:
            comp, n, mm, lit = m.group(1), int(m.group(2)), int(m.group(3)), m.group(4)
            e = table[cur].setdefault((n, mm), [0.0, 0.0])
            e[0 if comp == "x" else 1] = value(lit)
    with open(OUT, "w") as f:
        f.write("# ntheta n m uqx uqy: Fourier components of the relaxation profile\n")
        f.write("# (binary64 values as the reference stores them), written by tests/golden/make_relaxation_table.py\n")
        for t in sorted(table):
            for (n, mm) in sorted(table[t]):
                ux, uy = table[t][(n, mm)]
                f.write(f"{t} {n} {mm} {ux!r} {uy!r}\n")
    print("wrote", OUT, "for", len(table), "cells")
    return 0


if __name__ == "__main__":
    sys.exit(main())
