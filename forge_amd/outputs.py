"""The reference's per-iteration output files, written from a trajectory of
forge_scf_run / Engine.scf_run (rows: it, E_in, 2 Ekin_in, 2 EF_in, EH_in,
EU_in, max|dDensity|, Step, E_in - E_new, E_F, nOcc, nPart).

  output/Energy<parameters>       Main.f90:352-354, 626-631 (numS = 1):
      F5.1 it, F24.9 E_tot, 2 Ekin, 2 EF, EH, EU, separated by 3X
  output/Convergence<parameters>  Main.f90:356-358, 616-620:
      F5.1 it, F9.6 max|dDensity|, F5.3 Step, F24.9 E_in - E_new

<parameters> is the run-parameter string of Main.f90:180-183
(forge_amd.checkpoint.run_parameters).  A trailing 3X at the end of a
record writes nothing, so lines carry no trailing blanks.
"""
from __future__ import annotations

import os

import numpy as np


def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d output of a real: right-justified in w columns; the zero
    before the decimal point is dropped when only that makes it fit, and a
    value that does not fit prints w asterisks."""
    s = f"{x:.{d}f}"
    if len(s) > w:
        if s.startswith("0."):
            s = s[1:]
        elif s.startswith("-0."):
            s = "-" + s[2:]
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def energy_line(row) -> str:
    return "   ".join([fortran_f(row[0], 5, 1)] + [fortran_f(v, 24, 9) for v in row[1:6]])


def convergence_line(row) -> str:
    return "   ".join([fortran_f(row[0], 5, 1), fortran_f(row[6], 9, 6), fortran_f(row[7], 5, 3),
                       fortran_f(row[8], 24, 9)])


def write_outputs(outdir: str, parameters: str, traj: np.ndarray, append: bool = False):
    """Writes output/Energy<parameters> and output/Convergence<parameters>
    under outdir (created if needed); append continues a run's files.
    Returns the two paths."""
    d = os.path.join(outdir, "output")
    os.makedirs(d, exist_ok=True)
    paths = (os.path.join(d, "Energy" + parameters), os.path.join(d, "Convergence" + parameters))
    mode = "a" if append else "w"
    with open(paths[0], mode) as fe, open(paths[1], mode) as fc:
        for row in np.asarray(traj, dtype=float):
            fe.write(energy_line(row) + "\n")
            fc.write(convergence_line(row) + "\n")
    return paths
