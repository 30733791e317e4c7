"""The reference's per-iteration output files, written from a trajectory of
forge_scf_run / Engine.scf_run (rows: it, E_in, 2 Ekin_in, 2 EF_in, EH_in,
EU_in, max|dDensity|, Step, E_in - E_new, E_F, nOcc, nPart).

  output/Energy<parameters>       Main.f90:352-354, 626-631 (numS = 1):
      F5.1 it, F24.9 E_tot, 2 Ekin, 2 EF, EH, EU, separated by 3X
  output/Convergence<parameters>  Main.f90:356-358, 616-620:
      F5.1 it, F9.6 max|dDensity|, F5.3 Step, F24.9 E_in - E_new

  output/Bands-nspin1<parameters>  Main.f90:601-608, PlotBands
      (TightBinding.f90:275-369): F16.8 path coordinate and bands - E_F

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


def plot_bands_lines(bands: np.ndarray, fermi: float, flat: np.ndarray, numk: int,
                     g1: np.ndarray, g12: np.ndarray) -> list[str]:
    """The Bands file of PlotBands (TightBinding.f90:275-369): the window
    bands minus the Fermi energy along K - Gamma - M - K of the numk x numk
    grid, one line per k point: F16.8 path coordinate (normalised to the
    path length), then F16.8 for every band, separated by 3X.
    bands: (numb, Nk); flat[i1, i2]: 1-based flattened k index of grid point
    (i1, i2) (SampleBZ's nMomentaFlattened, Geometry.f90:96-127)."""
    numb = bands.shape[0]
    vK1 = (g1 + g12) / 3.0
    vM = (g1 + g12) / 2.0
    aKG = float(np.hypot(*vK1))
    aGM = float(np.hypot(*(vM - g12)))
    aMK = float(np.hypot(*(vK1 - vM)))
    aT = 3 * aKG + aGM + aMK
    n3, n6, n2 = numk // 3, numk // 6, numk // 2  # integer divisions, as in Fortran
    pts = []
    for ivk in range(0, n3):
        pts.append((ivk * aKG / n3 / aT, (2 * numk // 3 + ivk) % numk, (2 * numk // 3 - 2 * ivk) % numk))
    pts.append((aKG / aT, 0, 0))
    for ivk in range(1, n3 + 1):
        pts.append((aKG / aT + aKG * ivk / n3 / aT, ivk % numk, (numk - 2 * ivk) % numk))
    for ivk in range(1, n6 + 1):
        pts.append((2 * aKG / aT + aMK * ivk / n6 / aT, (n3 + ivk) % numk, (n3 + ivk) % numk))
    for ivk in range(1, n2):
        pts.append(((2 * aKG + aMK) / aT + aGM * ivk / n2 / aT, (n2 - ivk) % numk, (n2 + ivk) % numk))
    for ivk in range(0, n3 + 1):
        pts.append(((2 * aKG + aMK + aGM) / aT + ivk * aKG / n3 / aT, ivk % numk, ivk % numk))
    lines = []
    for x, i1, i2 in pts:
        col = int(flat[i1, i2]) - 1
        lines.append("   ".join([fortran_f(x, 16, 8)] +
                               [fortran_f(bands[b, col] - fermi, 16, 8) for b in range(numb)]))
    return lines


def flattened_momenta(nMomentaComponents: np.ndarray, numk: int) -> np.ndarray:
    """nMomentaFlattened of SampleBZ: 1-based flattened index of (i1, i2)."""
    flat = np.zeros((numk, numk), dtype=np.int64)
    for q, (i1, i2) in enumerate(np.asarray(nMomentaComponents, dtype=np.int64)):
        flat[i1, i2] = q + 1
    return flat


def write_bands(outdir: str, parameters: str, bands: np.ndarray, fermi: float, flat: np.ndarray,
                numk: int, g1: np.ndarray, g12: np.ndarray, nspin: int = 1) -> str:
    """output/Bands-nspin<s><parameters> (Main.f90:601-608)."""
    d = os.path.join(outdir, "output")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"Bands-nspin{nspin}" + parameters)
    with open(path, "w") as f:
        for line in plot_bands_lines(bands, fermi, flat, numk, g1, g12):
            f.write(line + "\n")
    return path
