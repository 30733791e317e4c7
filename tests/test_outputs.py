"""The reference's Energy / Convergence files (Main.f90:352-358, 616-631):
every line the reference program wrote for the fixture cases is reproduced
character for character from its parsed values, and the file name from the
case parameters (CPU)."""
import glob
import os

import numpy as np
import pytest

from forge_amd.checkpoint import run_parameters
from forge_amd.outputs import convergence_line, energy_line, fortran_f, write_outputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sections(path):
    head, energy, conv, cur = [], [], [], None
    for line in open(path):
        line = line.rstrip("\n")
        if line.startswith("# Energy:"):
            cur = energy
        elif line.startswith("# Convergence:"):
            cur = conv
        elif line.startswith("#"):
            head.append(line)
        elif line.strip() and cur is not None:
            cur.append(line)
    return head, energy, conv


CASES = sorted(glob.glob(os.path.join(GOLDEN, "*.traj.txt")))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p) for p in CASES])
def test_lines_match_the_reference(path):
    _, energy, conv = _sections(path)
    assert energy and len(energy) == len(conv)
    for e, c in zip(energy, conv):
        ev = [float(x) for x in e.split()]
        cv = [float(x) for x in c.split()]
        row = np.zeros(12)
        row[0:6] = ev
        row[6], row[7], row[8] = cv[1], cv[2], cv[3]
        assert energy_line(row) == e
        assert convergence_line(row) == c


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p) for p in CASES])
def test_file_name_matches_the_reference(path, tmp_path):
    head, energy, conv = _sections(path)
    meta = dict(kv.split("=") for kv in head[0].split(": ", 1)[1].split())
    fname = [h for h in head if h.startswith("# file: ")][0][len("# file: "):]
    params = run_parameters("SSS", float(meta["nfilling"]), int(meta.get("numS", 1)), int(meta["ntheta"]),
                            int(meta["nlayers"]), 0, 10.0, 4.0, 2,
                            float(meta["xi"].rstrip("nm")), 0.0, 1, 3, int(meta["numk"]))
    assert "Energy" + params == fname
    rows = np.zeros((len(energy), 12))
    for q, (e, c) in enumerate(zip(energy, conv)):
        rows[q, 0:6] = [float(x) for x in e.split()]
        cv = [float(x) for x in c.split()]
        rows[q, 6:9] = cv[1:4]
    pe, pc = write_outputs(str(tmp_path), params, rows)
    assert open(pe).read().splitlines() == energy
    assert open(pc).read().splitlines() == conv


def test_fortran_f_edge_cases():
    assert fortran_f(0.3, 5, 3) == "0.300"
    assert fortran_f(0.3, 4, 3) == ".300"        # leading zero dropped to fit
    assert fortran_f(-0.3, 5, 3) == "-.300"
    assert fortran_f(123.456, 5, 3) == "*****"   # does not fit
    assert fortran_f(1.0, 5, 1) == "  1.0"
    assert fortran_f(-337.235297886, 24, 9) == "          -337.235297886"


def test_bands_file_matches_the_reference():
    """The Bands file the reference program wrote (tests/golden/t1k6.bands.txt:
    28-orbital cell, 6 x 6 k, filling +1, 3 iterations from the reference's
    own start) against plot_bands_lines fed with the oracle's bands and Fermi
    energy of the same run: same lines, same columns, the path coordinates
    character for character, band energies within the F16.8 resolution."""
    from forge_amd.geometry import build_geometry
    from forge_amd.outputs import flattened_momenta, plot_bands_lines
    from oracle.pyoracle import System

    ref = [l.rstrip("\n") for l in open(os.path.join(GOLDEN, "t1k6.bands.txt")) if l.strip()]
    s = System(1, 6, 1.0, nlayers=2, xi_nm=40.0)
    _, traj, _, _ = s.scf(3, rho_start=None, occ_mode=0, mix_mode=0, keep_bands=True)
    g = build_geometry(1, 2, 6, 1)
    lines = plot_bands_lines(s.last_bands, traj[-1, 9], flattened_momenta(g.nMomentaComponents, 6),
                             6, g.g1, g.g12)
    assert len(lines) == len(ref) == 11
    for a, b in zip(lines, ref):
        va, vb = a.split(), b.split()
        assert len(va) == len(vb) == 21
        assert va[0] == vb[0]
        np.testing.assert_allclose([float(x) for x in va[1:]], [float(x) for x in vb[1:]],
                                   rtol=0, atol=1.5e-8)
