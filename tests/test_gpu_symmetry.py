"""GPU parity of the symmetry enforcement (nenforceT / nenforceC2 /
nenforceC2T, Main.f90:514-534): forge_symmetrize against the oracle on the
same density, and the device-resident loop with forge_scf_set_symmetry
against the trajectories of the reference built with those settings.

Tolerances: the symmetrisation is a sum and a halving of two elements, so the
same result to 1e-15; SCF energies 1e-8 eV as in the FP64 parity tests.
"""
import sys

import numpy as np
import pytest

from golden_io import load_traj, traj_meta
from forge_amd.hf import Engine, Setup
from forge_amd.seed import seeded_rho

pytestmark = pytest.mark.gpu

SYM_CASES = ["t1k3_T", "t1k3s_C2", "t1k3s_C2T", "t2k3s_TC2C2T"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _flags(m):
    return int(m["nenforceT"]) | 2 * int(m["nenforceC2"]) | 4 * int(m["nenforceC2T"])


def test_symmetrize_matches_oracle():
    from oracle.pyoracle import System
    say("symmetrize t1k3 geometry")
    s = System(1, 3, 1.0)
    e = Engine(Setup(ntheta=1, numk=3, nfilling=1.0, xi_nm=40.0))
    rng = np.random.default_rng(11)
    rho = rng.standard_normal((e.n, e.n, e.nnc)) + 1j * rng.standard_normal((e.n, e.n, e.nnc))
    rho = np.asfortranarray(rho)
    for flags in (1, 2, 4, 7):
        got = e.symmetrize(rho, flags)
        ref = s.symmetrize(rho, flags)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15)
    e.close()


@pytest.mark.parametrize("case", SYM_CASES)
def test_scf_symmetrised_trajectory_matches_reference(case):
    say("scf symmetrised", case)
    m = traj_meta(case)
    e = Engine(Setup(ntheta=int(m["ntheta"]), nlayers=int(m["nlayers"]), numk=int(m["numk"]),
                     nfilling=float(m["nfilling"]), xi_nm=float(m["xi"].rstrip("nm"))))
    e.set_options()
    e.set_symmetry(_flags(m))
    start = seeded_rho(e.n, e.nnc, float(m["nfilling"])) if m["start"] == "seeded" else None
    e.scf_init(start)
    traj = e.scf_run(int(m["itmax"]))
    E, Cv = load_traj(case)
    assert len(traj) == len(E)
    np.testing.assert_allclose(traj[:, 1:6], E[:, 1:6], rtol=0, atol=1e-8)
    np.testing.assert_allclose(traj[:, 6], Cv[:, 1], rtol=0, atol=2e-6)
    np.testing.assert_allclose(traj[:, 7], Cv[:, 2], rtol=0, atol=1e-3)
    np.testing.assert_allclose(traj[:, 8], Cv[:, 3], rtol=0, atol=1e-8)
    e.close()


def test_scf_c3_with_time_reversal_matches_reference():
    """C3-reduced k-set and T enforcement together in the device loop."""
    say("scf c3 + T")
    m = traj_meta("t1k3_c3T")
    e = Engine(Setup(ntheta=1, numk=3, nfilling=1.0, xi_nm=40.0))
    e.set_options()
    e.set_c3(True)
    e.set_symmetry(1)
    e.scf_init(None)
    traj = e.scf_run(int(m["itmax"]))
    E, Cv = load_traj("t1k3_c3T")
    assert len(traj) == len(E)
    np.testing.assert_allclose(traj[:, 1:6], E[:, 1:6], rtol=0, atol=1e-8)
    np.testing.assert_allclose(traj[:, 8], Cv[:, 3], rtol=0, atol=1e-8)
    e.close()
