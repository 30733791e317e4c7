"""Two spins (numS = 2) in the device-resident loop (forge_scf2_*), against
the trajectories of the reference built with numS = 2 (alone, with C2
enforcement and with the C3-reduced k-set), against the oracle's orc_scf2 on
the final densities, and the mixed precision against FP64.

Tolerances: energies to 1e-8 eV and E_in - E_new to 1e-8 eV (the reference
prints F24.9), the convergence column to 2e-6 and Step to 1e-3 (printed as
F9.6 and F5.3), as in the single-spin parity tests; densities within 1e-9 of
the oracle's.
"""
import sys

import numpy as np
import pytest

from golden_io import load_traj, traj_meta
from forge_amd.hf import Engine, Setup, SpinPair
from forge_amd.lib import ForgeError
from forge_amd.seed import seeded_rho_2spins

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(m, **opts):
    st = dict(ntheta=int(m["ntheta"]), nlayers=int(m["nlayers"]), numk=int(m["numk"]),
              nfilling=float(m["nfilling"]), xi_nm=float(m["xi"].rstrip("nm")))
    e1, e2 = Engine(Setup(**st)), Engine(Setup(**st))
    for e in (e1, e2):
        e.set_options(**opts)
        if m.get("nenforceC3") == "1":
            e.set_c3(True)
        flags = int(m.get("nenforceT", 0)) | 2 * int(m.get("nenforceC2", 0)) | 4 * int(m.get("nenforceC2T", 0))
        if flags:
            e.set_symmetry(flags)
    return e1, e2, SpinPair(e1, e2)


def _start(m, e):
    if m["start"] == "split":
        return seeded_rho_2spins(e.n, e.nnc, float(m["nfilling"]))
    return None


@pytest.mark.parametrize("case", ["t1k3_s2", "t1k3s_s2", "t2k3s_s2", "t1k3s_s2C2", "t1k3s_s2c3"])
def test_two_spin_trajectory_matches_reference(case):
    say("two spins vs reference", case)
    m = traj_meta(case)
    e1, e2, pr = _pair(m)
    pr.init(_start(m, e1))
    traj = pr.run(int(m["itmax"]))
    E, Cv = load_traj(case)
    assert len(traj) == len(E)
    say(f"  max |dE_tot| {np.abs(traj[:, 1] - E[:, 1]).max():.1e}")
    np.testing.assert_allclose(traj[:, 1:6], E[:, 1:6], rtol=0, atol=1e-8)
    np.testing.assert_allclose(traj[:, 6], Cv[:, 1], rtol=0, atol=2e-6)
    np.testing.assert_allclose(traj[:, 7], Cv[:, 2], rtol=0, atol=1e-3)
    np.testing.assert_allclose(traj[:, 8], Cv[:, 3], rtol=0, atol=1e-8)
    pr.close()
    e1.close()
    e2.close()


@pytest.mark.parametrize("case", ["t1k3s_s2", "t2k3s_s2"])
def test_two_spin_densities_match_oracle(case):
    """Both spins' final densities and every trajectory column (E_Fermi and the
    state counts included) against the oracle's two-spin loop."""
    from oracle.pyoracle import System
    say("two spins vs oracle", case)
    m = traj_meta(case)
    e1, e2, pr = _pair(m)
    start = _start(m, e1)
    pr.init(start)
    its = int(m["itmax"])
    traj = pr.run(its)
    rho = pr.density()
    s = System(int(m["ntheta"]), int(m["numk"]), float(m["nfilling"]), nlayers=int(m["nlayers"]))
    n_o, traj_o, rho_o = s.scf2(its, np.stack(start, axis=-1))
    assert n_o == its
    np.testing.assert_allclose(traj[:, 1:10], traj_o[:, 1:10], rtol=0, atol=1e-8)
    np.testing.assert_array_equal(traj[:, 10:12], traj_o[:, 10:12])
    say(f"  max |drho| {np.abs(rho - rho_o).max():.1e}")
    np.testing.assert_allclose(rho, rho_o, rtol=0, atol=1e-9)
    # the spins stay different (a spin-split start)
    assert np.abs(rho[..., 0] - rho[..., 1]).max() > 1e-4
    pr.close()
    e1.close()
    e2.close()


def test_two_spin_mixed_precision_matches_fp64():
    """n = 364 (6.01 deg TBG, 3 x 3 k), a spin-split start: the mixed phase
    against FP64, energies and E_Fermi within 1e-6 eV over the first
    iterations, and the FP64 iterations after the switch."""
    say("two spins mixed vs fp64")
    st = dict(ntheta=5, numk=3, nfilling=1.0, xi_nm=40.0)
    out = {}
    for prec in (0, 1):
        e1, e2 = Engine(Setup(**st)), Engine(Setup(**st))
        for e in (e1, e2):
            e.set_options(precision=prec, switch_de=1e-3)
        pr = SpinPair(e1, e2)
        pr.init(seeded_rho_2spins(e1.n, e1.nnc, 1.0))
        out[prec] = pr.run(6)
        if prec:
            assert 0 < int(e1.timings()[12]) < 6  # both phases ran
        pr.close()
        e1.close()
        e2.close()
    say(f"  max |dE_tot| {np.abs(out[1][:, 1] - out[0][:, 1]).max():.1e}")
    np.testing.assert_allclose(out[1][:, 1:6], out[0][:, 1:6], rtol=0, atol=1e-6)
    np.testing.assert_allclose(out[1][:, 9], out[0][:, 9], rtol=0, atol=1e-6)


def test_two_spin_pair_rejects_mismatched_contexts():
    say("two spins: argument checks")
    a = Engine(Setup(ntheta=1, numk=3, nfilling=1.0, xi_nm=40.0))
    b = Engine(Setup(ntheta=1, numk=3, nfilling=0.0, xi_nm=40.0))
    with pytest.raises(ForgeError, match="same system"):
        SpinPair(a, b)
    c = Engine(Setup(ntheta=1, numk=3, nfilling=1.0, xi_nm=40.0))
    a.set_options()
    c.set_options(occ_mode=1)
    pr = SpinPair(a, c)
    with pytest.raises(ForgeError, match="differ"):
        pr.init(None)
    c.set_options(occ_mode=2, kT=0.001)
    a.set_options(occ_mode=2, kT=0.001)
    with pytest.raises(ForgeError, match="T = 0 rule"):
        pr.init(None)
    with pytest.raises(ForgeError, match="init first"):
        pr.run(1)
    pr.close()
    for e in (a, b, c):
        e.close()
