"""HIP graphs (forge_scf_options.graph).

graph = 1: the Householder reduction's captured launch sequence is replayed
every SCF iteration, with the iteration's own eager work (assembly,
eigensolver, densities, host-side Fermi level and mixing, synchronisations)
between replays.  Every replayed iteration must reproduce the eager run
exactly (same kernels, same arguments, same order), also across a re-capture
(the switch from the FP32 copy to FP64 in the mixed mode changes the captured
shape) and in a sweep whose batch shrinks as systems converge.

graph = 2: the whole iteration on the device -- the sort, Fermi level,
occupations, OptimalStep or linear mixing and the convergence measure as
kernels -- against the host loop (graph = 0), on the cases and modes where the
host loop matches the oracle: the eigensolver and density kernels are the
same, the device sums the occupations in another order and sorts tied
energies by their flat index, so trajectories agree to 1e-8 eV and densities
to 1e-9 (the tolerances of the host loop against the oracle).  Through the
iteration graph the reference's own runs are reproduced to their printed
precision.
graph = 3: that iteration captured into one HIP graph per phase and replayed:
bit for bit the eager device iteration, over several replays, across the
mixed-to-FP64 switch, and for replays of a graph kept from an earlier solve.
"""
import sys

import numpy as np
import pytest

from forge_amd.hf import Engine, Setup, Sweep
from forge_amd.seed import seeded_rho

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _traj(e, nfill, iters, **opt):
    e.set_options(occ_mode=2, mix_mode=0, kT=0.001, **opt)
    e.scf_init(seeded_rho(e.n, e.nnc, nfill))
    return e.scf_run(iters)


@pytest.mark.parametrize("reduction", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(precision, reduction):
    """reduction 2: the two-stage reduction's launches (panel QR, products,
    bulge chasing) captured instead of the one-stage ones."""
    say("graph replay, precision", precision, "reduction", reduction)
    e = Engine(Setup(ntheta=5, numk=3, nfilling=1.0, xi_nm=40.0))  # n = 364
    sw = dict(precision=precision, switch_de=1e-3) if precision else {}
    sw["reduction"] = reduction
    eager = _traj(e, 1.0, 12, **sw)
    graph = _traj(e, 1.0, 12, graph=1, **sw)
    again = _traj(e, 1.0, 12, graph=1, **sw)  # replays of the graph kept from the last run
    if precision:
        assert 0 < int(e.timings()[12]) < 12  # both shapes were replayed
    np.testing.assert_array_equal(graph, eager)
    np.testing.assert_array_equal(again, eager)
    e.close()


def test_graph_sweep_matches_eager():
    say("graph replay in a sweep")
    systems = [(2, 1.0), (4, -1.0), (3, 2.0)]
    engs = [Engine(Setup(ntheta=t, numk=3, nfilling=f, xi_nm=40.0)) for t, f in systems]

    def run(graph):
        for e, (_, f) in zip(engs, systems):
            e.set_options(occ_mode=2, mix_mode=0, kT=0.001, etol=1e-8, graph=graph)
            e.scf_init(seeded_rho(e.n, e.nnc, f))
        sw = Sweep(engs)
        out = sw.run(40)
        sw.close()
        return out

    eager, graph = run(0), run(1)
    for a, b in zip(eager, graph):
        np.testing.assert_array_equal(b, a)
    for e in engs:
        e.close()


def _dev_run(e, iters, graph, opts, init=True, start="seeded", nfill=1.0):
    if init:
        e.set_options(**opts, graph=graph)
    e.scf_init(seeded_rho(e.n, e.nnc, nfill) if start == "seeded" else None)
    return e.scf_run(iters), e.density()


# (case, options): the cases and modes the host loop matches the oracle on
# (test_gpu_parity), plus the mixed precision
DEV_CASES = [("t2k3", dict(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3)),
             ("t2k3", dict(occ_mode=1, mix_mode=0)),
             ("t3k6", dict(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3)),
             ("t3k6", dict(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3, precision=1, switch_de=1e-5))]


@pytest.mark.parametrize("case,opts", DEV_CASES)
def test_device_iteration_matches_host_loop(case, opts):
    from golden_io import case_params, load_case
    say("device iteration vs host loop", case, opts)
    c = case_params(load_case(case))
    e = Engine(Setup(ntheta=c["ntheta"], nlayers=c["nlayers"], numk=c["numk"], nfilling=c["nfilling"],
                     xi_nm=c["xi_nm"]))
    host, rho_h = _dev_run(e, 6, 0, opts, nfill=c["nfilling"])
    dev, rho_d = _dev_run(e, 6, 2, opts, nfill=c["nfilling"])
    say(f"  max |dE_tot| {np.abs(dev[:, 1] - host[:, 1]).max():.1e}, max |drho| {np.abs(rho_d - rho_h).max():.1e}")
    # In the mixed mode the first iterations reduce an FP32 copy: densities
    # that differ in the last FP64 bit (the occupation sums' order) round to
    # different FP32 matrices, so both runs carry independent FP32 rounding
    # (unit roundoff 6e-8) until the switch to FP64.
    tol_e, tol_rho = (1e-7, 1e-7) if opts.get("precision") else (1e-8, 1e-9)
    np.testing.assert_allclose(dev[:, 1:10], host[:, 1:10], rtol=0, atol=tol_e)
    np.testing.assert_array_equal(dev[:, 10:12], host[:, 10:12])
    np.testing.assert_allclose(rho_d, rho_h, rtol=0, atol=tol_rho)
    e.close()


@pytest.mark.parametrize("case", ["t2k3", "t3k6"])
def test_device_iteration_matches_reference_trajectory(case):
    """The reference's own run (T = 0 rule with its zdeg weight, optimal
    damping, first-diagonalisation start) through the iteration graph."""
    from golden_io import case_params, load_case, load_traj
    say("iteration graph vs reference trajectory", case)
    c = case_params(load_case(case))
    e = Engine(Setup(ntheta=c["ntheta"], nlayers=c["nlayers"], numk=c["numk"], nfilling=c["nfilling"],
                     xi_nm=c["xi_nm"]))
    traj, _ = _dev_run(e, c["itmax"], 3, dict(), start="first")
    E, Cv = load_traj(case)
    assert len(traj) == len(E)
    np.testing.assert_allclose(traj[:, 1:6], E[:, 1:6], rtol=0, atol=1e-8)
    np.testing.assert_allclose(traj[:, 8], Cv[:, 3], rtol=0, atol=1e-8)
    e.close()


@pytest.mark.parametrize("reduction", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_iteration_graph_replays_the_device_iteration_bitwise(precision, reduction):
    say("iteration graph vs eager device iteration, precision", precision, "reduction", reduction)
    e = Engine(Setup(ntheta=5, numk=3, nfilling=1.0, xi_nm=40.0))  # n = 364
    opts = dict(occ_mode=2, mix_mode=0, kT=0.001, reduction=reduction)
    if precision:
        opts.update(precision=1, switch_de=1e-3)
    eager, rho_e = _dev_run(e, 12, 2, opts)
    graph, rho_g = _dev_run(e, 12, 3, opts)
    if precision:
        assert 0 < int(e.timings()[12]) < 12  # both phases ran
    again, rho_a = _dev_run(e, 12, 3, opts, init=False)  # replays of the kept graphs
    np.testing.assert_array_equal(graph, eager)
    np.testing.assert_array_equal(rho_g, rho_e)
    np.testing.assert_array_equal(again, eager)
    np.testing.assert_array_equal(rho_a, rho_e)
    e.close()


def test_device_iteration_refuses_the_c3_set():
    from forge_amd import lib as L
    e = Engine(Setup(ntheta=1, numk=3, nfilling=1.0, xi_nm=40.0))
    e.set_options(graph=3)
    e.set_c3(True)
    e.scf_init(None)
    with pytest.raises(L.ForgeError, match="plain loop"):
        e.scf_run(2)
    e.close()
