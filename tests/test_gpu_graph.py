"""HIP-graph replay of the Householder reduction (forge_scf_options.graph):
the captured launch sequence is replayed every SCF iteration, with the
iteration's own eager work (assembly, eigensolver, densities, host-side Fermi
level and mixing, synchronisations) between replays.  Every replayed
iteration must reproduce the eager run exactly (same kernels, same arguments,
same order), also across a re-capture (the switch from the FP32 copy to FP64
in the mixed mode changes the captured shape) and in a sweep whose batch
shrinks as systems converge.
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


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(precision):
    say("graph replay, precision", precision)
    e = Engine(Setup(ntheta=5, numk=3, nfilling=1.0, xi_nm=40.0))  # n = 364
    sw = dict(precision=precision, switch_de=1e-3) if precision else {}
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
