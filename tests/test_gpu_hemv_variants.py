"""Every variant of the Householder hemv (segment height and, for the FP32
copy, lane layout) forced in turn through FORGE_HEMV_FORCE, on a cell whose trailing orders cross every band of the
tuner (1.61 deg TBG, n = 5044, 2 x 2 k).

FP64 matrix: each variant's first SCF iteration equals the others' to 1e-10
eV (energies, Fermi level) -- they sum the same products in another order.
FP32 copy: each variant within 1e-6 eV of the FP64 iteration (the mixed
path's bound).  The FP32 variants differ from each other by as much as from
FP64 (measured: up to 2.7e-7 eV against 0.7-2.1e-7 eV): the trailing updates
are rounded to FP32 after FP64 sums taken in another order, so single
elements round the other way, and each variant carries its own FP32
rounding; they are held to the same 1e-6 eV among themselves.  The tuned plan (no variable) is run
last and must match as well.
"""
import os
import sys

import numpy as np
import pytest

from forge_amd.hf import Engine, Setup
from forge_amd.seed import seeded_rho

pytestmark = pytest.mark.gpu

FP64_VARIANTS = ["1024,0", "2048,0"]
FP32_VARIANTS = ["1024,0", "2048,0", "1024,1", "2048,1"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = Engine(Setup(ntheta=20, numk=2, nfilling=1.0, xi_nm=40.0))
    assert e.n == 5044
    yield e
    e.close()
    os.environ.pop("FORGE_HEMV_FORCE", None)


def _one_iteration(e, precision, force):
    if force is None:
        os.environ.pop("FORGE_HEMV_FORCE", None)
    else:
        os.environ["FORGE_HEMV_FORCE"] = force
    e.set_options(occ_mode=2, kT=0.001, precision=precision, switch_de=1e-12, timing=0)
    e.scf_init(seeded_rho(e.n, e.nnc, 1.0))
    r = e.scf_run(1)[0]
    code = int(e.timings()[14])
    os.environ.pop("FORGE_HEMV_FORCE", None)
    return r, code


def test_hemv_variants_agree(eng):
    rows64, rows32 = {}, {}
    for f in FP64_VARIANTS:
        say("fp64 variant", f)
        rows64[f], code = _one_iteration(eng, 0, f)
        assert code == 1111 * (int(f.split(",")[0]) // 1024)
    for f in FP32_VARIANTS:
        say("fp32 variant", f)
        rows32[f], code = _one_iteration(eng, 1, f)
        sg, lay = (int(x) for x in f.split(","))
        assert code == 1111 * (sg // 1024 + 2 * lay)
    tuned64, _ = _one_iteration(eng, 0, None)
    tuned32, code32 = _one_iteration(eng, 1, None)
    say(f"  tuned FP32 plan code {code32}")
    ref = rows64[FP64_VARIANTS[0]]
    cols = [1, 2, 3, 4, 5, 6, 7, 8, 9]  # the input state, max|dn|, Step, E_in - E_new, E_Fermi
    for f, r in list(rows64.items()) + [("tuned", tuned64)]:
        np.testing.assert_allclose(r[cols], ref[cols], rtol=0, atol=1e-10, err_msg=f"fp64 {f}")
    ref32 = rows32[FP32_VARIANTS[0]]
    for f, r in list(rows32.items()) + [("tuned", tuned32)]:
        d64 = np.abs(r[cols] - ref[cols]).max()
        d32 = np.abs(r[cols] - ref32[cols]).max()
        say(f"  fp32 {f}: vs FP64 {d64:.1e} eV, vs FP32 {FP32_VARIANTS[0]} {d32:.1e} eV")
        assert d64 < 1e-6, f
        assert d32 < 1e-6, f
