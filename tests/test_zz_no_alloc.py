"""forge_zdiagred_noalloc (no device allocation, scratch preallocated per
GPU) against forge_zdiagred on the same matrices, through the C ABI: two
calls in a row, the second smaller, so stale scratch contents from the first
call would show.  Eigenvalues within 1e-12 of ||A||; eigenvectors compared
column by column through |<z_a, z_b>| (the phase of an eigenvector is free),
within 1e-10 of 1, and the residual ||A Z - Z diag(w)|| within 1e-12 ||A||.
Sizes above the scratch's maximum order are refused."""
import sys

import numpy as np
import pytest

from forge_amd import lib as L
from forge_amd.hf import zdiagred, zdiagred_noalloc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def herm(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray((X + X.conj().T) / 2)


def check(A, il, iu):
    w0, z0, i0 = zdiagred(np.tril(A), il, iu)
    w1, z1, i1 = zdiagred_noalloc(np.tril(A), il, iu)
    nrm = np.linalg.norm(A, 2)
    assert i0 == 0 and i1 == 0
    assert np.abs(w1 - w0).max() < 1e-12 * nrm
    ov = np.abs(np.einsum("ij,ij->j", z0.conj(), z1))
    assert np.abs(ov - 1.0).max() < 1e-10
    assert np.linalg.norm(A @ z1 - z1 * w1) < 1e-12 * nrm * np.sqrt(z1.shape[1])
    return np.abs(w1 - w0).max() / nrm


def test_noalloc_matches_zdiagred_twice_second_smaller():
    say("zdiagred_noalloc n=512 (all eigenpairs), then n=300 (eigenpairs 5..160)")
    e1 = check(herm(512, 1), 1, 512)
    e2 = check(herm(300, 2), 5, 160)
    say(f"  max |w - w_ref| / ||A||: {e1:.1e}, {e2:.1e}")


def test_noalloc_refuses_orders_above_its_scratch():
    with pytest.raises(L.ForgeError, match="maximum order"):
        zdiagred_noalloc(herm(513, 3), 1, 10)
    assert "maximum order" in L.last_error()
