"""GPU tests of the two-stage reduction (band of width 32 by matrix-core
products, bulge chasing, both back-transformations), forced through the
environment variable FORGE_REDUCTION=2 at orders where the one-stage
reduction would be chosen by default.

Eigenvalues are checked against LAPACK (numpy.linalg.eigvalsh), eigenvectors
by the residual |A Z - Z diag(w)| and orthogonality; the SCF loop through the
two-stage reduction against the one-stage reduction on the same inputs.  Sizes
straddle the band width (32), the group width of the chase's reflectors (32)
and their multiples, where the last panel, the last chase steps and the last
group of reflectors are partial.
"""
import sys

import numpy as np
import pytest

from forge_amd.hf import Engine, Setup, zdiagred
from forge_amd.seed import seeded_rho

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_eig(A, w, Z, iu):
    n = A.shape[0]
    wr = np.linalg.eigvalsh(A)[:iu]
    scale = max(1.0, abs(wr).max())
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-13 * scale * max(1.0, np.sqrt(n)) * 10)
    assert np.abs(A @ Z - Z * w).max() < 1e-13 * scale * max(n, 10)
    assert np.abs(Z.conj().T @ Z - np.eye(iu)).max() < 1e-13 * max(n, 10)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 31, 32, 33, 34, 63, 64, 65, 66, 97, 100, 129, 200, 333,
                               700])
def test_twostage_random_hermitian(n, monkeypatch):
    monkeypatch.setenv("FORGE_REDUCTION", "2")
    say("two-stage random", n)
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    iu = min(n, max(1, n // 2 + 3))
    w, Z, info = zdiagred(np.tril(A), 1, iu)
    assert info == 0
    check_eig(A, w, Z, iu)


def _structured(kind, n, rng):
    if kind == "diagonal":
        return np.diag(rng.standard_normal(n)).astype(complex)
    if kind == "zero":
        return np.zeros((n, n), complex)
    if kind == "band":
        # already banded with width 5 (most panel reflectors and chase steps
        # act on zero columns)
        A = np.zeros((n, n), complex)
        for d in range(6):
            v = rng.standard_normal(n - d) + 1j * rng.standard_normal(n - d)
            A += np.diag(v, -d)
        return np.tril(A) + np.tril(A, -1).conj().T + np.diag(np.diag(A).real - np.diag(A))
    if kind == "degenerate":
        # identity plus rank 3: eigenvalue 1 with multiplicity n - 3
        X = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
        return np.eye(n) + X @ X.conj().T
    if kind == "phases":
        # a real symmetric tridiagonal matrix under a random diagonal unitary
        T = np.diag(rng.standard_normal(n)) + np.diag(rng.standard_normal(n - 1), 1)
        T = T + np.triu(T, 1).T
        D = np.exp(1j * rng.uniform(0, 2 * np.pi, n))
        return (D[:, None] * T) * D.conj()[None, :]
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["diagonal", "zero", "band", "degenerate", "phases"])
def test_twostage_structured(kind, monkeypatch):
    monkeypatch.setenv("FORGE_REDUCTION", "2")
    n = 150
    say("two-stage structured", kind, n)
    A = _structured(kind, n, np.random.default_rng(7))
    iu = 90
    w, Z, info = zdiagred(np.tril(A), 1, iu)
    assert info == 0
    check_eig(A, w, Z, iu)


def test_twostage_default_at_large_order():
    """n = 2500 without the override: the two-stage reduction is the default
    from order 1024 on."""
    n = 2500
    say("two-stage by default", n)
    rng = np.random.default_rng(2500)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    iu = n // 2
    w, Z, info = zdiagred(np.tril(A), 1, iu)
    assert info == 0
    check_eig(A, w, Z, iu)


@pytest.mark.parametrize("ntheta,numk,precision", [(2, 3, 0), (4, 3, 0), (4, 3, 1)])
def test_twostage_scf_matches_one_stage(ntheta, numk, precision):
    """A batch of numk^2 k-points through the SCF loop: the two-stage and the
    one-stage reductions give the same trajectory (FP64: to 1e-9 eV;
    mixed precision: the same stated bound as FP64 against mixed)."""
    setup = Setup(ntheta=ntheta, nlayers=2, numk=numk, nfilling=1.0, xi_nm=40.0)
    say("two-stage scf", ntheta, numk, precision)
    trajs = []
    for red in (1, 2):
        eng = Engine(setup)
        start = seeded_rho(eng.n, eng.nnc, setup.nfilling)
        eng.set_options(precision=precision, switch_de=0.0 if precision == 0 else 1e-12,
                        reduction=red)
        eng.scf_init(start)
        trajs.append((eng.scf_run(4), eng.density()))
        eng.close()
    (t1, r1), (t2, r2) = trajs
    tol = 1e-9 if precision == 0 else 1e-6
    np.testing.assert_allclose(t2[:, 1], t1[:, 1], rtol=0, atol=tol)
    assert np.abs(r2 - r1).max() < (1e-10 if precision == 0 else 1e-6)
