"""The matrix-core GEMMs (gemm.hip) through the C ABI's products of the
reference's LAPACK shim (LapackRoutines.f90:119-236), against numpy on the
same inputs.

Shapes cover single rows and columns, sizes that are not multiples of the
64- and 128-wide tiles, both tile variants (small outputs take the 64 x 64
kernel) and inner dimensions that are not multiples of the 16-deep stage.
Tolerances: FP64 products within 1e-13 of the largest |entry| of |A| |B|
(the sum of k rounded terms); FP32 within 2e-6 of it.  The fused density
product (herk_cells) and the real FP64 products of the divide-and-conquer
merges are covered by the Solve and zdiagRed parity tests.
"""
import sys

import numpy as np
import pytest

from forge_amd.hf import c_matrix_mul, z_matrix_mul, z_matrix_overlap, z_vector_mul

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape, dtype=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def bound(A, B):
    return np.abs(A) @ np.abs(B)


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (7, 3, 5), (64, 16, 64), (65, 17, 63), (200, 33, 129),
                                   (300, 257, 70), (517, 130, 1030), (1, 500, 300)])
def test_z_matrix_mul(m, k, n):
    say("z_matrix_mul", m, k, n)
    rng = np.random.default_rng(m * 7 + k * 3 + n)
    A, B = crand(rng, m, k), crand(rng, k, n)
    got = z_matrix_mul(A, B)
    err = np.abs(got - A @ B).max() / bound(A, B).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("m1,m0", [(1, 1), (40, 9), (1000, 70), (333, 257)])
def test_z_matrix_overlap(m1, m0):
    say("z_matrix_overlap", m1, m0)
    rng = np.random.default_rng(m1 + m0)
    A, B = crand(rng, m1, m0), crand(rng, m1, m0)
    z = 0.3 - 1.7j
    got = z_matrix_overlap(A, B, z)
    ref = z * (A.conj().T @ B)
    assert np.abs(got - ref).max() / (abs(z) * bound(A.conj().T, B).max()) < 1e-13
    np.testing.assert_allclose(z_matrix_overlap(A, B), A.conj().T @ B, atol=1e-12 * bound(A.T, B).max())


@pytest.mark.parametrize("n,k", [(1, 1), (50, 1), (300, 1), (129, 40), (700, 300)])
def test_z_vector_mul(n, k):
    """k = 1 is the reference's rank-1 update; k > 1 the same for k columns."""
    say("z_vector_mul", n, k)
    rng = np.random.default_rng(n + k)
    A, C = crand(rng, n, k), crand(rng, n, n)
    z = -0.25 + 0.5j
    got = z_vector_mul(A, C, z)
    ref = C + z * (A @ A.conj().T)
    scale = abs(z) * bound(A, A.conj().T).max()
    assert np.abs(got - ref).max() / scale < 1e-13
    # rank-k equals k successive rank-1 updates
    step = C.copy()
    for c in range(min(k, 3)):
        step = z_vector_mul(A[:, c], step, z)
    np.testing.assert_allclose(step, C + z * (A[:, :3] @ A[:, :3].conj().T) if k >= 3
                               else ref, atol=1e-12 * scale)


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (65, 17, 63), (300, 257, 700), (1030, 64, 513)])
def test_c_matrix_mul_fp32(m, k, n):
    say("c_matrix_mul", m, k, n)
    rng = np.random.default_rng(m + 2 * k + 3 * n)
    A, B = crand(rng, m, k, dtype=np.complex64), crand(rng, k, n, dtype=np.complex64)
    got = c_matrix_mul(A, B)
    ref = A.astype(np.complex128) @ B.astype(np.complex128)
    err = np.abs(got - ref).max() / bound(A, B).max()
    assert err < 2e-6, err
