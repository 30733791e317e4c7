"""GPU checks of the learned-guess MLP kernels against the FP64 CPU oracle.

Tolerances: the training kernels are FP32 (loss within 1e-5 relative;
gradients within 1e-3 relative to the largest gradient entry); the bf16
matrix-core inference rounds operands to 8-bit mantissas, so its outputs are
compared with the FP64 forward within 3 % of the output range and must
correlate above 0.999.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as O
from forge_amd import mlp as MLP

pytestmark = pytest.mark.gpu


def _data(N, din, dout, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, din)).astype(np.float32)
    T = rng.standard_normal((N, dout)).astype(np.float32)
    M = (rng.random((N, dout)) > 0.1).astype(np.float32)
    p = MLP.init_params(din, dout, seed)
    p[:] += 0.01 * rng.standard_normal(p.size).astype(np.float32)  # nonzero biases
    return p, X, T, M


@pytest.mark.parametrize("N,din,dout", [(3000, 18, 7), (777, 5, 16), (64, 32, 1)])
def test_loss_and_gradient_match_oracle(N, din, dout):
    print("mlp grad", N, din, dout, flush=True)
    p, X, T, M = _data(N, din, dout, N)
    loss, g = MLP.loss_grad(p, X, T, M)
    lo, go = O.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                         M.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-5)
    assert np.abs(g - go).max() <= 1e-3 * np.abs(go).max()


def test_inference_fp32_and_bf16_match_oracle():
    print("mlp inference", flush=True)
    p, X, _, _ = _data(5000, 18, 7, 11)
    ref, _ = O.forward(p.astype(np.float64), X.astype(np.float64), 7)
    y32 = MLP.infer(p, X, 7, bf16=False)
    assert np.abs(y32 - ref).max() <= 1e-4 * np.abs(ref).max()
    y16 = MLP.infer(p, X, 7, bf16=True)
    span = ref.max() - ref.min()
    assert np.abs(y16 - ref).max() <= 0.03 * span
    for c in range(7):
        assert np.corrcoef(y16[:, c], ref[:, c])[0, 1] > 0.999


def test_training_fits_a_smooth_function():
    """Adam on mini-batches drives the loss of a smooth 2-output target from
    its initial value to below 5 %, and two runs with one seed agree bitwise."""
    print("mlp training", flush=True)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (4096, 6)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], np.cos(X[:, 2] + X[:, 3])], 1).astype(np.float32)
    M = np.ones_like(T)
    p0 = MLP.init_params(6, 2, 5)
    l0, _ = MLP.loss_grad(p0, X, T, M)
    p1, hist = MLP.train(p0, X, T, M, steps=600, batch=512, lr=2e-3, seed=9)
    l1, _ = MLP.loss_grad(p1, X, T, M)
    assert l1 < 0.05 * l0
    assert hist.shape == (600,) and hist[-50:].mean() < hist[:50].mean()
    p2, _ = MLP.train(p0, X, T, M, steps=600, batch=512, lr=2e-3, seed=9)
    assert np.array_equal(p1, p2)
