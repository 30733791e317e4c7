"""Host side of the learned-guess MLP (kernels in csrc/mlp.hip).

`Mlp` holds the FP32 parameters and the input/output standardisation; the
network itself, its training (forward, backward, Adam) and its inference run
in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L

WIDTH = 128


def _p(a):
    return C.c_void_p(a.ctypes.data)


def param_count(din: int, dout: int) -> int:
    return int(L.load().forge_mlp_param_count(din, dout))


def init_params(din: int, dout: int, seed: int = 1) -> np.ndarray:
    """He-normal weights, zero biases, in the packed layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fi, fo in ((din, WIDTH), (WIDTH, WIDTH), (WIDTH, dout)):
        parts.append(rng.standard_normal(fi * fo) * np.sqrt(2.0 / fi))
        parts.append(np.zeros(fo))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == param_count(din, dout)
    return p


def loss_grad(params, X, T, mask):
    X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, mask))
    params = np.ascontiguousarray(params, np.float32)
    N, din = X.shape
    dout = T.shape[1]
    loss = np.zeros(1, np.float32)
    grad = np.zeros_like(params)
    L.check(L.load().forge_mlp_loss_grad(_p(params), din, dout, _p(X), _p(T), _p(M), N, _p(loss),
                                         _p(grad)), "forge_mlp_loss_grad")
    return float(loss[0]), grad


def train(params, X, T, mask, steps, batch=4096, lr=1e-3, seed=7):
    X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, mask))
    p = np.array(params, np.float32, copy=True)
    N, din = X.shape
    dout = T.shape[1]
    hist = np.zeros(max(steps, 1), np.float32)
    L.check(L.load().forge_mlp_train(_p(p), din, dout, _p(X), _p(T), _p(M), N, steps, batch, lr,
                                     seed, _p(hist)), "forge_mlp_train")
    return p, hist[:steps]


def infer(params, X, dout, bf16=True):
    X = np.ascontiguousarray(X, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    N, din = X.shape
    Y = np.zeros((N, dout), np.float32)
    L.check(L.load().forge_mlp_infer(_p(params), din, dout, _p(X), N, int(bool(bf16)), _p(Y)),
            "forge_mlp_infer")
    return Y


class Mlp:
    """Standardised regression network: the kernels see (X - xm) / xs and
    targets (T - ym) / ys."""

    def __init__(self, params, xm, xs, ym, ys):
        self.params = np.asarray(params, np.float32)
        self.xm, self.xs = np.asarray(xm, np.float32), np.asarray(xs, np.float32)
        self.ym, self.ys = np.asarray(ym, np.float32), np.asarray(ys, np.float32)

    @property
    def din(self):
        return self.xm.size

    @property
    def dout(self):
        return self.ym.size

    @classmethod
    def fit(cls, X, T, mask, steps, batch=4096, lr=1e-3, seed=7):
        xm, xs = X.mean(0), X.std(0)
        xs = np.where(xs > 1e-6, xs, 1.0)
        w = np.maximum(mask.sum(0), 1.0)
        ym = (T * mask).sum(0) / w
        ys = np.sqrt(((T - ym) ** 2 * mask).sum(0) / w)
        ys = np.where(ys > 1e-12, ys, 1.0)
        m = cls(init_params(X.shape[1], T.shape[1], seed), xm, xs, ym, ys)
        Xn = (X - m.xm) / m.xs
        Tn = (T - m.ym) / m.ys
        m.params, hist = train(m.params, Xn, Tn, mask, steps, batch, lr, seed)
        return m, hist

    def predict(self, X, bf16=True):
        Y = infer(self.params, (X - self.xm) / self.xs, self.dout, bf16)
        return Y * self.ys + self.ym

    def save(self, path):
        np.savez(path, params=self.params, xm=self.xm, xs=self.xs, ym=self.ym, ys=self.ys)

    @classmethod
    def load(cls, path):
        f = np.load(path)
        return cls(f["params"], f["xm"], f["xs"], f["ym"], f["ys"])
