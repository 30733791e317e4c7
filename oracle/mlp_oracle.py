"""CPU oracle of the learned-guess MLP (test infrastructure only).

FP64 numpy restatement of the network of forge_amd/csrc/mlp.hip: 3 layers of
width 128 with ReLU, packed parameters [W1, b1, W2, b2, W3, b3] (W row-major
in x out), masked mean-squared loss, hand-derived backward pass.  There is no
reference counterpart; the backward pass is pinned by central finite
differences of the forward loss (tests/test_mlp_oracle.py).
"""
import numpy as np

H = 128


def unpack(p, din, dout):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for fi, fo in ((din, H), (H, H), (H, dout)):
        out.append(p[o:o + fi * fo].reshape(fi, fo)); o += fi * fo
        out.append(p[o:o + fo]); o += fo
    assert o == p.size
    return out


def forward(p, X, dout):
    W1, b1, W2, b2, W3, b3 = unpack(p, X.shape[1], dout)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1 + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2 + b2
    A2 = np.maximum(Z2, 0)
    return A2 @ W3 + b3, (X, Z1, A1, Z2, A2)


def loss_grad(p, X, T, mask):
    dout = T.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(p, X.shape[1], dout)
    Y, (X, Z1, A1, Z2, A2) = forward(p, X, dout)
    M = np.asarray(mask, np.float64)
    d = Y - T
    cnt = M.sum()
    loss = (M * d * d).sum() / cnt
    dY = 2 * M * d / cnt
    gW3 = A2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * (Z2 > 0)
    gW2 = A1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
