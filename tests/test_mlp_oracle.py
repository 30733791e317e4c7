"""CPU checks of the learned-guess pieces: the MLP oracle's backward pass
against central finite differences, the nearest-neighbour bond bookkeeping and
the descriptor/target/guess round trip."""
import numpy as np
import pytest

from oracle import mlp_oracle as O
from forge_amd import learned as Lr
from forge_amd.geometry import build_geometry


def _params(din, dout, rng):
    parts = []
    for fi, fo in ((din, O.H), (O.H, O.H), (O.H, dout)):
        parts += [rng.standard_normal(fi * fo) / np.sqrt(fi), 0.1 * rng.standard_normal(fo)]
    return np.concatenate(parts)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    din, dout, N = 5, 3, 40
    p = _params(din, dout, rng)
    X = rng.standard_normal((N, din))
    T = rng.standard_normal((N, dout))
    M = (rng.random((N, dout)) > 0.2).astype(float)
    loss, g = O.loss_grad(p, X, T, M)
    idx = rng.choice(p.size, 80, replace=False)
    # every block of the packing is probed (first and last entries of each)
    W = [din * O.H, O.H, O.H * O.H, O.H, O.H * dout, dout]
    starts = np.cumsum([0] + W[:-1])
    idx = np.unique(np.r_[idx, starts, starts + np.array(W) - 1])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.loss_grad(pp, X, T, M)[0] - O.loss_grad(pm, X, T, M)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9), i


def test_oracle_loss_is_masked_mean_square():
    rng = np.random.default_rng(4)
    din, dout, N = 4, 2, 10
    p = _params(din, dout, rng)
    X = rng.standard_normal((N, din))
    Y, _ = O.forward(p, X, dout)
    M = np.zeros((N, dout)); M[::2, 0] = 1
    T = Y.copy(); T[::2, 0] += 0.5
    loss, _ = O.loss_grad(p, X, T, M)
    assert loss == pytest.approx(0.25, rel=1e-12)


@pytest.mark.parametrize("m,nl", [(2, 2), (5, 2), (1, 3)])
def test_nearest_neighbour_bonds(m, nl):
    g = build_geometry(m, nl, 3)
    b = Lr.nn_bonds(g)
    n = g.ndim
    assert len(b.i) == 3 * n
    assert np.all(np.bincount(b.i, minlength=n) == 3)
    assert np.all(np.bincount(b.slot) == n)
    # every bond is seen from both ends: a cross-cell entry is stored once
    # plainly and once conjugated, a home-cell entry comes with its transpose
    cross = b.cell > 0
    keys = {}
    for r, c, cc, cj in zip(b.row[cross], b.col[cross], b.cell[cross], b.conj[cross]):
        keys.setdefault((r, c, cc), []).append(bool(cj))
    assert all(sorted(v) == [False, True] for v in keys.values())
    home = set(zip(b.row[~cross], b.col[~cross]))
    assert all((c, r) in home for r, c in home)
    assert not b.conj[~cross].any()
    # bond length a0 in the stored orientation: r_row - (r_col - T_cell)
    C = np.asarray(g.Coords)
    T = Lr.cell_shifts(g)
    d = C[b.row, :2] - (C[b.col, :2] - T[b.cell])
    assert np.allclose(np.hypot(d[:, 0], d[:, 1]), Lr.A0, atol=1e-9)


def test_descriptors_shape_and_registry():
    g = build_geometry(4, 2, 3)
    b = Lr.nn_bonds(g)
    X = Lr.descriptors(g, 1.0, b)
    layer, sub = Lr.site_labels(g)
    assert X.shape == (g.ndim, Lr.D_IN)
    assert np.allclose(X[:, :2].sum(1), 1) and np.allclose(X[:, 2:5].sum(1), 1)
    # bilayer: layer 0 is the bottom (only a layer above), layer 1 the top
    assert np.all(X[layer == 0, 2] == 1) and np.all(X[layer == 1, 4] == 1)
    assert np.all(X[layer == 0, 17:19] == [0, 1]) and np.all(X[layer == 1, 17:19] == [1, 0])
    assert np.all(X[layer == 0, 5:11] == 0) and np.all(X[layer == 1, 11:17] == 0)
    # cos^2 + sin^2 = 1 for every phase present
    for col in (5, 11):
        has = X[:, 17 + (col - 5) // 6] == 1
        c, s = X[has, col:col + 6:2], X[has, col + 1:col + 6:2]
        assert np.allclose(c * c + s * s, 1, atol=1e-6)
    assert np.allclose(X[:, 19], 0.25) and np.allclose(X[:, 20], 2.5)
    # the cell is centred on an AA point: some site sits on top of an A site of
    # the other layer (all three phases 0)
    assert np.max(X[layer == 0, 11:17:2].sum(1)) > 3 - 1e-6


def test_layer_lattice_phases():
    """G_k . (r - o) is a multiple of 2 pi on every A site of the layer itself
    and takes the one value G_k . (d) on its B sites (d = r_B - r_A of a bond),
    for both layers of a bilayer and every layer of a trilayer."""
    for L in (2, 3):
        g = build_geometry(3, L, 3)
        b = Lr.nn_bonds(g)
        layer, sub = Lr.site_labels(g)
        C = np.asarray(g.Coords)[:, :2]
        for l, (G, o) in enumerate(Lr.layer_lattices(g, b)):
            ph = (C[(layer == l) & (sub == 0)] - o) @ G.T
            assert np.allclose(np.cos(ph), 1, atol=1e-9)
            phb = (C[(layer == l) & (sub == 1)] - o) @ G.T
            assert np.allclose(np.cos(phb - phb[0]), 1, atol=1e-9)
            assert not np.allclose(np.cos(phb[0]), 1)
        X = Lr.descriptors(g, 0.0, b)
        if L == 3:
            mid = layer == 1
            assert np.all(X[mid, 3] == 1) and np.all(X[mid, 17:19] == 1)


def test_targets_guess_round_trip():
    """guess_density(rho_tb, targets(rho, rho_tb)) reproduces rho's diagonal and
    nearest-neighbour bonds, and leaves every other entry of rho_tb alone."""
    rng = np.random.default_rng(5)
    g = build_geometry(2, 2, 3)
    n, nnc = g.ndim, g.nnc
    def rand():
        return rng.standard_normal((n, n, nnc)) + 1j * rng.standard_normal((n, n, nnc))
    tb = rand()
    tb[:, :, 0] = tb[:, :, 0] + tb[:, :, 0].conj().T
    d = 0.01 * rand()
    d[:, :, 0] = d[:, :, 0] + d[:, :, 0].conj().T
    rho = tb + d
    b = Lr.nn_bonds(g)
    Y, M = Lr.targets(g, rho, tb, b)
    gr = Lr.guess_density(g, tb, Y.astype(np.float64), b)
    idx = np.arange(n)
    assert np.allclose(np.real(gr[idx, idx, 0]), np.real(rho[idx, idx, 0]), atol=1e-6)
    touched = np.zeros(gr.shape, bool)
    touched[b.row, b.col, b.cell] = True
    home = b.cell == 0
    touched[b.col[home], b.row[home], 0] = True
    assert np.allclose(gr[touched], rho[touched], atol=1e-6)
    off = ~touched
    off[idx, idx, 0] = False
    assert np.array_equal(gr[off], tb[off])


def test_bond_harmonics():
    """b_0 is the shift common to the three bonds, b_1 / b_2 the patterns
    w^-s / w^-2s over the angle slots s (w = exp(2 pi i / 3)); the transform
    inverts exactly."""
    w = np.exp(2j * np.pi / 3)
    s = np.arange(3)
    B = np.stack([np.full(3, 0.3 - 0.2j), 0.5 * w ** (-s), (0.1 + 0.4j) * w ** (-2 * s)])
    H = Lr.to_harmonics(B)
    np.testing.assert_allclose(H[0], [0.3, -0.2, 0, 0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(H[1], [0, 0, 0.5, 0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(H[2], [0, 0, 0, 0, 0.1, 0.4], atol=1e-15)
    rng = np.random.default_rng(3)
    R = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    np.testing.assert_allclose(Lr.from_harmonics(Lr.to_harmonics(R)), R, atol=1e-14)
