"""Pins the CPU oracle of the order network.

The network has no counterpart in the reference, so the oracle is checked
against what the mathematics fixes: the backpropagated gradient against
central finite differences of the loss (float64), the masked loss against
its definition on a case with a hand-computed value, and the Adam step
against its closed form after one step (|update| = lr for every parameter
with a non-zero gradient).
"""
import numpy as np

from oracle import mlp_oracle as M


def _batch(seed, B=7):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, M.F))
    T = rng.normal(size=(B, M.O))
    mask = (rng.random((B, M.O)) < 0.7).astype(np.float64)
    return X, T, mask


def test_gradient_matches_finite_differences():
    X, T, mask = _batch(1)
    p = M.init(3)
    # biases away from zero so that no ReLU sits at its kink
    d = M.unpack(p)
    d["b1"] += 0.05
    d["b2"] += 0.05
    p = M.pack(d)
    L, g = M.grad(X, T, mask, p)
    assert abs(L - M.loss(X, T, mask, p)) < 1e-14 * max(1.0, L)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(M.NPARAMS, 200, replace=False), [M.NPARAMS - 1, 0]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(X, T, mask, pp) - M.loss(X, T, mask, pm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_masked_loss_definition():
    # zero weights: y = b3; with b3 = 1 and targets 0, every valid output
    # contributes 1, so the loss is exactly 1 whatever the mask
    p = np.zeros(M.NPARAMS)
    d = M.unpack(p)
    d["b3"][:] = 1.0
    p = M.pack(d)
    X, _, mask = _batch(2)
    T = np.zeros((X.shape[0], M.O))
    assert M.loss(X, T, mask, p) == 1.0
    T[mask == 0] = 100.0  # masked outputs do not count
    assert M.loss(X, T, mask, p) == 1.0


def test_adam_first_step_moves_by_lr():
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    g[3] = 0.0
    p1, m, v = M.adam(p, g, np.zeros(50), np.zeros(50), 1e-3, 0.9, 0.999, 1e-300, 1)
    step = np.abs(p1 - p)
    assert np.allclose(np.delete(step, 3), 1e-3, rtol=1e-12)
    assert p1[3] == p[3]
    assert np.all(np.sign(p - p1)[g != 0] == np.sign(g[g != 0]))


def test_order_choice():
    y = np.array([[-3, -5, -6.2, -7, -8], [-1, -2, -3, -4, -5.0]])
    q = M.choose_order(y, -6.0, 0.1, 2, 2, 12)
    assert list(q) == [4, 12]
    assert list(M.choose_order(y, -6.0, 0.0, 2, 5, 12)) == [5, 12]
