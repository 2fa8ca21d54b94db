"""CPU oracle of the order network (test infrastructure only).

A float64 numpy restatement of the network the HIP kernels implement
(fmmgen_amd/csrc/mlp.hip): two hidden ReLU layers, a linear output layer,
the masked mean-square loss (mean over the outputs whose mask is non-zero)
and its gradient by backpropagation, and one bias-corrected Adam step.
The network has no counterpart in the reference, so nothing here follows
a reference file; the gradient is pinned by central finite differences
(tests/test_mlp_oracle.py), which a dropped ReLU mask, a transposed weight
or a missing factor of the loss would fail.

Parameter layout (flat): W1[H][F], b1[H], W2[H][H], b2[H], W3[O][H], b3[O].
"""
import numpy as np

F, H, O = 12, 64, 11
SIZES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)), ("b3", (O,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(params):
    out, o = {}, 0
    for nm, shp in SIZES:
        k = int(np.prod(shp))
        out[nm] = np.asarray(params[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[nm], dtype=np.float64).ravel() for nm, _ in SIZES])


def init(seed=0):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for nm, shp in SIZES:
        if nm.startswith("W"):
            lim = np.sqrt(6.0 / shp[1])
            d[nm] = rng.uniform(-lim, lim, shp)
        else:
            d[nm] = np.zeros(shp)
    return pack(d)


def forward(X, params):
    p = unpack(params)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["W3"].T + p["b3"]
    return y, (X, a1, h1, a2, h2)


def loss(X, T, mask, params):
    y, _ = forward(X, params)
    m = (np.asarray(mask) != 0).astype(np.float64)
    cnt = m.sum()
    return float(((y - T) ** 2 * m).sum() / cnt) if cnt > 0 else 0.0


def grad(X, T, mask, params):
    """(loss, gradient) of the masked mean-square loss."""
    p = unpack(params)
    y, (X, a1, h1, a2, h2) = forward(X, params)
    m = (np.asarray(mask) != 0).astype(np.float64)
    cnt = m.sum()
    inv = 1.0 / cnt if cnt > 0 else 0.0
    e = (y - T) * m
    L = float((e * e).sum() * inv)
    dy = 2.0 * e * inv
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ p["W3"]) * (a2 > 0)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g["W1"], g["b1"] = d1.T @ X, d1.sum(0)
    return L, pack(g)


def adam(params, g, m, v, lr, b1, b2, eps, step):
    """One bias-corrected Adam step; returns (params, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def choose_order(y, log10_tol, margin, order0, min_order, max_order):
    """Lowest order order0 + o with y[o] + margin <= log10_tol, else max_order, clamped."""
    y = np.atleast_2d(y)
    ok = y + margin <= log10_tol
    q = np.where(ok.any(1), ok.argmax(1) + order0, max_order)
    return np.clip(q, min_order, max_order)
