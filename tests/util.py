"""Shared helpers of the test suite: particle distributions and canonical
tree forms for comparing the GPU tree with the oracle tree."""
import numpy as np


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    q = rng.uniform(-1.0, 1.0, (n, 1))
    return pos, q


def plummer(n, seed, a=0.1, rmax=1.0):
    """Plummer sphere, scale radius a, truncated at rmax, isotropic."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = 2 * (n - out.shape[0]) + 16
        X = rng.random(m)
        X = X[X > 0]
        r = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        r = r[r <= rmax]
        v = rng.normal(size=(r.shape[0], 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        out = np.vstack([out, v * r[:, None]])
    return out[:n]


def dipoles(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * rng.lognormal(0.0, 0.5, size=(n, 1))


def blobs(n, seed, nblob=16, frac_bg=0.2):
    rng = np.random.default_rng(seed)
    nbg = int(round(frac_bg * n))
    nb = n - nbg
    centres = rng.random((nblob, 3))
    sig = rng.uniform(0.01, 0.1, nblob)
    which = rng.integers(0, nblob, nb)
    pts = centres[which] + rng.normal(size=(nb, 3)) * sig[which, None]
    pos = np.vstack([pts, rng.random((nbg, 3))])
    return pos[rng.permutation(n)]


def oracle_canonical(T):
    """Oracle tree keyed by octant path: path -> (x, y, z, r, rmax, nleaf, is_leaf)."""
    geo, meta = T.cells()
    nc = geo.shape[0]
    paths = [None] * nc
    paths[0] = ""
    for c in range(nc):
        for o in range(8):
            ch = meta[c, 4 + o]
            if ch >= 0:
                paths[ch] = paths[c] + str(o)
    cells = {}
    for c in range(nc):
        cells[paths[c]] = (tuple(geo[c]), int(meta[c, 2]), int(meta[c, 3]) == 0)
    return paths, cells


def gpu_canonical(geo, meta):
    """GPU tree (canonical numbering) keyed by octant path."""
    nc = geo.shape[0]
    paths = [None] * nc
    paths[0] = ""
    for c in range(nc):
        k = int(meta[c, 4])
        if k == 0:
            continue
        fc = int(meta[c, 5])
        mask = int(meta[c, 6])
        j = 0
        for o in range(8):
            if mask & (1 << o):
                paths[fc + j] = paths[c] + str(o)
                j += 1
        assert j == k
    cells = {}
    for c in range(nc):
        cells[paths[c]] = (tuple(geo[c]), int(meta[c, 3]), int(meta[c, 4]) == 0)
    return paths, cells


def rel_norm(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
