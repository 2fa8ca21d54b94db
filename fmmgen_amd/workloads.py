"""Synthetic particle sets of the benchmark workloads (numpy, host side).

  uniform   positions U[0,1)^3
  plummer   Plummer sphere, scale radius a, truncated at rmax, isotropic
            (r = a / sqrt(X^(-2/3) - 1), X ~ U(0,1), rejected above rmax)
  blobs     nblob isotropic Gaussian blobs (sigma ~ U[0.01, 0.1], centres
            U[0,1)^3, equal weights) plus a fraction of uniform background
  mixed     half uniform, half Plummer (a = 0.05, truncated at r = 1)
            centred in the unit cube
Charges are U[-1, 1]; dipoles are random unit vectors times lognormal(0, 0.5).
"""
import numpy as np


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    q = rng.uniform(-1.0, 1.0, (n, 1))
    return pos, q


def plummer(n, seed, a=0.1, rmax=1.0):
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


def blobs(n, seed, nblob=16, frac_bg=0.2, sigma=(0.01, 0.1)):
    rng = np.random.default_rng(seed)
    nbg = int(round(frac_bg * n))
    nb = n - nbg
    centres = rng.random((nblob, 3))
    sig = rng.uniform(sigma[0], sigma[1], nblob)
    which = rng.integers(0, nblob, nb)
    pts = centres[which] + rng.normal(size=(nb, 3)) * sig[which, None]
    pos = np.vstack([pts, rng.random((nbg, 3))])
    return pos[rng.permutation(n)]


def mixed(n, seed):
    rng = np.random.default_rng(seed)
    h = n // 2
    pos = np.vstack([rng.random((h, 3)), plummer(n - h, seed, a=0.05) + 0.5])
    return pos[rng.permutation(n)]


def positions(kind, n, seed):
    if kind == "uniform":
        return np.random.default_rng(seed).random((n, 3))
    if kind == "plummer":
        return plummer(n, seed, a=0.1)
    if kind == "blobs":
        return blobs(n, seed)
    if kind == "mixed":
        return mixed(n, seed)
    raise ValueError(f"unknown distribution {kind!r}")


def charges(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, (n, 1))
