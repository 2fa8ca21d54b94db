"""Precision model of the FP32 near field (CPU only).

The FP32 near field (`k_p2p_leaf<float>`) forms each pair's displacement as
the FP64 difference x_t - x_s rounded to FP32 and evaluates the rest of the
pair (r^2, 1/r, the charge's potential and field) in FP32.  The cheaper form
-- FP32 offsets of both particles from the target leaf's centre, subtracted
in FP32 -- carries an absolute rounding error of ~2^-24 times the leaf size
in every displacement, which the closest pairs, whose terms dominate the
field on clustered sets, turn into a large relative error.

This test restates both forms in numpy over the oracle's leaves and P2P lists
(example/tree.cpp:130-233 lists; example/calculate.cpp:24-40 pair body) on
2^16 clustered charges, against the FP64 pair sums, and measures the
normwise error relative to the full field (the oracle's FMM at order 10):
the kernel's form stays near 1e-7, the offset form exceeds the 1e-6
tolerance of the mixed-precision modes.
"""
import numpy as np

from fmmgen_amd import workloads
from oracle import oracle as O


def _near_fields(pos, q, geo, leaf, p2p):
    order = np.argsort(leaf, kind="stable")
    u, first, cnt = np.unique(leaf[order], return_index=True, return_counts=True)
    members = {int(c): order[i:i + k] for c, i, k in zip(u, first, cnt)}
    p2p = p2p[np.argsort(p2p[:, 0], kind="stable")]
    tg, ti, tc = np.unique(p2p[:, 0], return_index=True, return_counts=True)
    n = pos.shape[0]
    out = {k: np.zeros((n, 4)) for k in ("fp64", "kernel", "offsets")}

    def pair_sums(d, qs, selfm, dt):
        dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
        r2 = (dx * dx + dy * dy + dz * dz).astype(dt)
        with np.errstate(divide="ignore"):
            ri = (dt(1) / np.sqrt(r2)).astype(dt)
        ri[selfm] = 0
        qr = (qs * ri).astype(dt)
        qr3 = (qr * ri * ri).astype(dt)
        return np.stack([qr.sum(1, dtype=np.float64), (qr3 * dx).sum(1, dtype=np.float64),
                         (qr3 * dy).sum(1, dtype=np.float64), (qr3 * dz).sum(1, dtype=np.float64)], 1)

    f32 = np.float32
    for c, i, k in zip(tg, ti, tc):
        t = members[int(c)]
        s = np.concatenate([members[int(j)] for j in p2p[i:i + k, 1]])
        selfm = t[:, None] == s[None, :]
        qs = q[s, 0][None, :]
        d = pos[t][:, None, :] - pos[s][None, :, :]
        out["fp64"][t] += pair_sums(d, qs, selfm, np.float64)
        out["kernel"][t] += pair_sums(d.astype(f32), qs.astype(f32), selfm, f32)
        ct = geo[int(c), :3]
        to = (pos[t] - ct).astype(f32)[:, None, :]
        so = (pos[s] - ct).astype(f32)[None, :, :]
        out["offsets"][t] += pair_sums((to - so).astype(f32), qs.astype(f32), selfm, f32)
    return out


def test_fp64_difference_keeps_near_field_below_tolerance():
    n = 1 << 16
    pos = workloads.blobs(n, 3)
    q = np.random.default_rng(5).uniform(-1, 1, (n, 1))
    T = O.OracleTree(pos, q, 64, 10, 0.5)
    geo, _ = T.cells()
    _, p2p = T.lists()
    F = T.compute_field_fmm()
    near = _near_fields(pos, q, geo, T.leaf_of(), p2p)
    scale = np.linalg.norm(F[:, 1:])
    err = {k: np.linalg.norm(near[k][:, 1:] - near["fp64"][:, 1:]) / scale for k in ("kernel", "offsets")}
    assert err["kernel"] < 3e-7, err
    assert err["offsets"] > 2e-6, err
