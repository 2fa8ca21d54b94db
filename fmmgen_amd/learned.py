"""Learned per-leaf expansion order: training data, training, model file.

The order network (three layers, 64 hidden ReLU units, FP32 weights, FP16
inference; kernels in csrc/mlp.hip) maps twelve features of a leaf cell
(csrc/learned.h) to the predicted far-field error of the leaf at every
order 2..12, as log10(rms |dF| over the leaf's particles / F_s), F_s the
field scale of the judge's sample of targets (Tree.field_scale).  The
learned evaluation (mode "learned") picks the leaf orders with the fewest
M2L multiply-adds whose predicted norm-wise error (predictions raised by a
calibrated safety margin) meets the tolerance, runs every M2L tile
at the largest order of its target cells, and checks a sample of targets
against an FP64 direct sum, rerunning leaves that fail at the full order.

Training data are made by the plain FP64 FMM itself: one tree per
particle set (uniform, Plummer, Gaussian blobs, mixed; N = 2^16 .. 2^22;
ncrit 32 .. 128; opening angle 0.2 .. 0.6; random-sign strengths),
evaluated at every order against the FP64 direct sum of all targets;
each leaf gives one sample (its features, its 11 errors).  The safety
margin is the smallest for which every held-out set that the highest
order can meet meets the tolerance at the allocated orders.  The
network is trained on the GPU by the library's own forward, backward and
Adam kernels (fmm_mlp_grad / fmm_mlp_adam).  Feature normalisation is
folded into the first layer, so the saved parameters act on raw features.

    python -m fmmgen_amd.learned --out model.npz [--samples 200000]
"""
import argparse
import ctypes
import os
import time

import numpy as np
import torch

from . import _lib, workloads
from .tree import build_tree, _ptr, _stream

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "order_mlp.npz")
F, H, O, ORDER0 = _lib.MLP_FEATURES, _lib.MLP_HIDDEN, _lib.MLP_OUTPUTS, _lib.MLP_ORDER0
ORDERS = list(range(ORDER0, ORDER0 + O))
FLOOR = -16.0  # log10 of the smallest relative error kept (FP64 round-off)


# ------------------------------------------------------------------ data

def error_samples(pos, S, ncrit=64, theta=0.5, pmax=None, max_leaves=8000, seed=0):
    """(X [m, 12], Y [m, 11]) float32 device tensors of one particle set.

    Y[:, o] = log10(rms over the leaf's particles of |F_q - F_direct| /
    F_s) for q = ORDERS[o] (field components), from the plain FP64 FMM
    at order q and the FP64 direct sum; F_s is the field scale of the
    judge sample (Tree.field_scale), the scale the features use too, known
    before the far field runs.  Sets with more than ``max_leaves`` leaves
    contribute a
    random sample of that many leaves (the direct sum runs at their
    particles only).
    """
    from .tree import direct_sum
    pmax = pmax or ORDERS[-1]
    n = pos.shape[0]
    dpos = torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
    dS = torch.from_numpy(np.ascontiguousarray(S, dtype=np.float64)).cuda()
    tree = build_tree(dpos, dS, n, ncrit, pmax, theta)
    leaves = tree.leaves()
    nl = len(leaves)
    F = torch.empty(n, 4, dtype=torch.float64, device="cuda")
    if nl <= max_leaves:
        rows = np.arange(nl)
        Fd = tree.compute_field_exact(torch.empty(n, 4, dtype=torch.float64, device="cuda"))
        tg = None
    else:
        rows = np.sort(np.random.default_rng(seed).choice(nl, max_leaves, replace=False))
        _, leaf_of = tree.export_particles()
        pick = np.zeros(int(leaves.max()) + 1, dtype=bool)
        pick[leaves[rows]] = True
        tg = torch.from_numpy(np.nonzero(pick[leaf_of])[0].astype(np.int64)).cuda()
        Fsub = direct_sum(dpos, dS, tg)
    tree.compute_field_fmm(F, mode="plain", phases=_lib.PHASE_P2P)  # gathers the sources
    X = tree.leaf_features()[torch.from_numpy(rows).cuda()]
    fscale = tree.field_scale()
    Y = []
    for q in ORDERS:
        tree.set_order(q)
        tree.compute_field_fmm(F, mode="plain")
        if tg is not None:  # reference: the FMM itself, direct sums at the sampled leaves
            Fd = F.clone()
            Fd[tg] = Fsub
        e = tree.leaf_error(F, Fd)[torch.from_numpy(rows).cuda()] / fscale
        Y.append(torch.log10(torch.clamp(e, min=10.0 ** FLOOR)))
    return X, torch.stack(Y, 1).float()


THETAS = (0.2, 0.3, 0.4, 0.45, 0.5, 0.5, 0.5, 0.55, 0.6)


def problem_set(total, seed=0, nmin=1 << 16, nmax=1 << 22):
    """Particle sets of mixed kinds, sizes, leaf capacities (ncrit), opening
    angles (the reference's example sweeps theta over 0.1 .. 0.5,
    example/run.sh) and source-strength laws with random signs: charges
    uniform in [-1, 1] (85%, the example's law, example/main.cpp:86-87)
    and lognormal magnitudes (15%)."""
    rng = np.random.default_rng(seed)
    kinds = ["blobs", "blobs", "uniform", "plummer", "mixed", "blobs_tight", "blobs_wide"]
    k = 0
    while True:
        kind = kinds[k % len(kinds)]
        n = int(2 ** rng.uniform(np.log2(nmin), np.log2(nmax)))
        s = int(rng.integers(1 << 30))
        if kind == "blobs_tight":
            pos = workloads.blobs(n, s, nblob=int(rng.integers(4, 32)), sigma=(0.003, 0.03))
        elif kind == "blobs_wide":
            pos = workloads.blobs(n, s, nblob=int(rng.integers(4, 32)), sigma=(0.03, 0.2),
                                  frac_bg=float(rng.uniform(0.0, 0.5)))
        else:
            pos = workloads.positions(kind, n, s)
        ncrit = int(rng.choice([32, 64, 64, 128]))
        theta = float(rng.choice(THETAS))
        u = rng.random()
        crng = np.random.default_rng(s + 1)
        if u < 0.85:
            law, S = "pm", crng.uniform(-1.0, 1.0, (n, 1))
        else:
            law, S = "lognorm", crng.lognormal(0.0, 1.0, (n, 1)) * crng.choice([-1.0, 1.0], (n, 1))
        yield kind, pos, S, ncrit, theta, law
        k += 1


def make_dataset(total, seed=0, log=print, **kw):
    """(X, Y, problem id per sample, per-problem description) over particle
    sets until ``total`` samples."""
    Xs, Ys, Ps, meta, got = [], [], [], [], 0
    t0 = time.time()
    for k, (kind, pos, S, ncrit, theta, law) in enumerate(problem_set(total, seed, **kw)):
        X, Y = error_samples(pos, S, ncrit=ncrit, theta=theta, seed=seed + k)
        meta.append(f"{kind}/{law} n={pos.shape[0]} ncrit={ncrit} theta={theta:.2f}")
        Xs.append(X)
        Ys.append(Y)
        Ps.append(torch.full((X.shape[0],), k, dtype=torch.int32, device=X.device))
        got += X.shape[0]
        log(f"  data: {kind:12s} {law:7s} n={pos.shape[0]:7d} ncrit={ncrit:3d} theta={theta:.2f} leaves={X.shape[0]:6d} "
            f"total={got:7d} "
            f"({time.time() - t0:.1f}s)")
        if got >= total:
            break
    return torch.cat(Xs), torch.cat(Ys), torch.cat(Ps), meta


# ------------------------------------------------------------------ network

def init_params(seed=0):
    """He-uniform weights, zero biases (flat FP32 layout of the C ABI)."""
    g = torch.Generator().manual_seed(seed)
    parts = []
    for nin, nout in ((F, H), (H, H), (H, O)):
        lim = (6.0 / nin) ** 0.5
        parts.append((torch.rand(nout, nin, generator=g) * 2 - 1) * lim)
        parts.append(torch.zeros(nout))
    return torch.cat([p.reshape(-1) for p in parts]).float().cuda()


def fold_normalisation(params, mu, sigma):
    """Parameters acting on raw features x of a network trained on (x - mu) / sigma."""
    p = params.clone()
    W1 = p[:H * F].view(H, F)
    b1 = p[H * F:H * F + H]
    W1n = W1 / sigma[None, :]
    b1n = b1 - W1n @ mu
    p[:H * F] = W1n.reshape(-1)
    p[H * F:H * F + H] = b1n
    return p


def grad(X, T, mask, params):
    """(loss, gradient) on the GPU (fmm_mlp_grad)."""
    B = X.shape[0]
    work = torch.empty(int(_lib.load().fmm_mlp_work_floats(B)), dtype=torch.float32, device="cuda")
    g = torch.empty(_lib.MLP_PARAMS, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().fmm_mlp_grad(_ptr(X), _ptr(T), _ptr(mask), B, _ptr(params), _ptr(g), _ptr(loss),
                                        _ptr(work), _stream()))
    return loss, g


def forward(X, params):
    Y = torch.empty(X.shape[0], O, dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().fmm_mlp_forward(_ptr(X), X.shape[0], _ptr(params), _ptr(Y), _stream()))
    return Y


def infer_f16(X, params, tol=1e-6, margin=0.0, min_order=ORDER0, max_order=ORDERS[-1]):
    """(Y [B, 11], order [B]) from the FP16 inference kernel."""
    B = X.shape[0]
    Y = torch.empty(B, O, dtype=torch.float32, device="cuda")
    q = torch.empty(B, dtype=torch.int32, device="cuda")
    _lib.check(_lib.load().fmm_mlp_infer_f16(_ptr(X), B, _ptr(params), _ptr(Y), _ptr(q), float(tol),
                                             float(margin), int(min_order), int(max_order), _stream()))
    return Y, q


def train(X, Y, epochs=40, batch=2048, lr=3e-3, seed=0, log=print):
    """Adam on the masked mean-square loss, on the GPU; returns raw-feature parameters."""
    X = X.float().contiguous()
    Y = Y.float().contiguous()
    mu = X.mean(0)
    sigma = X.std(0) + 1e-3
    Xn = ((X - mu) / sigma).contiguous()
    mask = torch.ones_like(Y)
    params = init_params(seed)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    B = min(batch, X.shape[0])
    work = torch.empty(int(_lib.load().fmm_mlp_work_floats(B)), dtype=torch.float32, device="cuda")
    g = torch.empty_like(params)
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    L = _lib.load()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    nb = X.shape[0] // B
    step = 0
    total = epochs * nb
    for ep in range(epochs):
        perm = torch.randperm(X.shape[0], device="cuda", generator=gen)
        acc = 0.0
        for b in range(nb):
            idx = perm[b * B:(b + 1) * B]
            xb, yb, mb = Xn[idx].contiguous(), Y[idx].contiguous(), mask[idx].contiguous()
            step += 1
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * step / total))
            _lib.check(L.fmm_mlp_grad(_ptr(xb), _ptr(yb), _ptr(mb), B, _ptr(params), _ptr(g), _ptr(loss),
                                      _ptr(work), _stream()))
            _lib.check(L.fmm_mlp_adam(_ptr(g), _ptr(params), _ptr(m), _ptr(v), ctypes.c_float(lr_t),
                                      ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), step,
                                      _stream()))
            if b % 16 == 0:
                acc += loss.item()
        if ep % 5 == 0 or ep == epochs - 1:
            log(f"  epoch {ep:3d} loss {acc / max(1, (nb + 15) // 16):.4f}")
    return fold_normalisation(params, mu, sigma)


def allocate(Yp, counts, w, tol, margin, qmin=ORDER0, qmax=ORDERS[-1]):
    """Column (order - ORDER0) per leaf chosen as the device does
    (csrc/learned.hip, allocate_orders): the cheapest orders, by
    C(q + 6, 6) w_l + lam n_l e_l(q)^2, whose predicted norm-wise error
    meets tol; log10 lam is the first point meeting it on a grid of 64 over
    [-10, 40], refined by a grid of 64 inside the bracket."""
    from math import comb
    cols = np.array([o for o, q in enumerate(ORDERS) if qmin <= q <= qmax])
    macs = np.array([comb(ORDERS[o] + 6, 6) for o in cols], dtype=np.float64)
    e2 = 10.0 ** (2.0 * (Yp[:, cols] + margin))
    target = tol * tol * counts.sum()
    K = 64

    def pick(loglam):
        c = macs[None, :] * w[:, None] + 10.0 ** loglam * counts[:, None] * e2
        o = c.argmin(1)
        return o, (counts * e2[np.arange(len(o)), o]).sum()

    lo, step = -10.0, 50.0 / (K - 1)
    errs = [pick(lo + k * step)[1] for k in range(K)]
    k = next((i for i, e in enumerate(errs) if e <= target), K)
    if k == K:
        loglam = lo + (K - 1) * step
    elif k == 0:
        loglam = lo
    else:
        flo, fstep = lo + (k - 1) * step, step / (K - 1)
        errs = [pick(flo + j * fstep)[1] for j in range(K)]
        j = next((i for i, e in enumerate(errs) if e <= target), K - 1)
        loglam = flo + j * fstep
    return cols[pick(loglam)[0]]


def problem_errors(Y_true, Y_pred, X, pid, tol, margin):
    """True norm-wise relative field error of every particle set when its
    leaves run at the allocated orders: sqrt(sum_l n_l e_l^2 / sum_l n_l),
    e_l the leaf's true rms error relative to the set's rms field."""
    counts = np.exp2(X[:, 6].astype(np.float64))
    w = np.maximum(10.0 ** X[:, 5].astype(np.float64) - 1.0, 1.0)
    out = []
    for k in np.unique(pid):
        s = pid == k
        o = allocate(Y_pred[s], counts[s], w[s], tol, margin)
        e = 10.0 ** Y_true[s][np.arange(s.sum()), o]
        out.append(np.sqrt((counts[s] * e * e).sum() / counts[s].sum()))
    return np.array(out)


def full_order_errors(Y_true, X, pid):
    """Norm-wise error of every particle set with all leaves at the highest
    order (the fixed-order fallback's error)."""
    counts = np.exp2(X[:, 6].astype(np.float64))
    out = []
    for k in np.unique(pid):
        s = pid == k
        e = 10.0 ** Y_true[s][:, -1]
        out.append(np.sqrt((counts[s] * e * e).sum() / counts[s].sum()))
    return np.array(out)


def calibrate_margin(Y_true, Y_pred, X, pid, tol, target=1.0):
    """Smallest margin (log10 units, step 0.05) for which every held-out
    particle set meets target * tol norm-wise at the allocated orders --
    every set that the highest order meets: the others no order choice
    can save (the judge reruns them at the full order)."""
    yt = Y_true.cpu().numpy().astype(np.float64)
    yp = Y_pred.cpu().numpy().astype(np.float64)
    x = X.cpu().numpy()
    p = pid.cpu().numpy()
    feasible = full_order_errors(yt, x, p) <= target * tol
    for margin in np.arange(-0.5, 3.0, 0.05):
        if np.all(problem_errors(yt, yp, x, p, tol, margin)[feasible] <= target * tol):
            return float(margin)
    return 3.0


def save(path, params, margin, tol, info):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=params.detach().cpu().numpy().astype(np.float32), margin=np.float64(margin),
             tol=np.float64(tol), info=np.array(info))


def load(path=MODEL_PATH):
    """(params float32 device tensor, margin, tol)."""
    with np.load(path, allow_pickle=False) as z:
        p = torch.from_numpy(z["params"].astype(np.float32)).cuda()
        return p, float(z["margin"]), float(z["tol"])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--samples", type=int, default=200_000)
    ap.add_argument("--holdout", type=int, default=40_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    t0 = time.time()
    print("training data", flush=True)
    X, Y, _, _ = make_dataset(a.samples, seed=a.seed, log=lambda s: print(s, flush=True))
    print("held-out data", flush=True)
    Xh, Yh, ph, meta_h = make_dataset(a.holdout, seed=a.seed + 1000, log=lambda s: print(s, flush=True))
    print(f"training on {X.shape[0]} samples", flush=True)
    params = train(X, Y, epochs=a.epochs, seed=a.seed, log=lambda s: print(s, flush=True))
    Yp, _ = infer_f16(Xh, params)
    rmse = torch.sqrt(((Yp - Yh) ** 2).mean()).item()
    margin = calibrate_margin(Yh, Yp, Xh, ph, a.tol)
    errs = problem_errors(Yh.cpu().numpy().astype(np.float64), Yp.cpu().numpy().astype(np.float64),
                          Xh.cpu().numpy(), ph.cpu().numpy(), a.tol, margin)
    full = full_order_errors(Yh.cpu().numpy().astype(np.float64), Xh.cpu().numpy(), ph.cpu().numpy())
    print(f"held-out rmse {rmse:.3f} (log10 units), margin {margin:.2f}, held-out set errors "
          f"max {errs.max():.2e} median {np.median(errs):.2e}", flush=True)
    for d, e, f in zip(meta_h, errs, full):
        print(f"  held-out {d}: {e:.2e} (highest order: {f:.2e})", flush=True)
    per = {}
    for d, k in zip(meta_h, np.unique(ph.cpu().numpy())):
        s = (ph == int(k))
        per[d] = torch.sqrt(((Yp[s] - Yh[s]) ** 2).mean()).item()
    print("  held-out rmse per set: " + ", ".join(f"{d.split(' ')[0]} {v:.3f}" for d, v in per.items()),
          flush=True)
    save(a.out, params, margin, a.tol,
         f"samples={X.shape[0]} holdout={Xh.shape[0]} rmse={rmse:.4f} epochs={a.epochs} seed={a.seed}")
    print(f"saved {a.out} ({time.time() - t0:.1f}s)", flush=True)


if __name__ == "__main__":
    main()
