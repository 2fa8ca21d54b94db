#!/usr/bin/env python3
"""Benchmark of the octree FMM evaluation on one MI355X.

Metric: FMM particle-evals/s = N_targets / t_eval, where t_eval is one full
evaluation (P2M, M2M, M2L, P2P, L2L, L2P and the final per-target sum) with
the inputs resident in HBM; the tree and list build is excluded, as the
reference's timer excludes it (example/main.cpp:138 vs 154-162).  The field
error is ||F - F_ref||_2 / ||F_ref||_2 over sampled targets against an FP64
direct sum on the GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--mode plain]

Prints one JSON line (rank 0).  A run asking for more than one GPU exits:
this evaluator runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Workloads (sizes from the benchmark definitions; data are synthetic).
CONFIGS = {
    "c1": dict(n=100_000, dist="uniform", order=4, theta=0.5, ncrit=64, so=0, ntargets=1000),
    "c2": dict(n=1 << 20, dist="uniform", order=8, theta=0.5, ncrit=64, so=0, ntargets=4096),
    "c3": dict(n=1_000_000, dist="plummer", order=6, theta=0.4, ncrit=32, so=1, ntargets=4096),
    "c4": dict(n=1 << 22, dist="blobs", order=8, theta=0.5, ncrit=64, so=0, ntargets=4096),
    "c5": dict(n=1 << 24, dist="mixed", order=8, theta=0.5, ncrit=128, so=0, ntargets=4096),
}

PEAK = {  # MI355X published peaks
    "hbm_GBs": 8000.0,
    "fp64_vector_TF": 78.6, "fp32_vector_TF": 157.3,
    "fp64_mfma_TF": 78.6, "fp32_mfma_TF": 157.3, "bf16_mfma_TF": 2500.0,
}
P2P_FLOP_PER_PAIR = 20  # potential + field, charges


def make_inputs(cfg, seed=1):
    from tests import util
    n = cfg["n"]
    rng = np.random.default_rng(seed)
    if cfg["dist"] == "uniform":
        pos = rng.random((n, 3))
    elif cfg["dist"] == "plummer":
        pos = util.plummer(n, seed, a=0.1)
    elif cfg["dist"] == "blobs":
        pos = util.blobs(n, seed)
    else:  # half uniform, half Plummer a=0.05 truncated at r=1, centred in the cube
        h = n // 2
        pos = np.vstack([rng.random((h, 3)), util.plummer(n - h, seed, a=0.05) + 0.5])
        pos = pos[rng.permutation(n)]
    if cfg["so"] == 0:
        S = rng.uniform(-1.0, 1.0, (n, 1))
    else:
        S = util.dipoles(n, seed + 1)
    return np.ascontiguousarray(pos), np.ascontiguousarray(S)


def mac_flops_m2l(p, so):
    from fmmgen_amd.codegen.layout import mac_count_m2l
    return 2 * mac_count_m2l(p, so)


def cpu_baseline(cfg, budget_s=30.0):
    """The oracle (C++/OpenMP restatement of the example) on a bounded sample.

    Sample: the same distribution, order, theta and ncrit at n = min(N, 2^20)
    (the whole workload up to 2^20 particles), one evaluation timed."""
    from oracle import oracle as O
    import ctypes
    so = cfg["so"]
    outsz = 4 if so == 0 else 3
    n = min(cfg["n"], 1 << 20)
    sub = dict(cfg, n=n)
    pos, S = make_inputs(sub, seed=2)
    t0 = time.perf_counter()
    T = O.OracleTree(pos, S, cfg["ncrit"], cfg["order"], cfg["theta"], source_order=so, outputsize=outsz)
    t1 = time.perf_counter()
    T.compute_field_fmm()
    t2 = time.perf_counter()
    cores = O.lib(so, outsz).oracle_num_threads()
    return {"value": n / (t2 - t1), "unit": "particle-evals/s", "cores": int(cores), "kind": "port",
            "sample": f"oracle compute_field_fmm, n={n} {cfg['dist']}, p={cfg['order']}, "
                      f"theta={cfg['theta']}, ncrit={cfg['ncrit']}, one evaluation "
                      f"({t2 - t1:.1f} s; tree build {t1 - t0:.1f} s excluded)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--mode", default="plain", choices=["plain", "f32", "bf16x3"])
    ap.add_argument("--order", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this FMM evaluator runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from fmmgen_amd import build_tree, direct_sum
    cfg = dict(CONFIGS[args.config])
    if args.order:
        cfg["order"] = args.order
    if args.n:
        cfg["n"] = args.n
    n, so = cfg["n"], cfg["so"]
    outsz = 4 if so == 0 else 3
    dev = torch.device("cuda:0")
    pos_h, S_h = make_inputs(cfg)
    pos = torch.from_numpy(pos_h).to(dev)
    S = torch.from_numpy(S_h).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = build_tree(pos, S, n, cfg["ncrit"], cfg["order"], cfg["theta"], source_order=so, potential=(so == 0))
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    info = tree.info()
    print(f"# tree: {info} build {build_ms:.1f} ms", file=sys.stderr, flush=True)

    F = torch.zeros(n, outsz, dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        tree.compute_field_fmm(F, mode=args.mode)
    torch.cuda.synchronize()
    tree.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tree.compute_field_fmm(F, mode=args.mode)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = tree.timing_read()
    tree.timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    print(f"# {ms_per_step:.3f} ms/eval; phases (ms/eval): "
          + ", ".join(f"{k} {v[0] / max(v[1], 1):.3f}" for k, v in phases.items()), file=sys.stderr, flush=True)

    # accuracy on sampled targets
    rng = np.random.default_rng(7)
    tg = torch.from_numpy(rng.choice(n, min(cfg["ntargets"], n), replace=False).astype(np.int64)).to(dev)
    Fref = direct_sum(pos, S, tg, source_order=so, outsz=outsz)
    Fs = F[tg]
    fo = 1 if so == 0 else 0
    ferr = (torch.linalg.norm(Fs[:, fo:] - Fref[:, fo:]) / torch.linalg.norm(Fref[:, fo:])).item()
    perr = (torch.linalg.norm(Fs[:, 0] - Fref[:, 0]) / torch.linalg.norm(Fref[:, 0])).item() if so == 0 else None

    # roofline of the dominant kernel (phase timed with hipEvents on the evaluation stream)
    avg = {k: v[0] / max(v[1], 1) for k, v in phases.items()}
    dom = max(("P2P", "M2L"), key=lambda k: avg[k])
    if dom == "P2P":
        flops = P2P_FLOP_PER_PAIR * info["p2p_particle_pairs"]
        peak = PEAK["fp64_vector_TF"] if args.mode == "plain" else PEAK["fp32_vector_TF"]
        bound = "valu"
    else:
        flops = mac_flops_m2l(cfg["order"], so) * info["n_m2l"]
        peak = {"plain": PEAK["fp64_mfma_TF"], "f32": PEAK["fp32_mfma_TF"], "bf16x3": PEAK["bf16_mfma_TF"]}[args.mode]
        bound = "mfma"
    achieved = flops / (avg[dom] * 1e-3) / 1e12
    roofline = {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "avg_ms": avg[dom], "flop_per_launch": flops}

    base = None
    if not args.no_cpu_baseline:
        try:
            base = cpu_baseline(cfg)
        except Exception as e:  # the baseline is reported, never the target
            base = {"value": None, "error": str(e)}

    value = n / (ms_per_step * 1e-3)
    out = {
        "metric": "FMM particle-evals/sec at 1e-6 rel field err, mixed+learned vs FP64 plain",
        "value": value, "unit": "particle-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.mode == "plain" else ("f32" if args.mode == "f32" else "bf16x3+f32"),
        "data": "synthetic",
        "config": {"workload": f"{args.config}: N={n} {cfg['dist']}, order {cfg['order']}, theta {cfg['theta']}, "
                               f"ncrit {cfg['ncrit']}, source_order {so}", "mode": args.mode,
                   "n": n, "order": cfg["order"], "theta": cfg["theta"], "ncrit": cfg["ncrit"]},
        "field_rel_err": ferr, "potential_rel_err": perr, "meets_1e-6": bool(ferr <= 1e-6),
        "tree_build_ms": build_ms, "phase_ms": avg, "tree": info,
        "roofline": roofline, "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
