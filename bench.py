#!/usr/bin/env python3
"""Benchmark of the octree FMM evaluation on one MI355X.

Metric: FMM particle-evals/s = N_targets / t_eval, where t_eval is one full
evaluation (P2M, M2M, M2L, P2P, L2L, L2P and the final per-target sum) with
the inputs resident in HBM; the tree and list build is excluded, as the
reference's timer excludes it (example/main.cpp:138 vs 154-162).  The field
error is ||F - F_ref||_2 / ||F_ref||_2 over sampled targets against an FP64
direct sum on the GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4] [--mode headline]

The default (headline) run is the learned-order workload: N = 2^22 blobs,
tolerance 1e-6.  It finds the lowest fixed order at which the plain FP64
evaluation meets the tolerance on the sampled targets, times it, then times
the learned evaluation (FP32 mixed precision, per-cell order from the
network, sampled judge inside every evaluation) and reports the learned
throughput as ``value`` with the plain one and the speed-up beside it.

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
    "c4": dict(n=1 << 22, dist="blobs", order=10, theta=0.5, ncrit=64, so=0, ntargets=4096),
    "c5": dict(n=1 << 24, dist="mixed", order=10, theta=0.5, ncrit=128, so=0, ntargets=4096),
}

PEAK = {  # MI355X published peaks
    "hbm_GBs": 8000.0,
    "fp64_vector_TF": 78.6, "fp32_vector_TF": 157.3,
    "fp64_mfma_TF": 78.6, "fp32_mfma_TF": 157.3, "bf16_mfma_TF": 2500.0,
}
P2P_FLOP_PER_PAIR = 20  # potential + field, charges


def make_inputs(cfg, seed=1):
    from fmmgen_amd import workloads as W
    n = cfg["n"]
    pos = W.positions(cfg["dist"], n, seed)
    S = W.charges(n, seed + 1) if cfg["so"] == 0 else W.dipoles(n, seed + 1)
    return np.ascontiguousarray(pos), np.ascontiguousarray(S)


def mac_flops_m2l(p, so, harmonic=False):
    """Useful flops of one M2L pair: the full form's multiply-adds, or the
    harmonic form's (rows and columns with n_z <= 1)."""
    if harmonic:
        return 2 * sum((2 * n + 1) * (p - n + 1) ** 2 for n in range(p + 1))
    from fmmgen_amd.codegen.layout import mac_count_m2l
    return 2 * mac_count_m2l(p, so)


def harmonic_m2l(form, mode, so):
    """Whether the M2L of this mode runs in the harmonic form."""
    if so != 0 or form == "full" or mode == "bh":
        return False
    return form == "harmonic" or mode != "plain"


def cpu_baseline(cfg):
    """The oracle (C++/OpenMP restatement of the example) on a bounded sample.

    Sample: the same distribution, order, theta and ncrit at n = min(N, 2^20),
    one evaluation timed (10-20 s on the host's cores at order 10), with the
    example's run schedule OMP_SCHEDULE=dynamic,16 (example/run.sh:5) for its
    schedule(runtime) loops."""
    from oracle import oracle as O
    so = cfg["so"]
    outsz = 4 if so == 0 else 3
    n = min(cfg["n"], 1 << 20)
    sub = dict(cfg, n=n)
    pos, S = make_inputs(sub, seed=2)
    L = O.lib(so, outsz)
    L.oracle_set_runtime_schedule(16)
    t0 = time.perf_counter()
    T = O.OracleTree(pos, S, cfg["ncrit"], cfg["order"], cfg["theta"], source_order=so, outputsize=outsz)
    t1 = time.perf_counter()
    T.compute_field_fmm()
    t2 = time.perf_counter()
    cores = L.oracle_num_threads()
    return {"value": n / (t2 - t1), "unit": "particle-evals/s", "cores": int(cores), "kind": "port",
            "sample": f"oracle compute_field_fmm, n={n} {cfg['dist']}, p={cfg['order']}, "
                      f"theta={cfg['theta']}, ncrit={cfg['ncrit']}, OMP_SCHEDULE dynamic,16, one evaluation "
                      f"({t2 - t1:.1f} s; tree build {t1 - t0:.1f} s excluded)"}


def timed(tree, F, mode, steps, warmup):
    """Times ``steps`` evaluations after ``warmup`` untimed ones.

    Returns (median ms per evaluation from hipEvents recorded around each
    evaluation on the evaluation stream, mean wall-clock ms per evaluation
    of the synchronised run, per-phase ms per evaluation)."""
    import torch

    def run():
        if mode == "bh":
            tree.compute_field_bh(F)
        else:
            tree.compute_field_fmm(F, mode=mode)
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    tree.timing(True)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for k in range(steps):
        run()
        ev[k + 1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = tree.timing_read()
    tree.timing(False)
    per = sorted(ev[k].elapsed_time(ev[k + 1]) for k in range(steps))
    med = per[steps // 2] if steps % 2 else 0.5 * (per[steps // 2 - 1] + per[steps // 2])
    return med, (t1 - t0) * 1e3 / steps, {k: v[0] / steps for k, v in phases.items()}


def field_error(F, Fref, tg, so):
    import torch
    Fs = F[tg]
    fo = 1 if so == 0 else 0
    ferr = (torch.linalg.norm(Fs[:, fo:] - Fref[:, fo:]) / torch.linalg.norm(Fref[:, fo:])).item()
    perr = (torch.linalg.norm(Fs[:, 0] - Fref[:, 0]) / torch.linalg.norm(Fref[:, 0])).item() if so == 0 else None
    return ferr, perr


def error_detail(F, Fref, tg, so):
    """Per-target and per-component error figures at the sampled targets:
    the 99th percentile of |dF_i| / |F_i| (field vectors) and the example's
    per-component mean |(exact - approx) / exact| (example/main.cpp:166-188)."""
    import torch
    Fs = F[tg]
    fo = 1 if so == 0 else 0
    d = torch.linalg.norm(Fs[:, fo:] - Fref[:, fo:], dim=1) / torch.linalg.norm(Fref[:, fo:], dim=1)
    mean_abs_rel = ((Fref - Fs) / Fref).abs().mean(dim=0)
    return {"p99_target_field_rel_err": torch.quantile(d, 0.99).item(),
            "max_target_field_rel_err": d.max().item(),
            "mean_abs_rel_err_per_component": [float(v) for v in mean_abs_rel.tolist()]}


def roofline_of(phases, info, order, so, mode, m2l_macs=None, config="c4", harmonic=False):
    """Roofline entry of the dominant kernel (P2P: VALU, M2L: MFMA); times are
    hipEvent phase times on the evaluation stream.  M2L flops are the useful
    multiply-adds of the form that ran (harmonic: n_z <= 1 rows and columns)."""
    if mode == "bh":  # the walk: per particle M2P and P2P, FP64 VALU
        return {"bound": "valu", "kernel": "BH", "achieved": None, "peak": PEAK["fp64_vector_TF"],
                "unit": "TFLOP/s", "frac": None, "traffic": None, "avg_ms": phases["bh"],
                "flop_per_launch": None}
    dom = max(("P2P", "M2L"), key=lambda k: phases[k])
    if dom == "P2P":
        flops = P2P_FLOP_PER_PAIR * info["p2p_particle_pairs"]
        peak = PEAK["fp64_vector_TF"] if mode == "plain" else PEAK["fp32_vector_TF"]
        bound = "valu"
    else:
        flops = 2 * m2l_macs if m2l_macs is not None else mac_flops_m2l(order, so, harmonic) * info["n_m2l"]
        peak = {"plain": PEAK["fp64_mfma_TF"], "f32": PEAK["fp32_mfma_TF"], "learned": PEAK["fp32_mfma_TF"],
                "bf16x3": PEAK["bf16_mfma_TF"], "learned_bf16x3": PEAK["bf16_mfma_TF"]}[mode]
        bound = "mfma"
    achieved = flops / (phases[dom] * 1e-3) / 1e12
    return {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": measured_traffic(config, mode, dom, harmonic), "avg_ms": phases[dom],
            "flop_per_launch": flops, "m2l_form": ("harmonic" if harmonic else "full") if dom == "M2L" else None}


def measured_traffic(config, mode, kernel, harmonic=False):
    """HBM bytes per evaluation of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurement of the same workload and
    M2L form (profiles/*_traffic.json, scaled by a copy kernel of known
    size), or None where no such measurement exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{kernel.lower()}_{mode}_{config}_blobs_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    form = "harmonic" if harmonic else "full"
    same = d.get("workload") == config and d.get("mode") == mode and d.get("m2l_form", "full") == form
    return d["hbm_bytes_per_eval"] if same else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--mode", default="headline",
                    choices=["headline", "plain", "f32", "bf16x3", "learned", "learned_bf16x3", "bh"])
    ap.add_argument("--order", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--theta", type=float, default=None)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--model", default=None)
    ap.add_argument("--learned-thetas", default="0.5,0.55",
                    help="opening angles the learned modes may choose from (comma separated; empty: the tree's)")
    ap.add_argument("--m2l-form", default="auto", choices=["auto", "full", "harmonic"],
                    help="M2L operand form of the timed mode (auto: harmonic for the mixed modes)")
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
    if args.theta is not None:
        cfg["theta"] = args.theta
    if args.n:
        cfg["n"] = args.n
    n, so = cfg["n"], cfg["so"]
    outsz = 4 if so == 0 else 3
    learned = args.mode in ("headline", "learned", "learned_bf16x3")
    full = 12 if learned else cfg["order"]
    dev = torch.device("cuda:0")
    pos_h, S_h = make_inputs(cfg)
    pos = torch.from_numpy(pos_h).to(dev)
    S = torch.from_numpy(S_h).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = build_tree(pos, S, n, cfg["ncrit"], full, cfg["theta"], source_order=so, potential=(so == 0))
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    info = tree.info()
    print(f"# tree: {info} build {build_ms:.1f} ms", file=sys.stderr, flush=True)

    rng = np.random.default_rng(7)
    tg = torch.from_numpy(rng.choice(n, min(cfg["ntargets"], n), replace=False).astype(np.int64)).to(dev)
    Fref = direct_sum(pos, S, tg, source_order=so, outsz=outsz)
    F = torch.zeros(n, outsz, dtype=torch.float64, device=dev)
    extra = {}
    if args.mode == "headline":
        # the method without the network: plain FP64 at the lowest fixed order meeting tol
        pfix = None
        for p in range(6, 13):
            tree.set_order(p)
            tree.compute_field_fmm(F, mode="plain")
            e, _ = field_error(F, Fref, tg, so)
            print(f"# plain FP64 order {p}: field error {e:.3e}", file=sys.stderr, flush=True)
            if e <= args.tol:
                pfix = p
                break
        pfix = pfix or 12
        tree.set_order(pfix)
        ms_plain, wall_plain, ph_plain = timed(tree, F, "plain", args.steps, args.warmup)
        e_plain, _ = field_error(F, Fref, tg, so)
        print(f"# plain FP64 order {pfix}: {ms_plain:.3f} ms/eval (median), {wall_plain:.3f} (wall mean)",
              file=sys.stderr, flush=True)
        extra["plain_f64"] = {"order": pfix, "m2l_form": "full", "ms_per_step": ms_plain,
                              "ms_per_step_wall": wall_plain, "value": n / (ms_plain * 1e-3),
                              "field_rel_err": e_plain, "errors": error_detail(F, Fref, tg, so),
                              "phase_ms": ph_plain}
        # the same FP64 method with the harmonic-form M2L (reported beside it)
        tree.set_m2l_form("harmonic")
        ms_h, wall_h, ph_h = timed(tree, F, "plain", args.steps, args.warmup)
        e_h, _ = field_error(F, Fref, tg, so)
        print(f"# plain FP64 order {pfix}, harmonic M2L: {ms_h:.3f} ms/eval (median), field error {e_h:.3e}",
              file=sys.stderr, flush=True)
        extra["plain_f64_harmonic"] = {"order": pfix, "m2l_form": "harmonic", "ms_per_step": ms_h,
                                       "ms_per_step_wall": wall_h, "value": n / (ms_h * 1e-3),
                                       "field_rel_err": e_h, "phase_ms": ph_h}
        tree.set_m2l_form("auto")
        tree.set_order(full)
        mode = "learned"
    else:
        mode = args.mode
    tree.set_m2l_form(args.m2l_form)
    harm = harmonic_m2l(args.m2l_form, mode, so)
    if learned:
        from fmmgen_amd import learned as LN
        params, margin, _ = LN.load(args.model or LN.MODEL_PATH)
        thetas = [float(v) for v in args.learned_thetas.split(",") if v.strip()]
        tree.set_model(params, args.tol, margin, thetas=thetas or None)
        extra["model_margin"] = margin
        extra["learned_theta"] = tree.learned_theta
        print(f"# learned opening angle {tree.learned_theta} (candidates {thetas or [cfg['theta']]})",
              file=sys.stderr, flush=True)
    ms_per_step, ms_wall, avg = timed(tree, F, mode, args.steps, args.warmup)
    print(f"# {mode}: {ms_per_step:.3f} ms/eval (median), {ms_wall:.3f} (wall mean); phases (ms/eval): "
          + ", ".join(f"{k} {v:.3f}" for k, v in avg.items()), file=sys.stderr, flush=True)
    ferr, perr = field_error(F, Fref, tg, so)
    errs = error_detail(F, Fref, tg, so)
    stats = tree.learned_stats() if learned else None
    if stats:
        extra["learned"] = stats
        print(f"# learned stats {stats}", file=sys.stderr, flush=True)
    macs = (stats["m2l_macs_harmonic"] if harm else stats["m2l_macs"]) if stats else None
    roofline = roofline_of(avg, info, cfg["order"], so, mode, macs, args.config, harm)

    base = None
    if not args.no_cpu_baseline:
        try:
            bcfg = dict(cfg, order=extra["plain_f64"]["order"]) if "plain_f64" in extra else cfg
            base = cpu_baseline(bcfg)
        except Exception as e:  # the baseline is reported, never the target
            base = {"value": None, "error": str(e)}

    value = n / (ms_per_step * 1e-3)
    dtype = {"plain": "f64", "f32": "f32", "bf16x3": "bf16x3+f32", "learned": "f32",
             "learned_bf16x3": "bf16x3+f32", "bh": "f64"}[mode]
    workload = (f"{args.config}: N={n} {cfg['dist']}, theta {cfg['theta']}, ncrit {cfg['ncrit']}, "
                f"source_order {so}, " + (f"learned order 2-{full} and opening angle "
                                          f"{extra.get('learned_theta', cfg['theta'])}, tol {args.tol:g}"
                                          if learned else f"order {cfg['order']}"))
    out = {
        "metric": "FMM particle-evals/sec at 1e-6 rel field err, mixed+learned vs FP64 plain",
        "value": value, "unit": "particle-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "ms_per_step_wall": ms_wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic",
        "config": {"workload": workload, "mode": mode, "n": n, "order": full if learned else cfg["order"],
                   "theta": cfg["theta"], "ncrit": cfg["ncrit"],
                   "m2l_form": "harmonic" if harm else "full"},
        "field_rel_err": ferr, "potential_rel_err": perr, "meets_1e-6": bool(ferr <= 1e-6),
        "errors": errs, "timing": "median of hipEvent-timed evaluations on the evaluation stream",
        "tree_build_ms": build_ms, "phase_ms": avg, "tree": info,
        "roofline": roofline, "cpu_baseline": base,
    }
    if "plain_f64" in extra:
        out["speedup_vs_plain_f64"] = extra["plain_f64"]["ms_per_step"] / ms_per_step
        out["speedup_vs_plain_f64_harmonic"] = extra["plain_f64_harmonic"]["ms_per_step"] / ms_per_step
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
