"""The allocation-free gradient fmm_mlp_grad_noalloc (fixed workspace inside
the library) against fmm_mlp_grad with a caller workspace, through the C
ABI, on two calls in a row, the second with a smaller batch (the fixed
workspace keeps the first call's rows), and against the oracle's FP64
gradient."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[no_alloc] {msg}", flush=True)
    sys.stdout.flush()


def _inputs(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, MO.F)).astype(np.float32)
    T = rng.normal(size=(B, MO.O)).astype(np.float32)
    mask = (rng.random((B, MO.O)) < 0.8).astype(np.float32)
    return X, T, mask


def test_grad_noalloc_matches_grad_with_workspace(gpu):
    from fmmgen_amd import _lib
    from fmmgen_amd import learned as LN
    L = _lib.load()
    p = MO.init(3).astype(np.float32)
    pd = torch.from_numpy(p).to(gpu)
    st = LN._stream()
    for B, seed in ((6000, 1), (1000, 2)):
        assert B <= _lib.MLP_NOALLOC_MAX_BATCH
        X, T, mask = _inputs(B, seed)
        Xd, Td, Md = (torch.from_numpy(a).to(gpu) for a in (X, T, mask))
        work = torch.empty(int(L.fmm_mlp_work_floats(B)), dtype=torch.float32, device=gpu)
        g_ref = torch.empty(_lib.MLP_PARAMS, dtype=torch.float32, device=gpu)
        loss_ref = torch.empty(1, dtype=torch.float32, device=gpu)
        g = torch.full((_lib.MLP_PARAMS,), float("nan"), dtype=torch.float32, device=gpu)
        loss = torch.full((1,), float("nan"), dtype=torch.float32, device=gpu)
        _log(f"B={B}: fmm_mlp_grad (caller workspace)")
        _lib.check(L.fmm_mlp_grad(LN._ptr(Xd), LN._ptr(Td), LN._ptr(Md), B, LN._ptr(pd), LN._ptr(g_ref),
                                  LN._ptr(loss_ref), LN._ptr(work), st))
        _log(f"B={B}: fmm_mlp_grad_noalloc")
        _lib.check(L.fmm_mlp_grad_noalloc(LN._ptr(Xd), LN._ptr(Td), LN._ptr(Md), B, LN._ptr(pd), LN._ptr(g),
                                          LN._ptr(loss), st))
        torch.cuda.synchronize()
        # the same kernels in the same order: the gradient (per-thread
        # sequential sums) is bit-identical; the loss is summed with float
        # atomics across blocks, whose order varies: FP32 rounding only
        assert torch.equal(g, g_ref)
        assert abs(loss.item() - loss_ref.item()) <= 1e-6 * abs(loss_ref.item())
        # and both agree with the FP64 oracle
        Lo, go = MO.grad(X.astype(np.float64), T.astype(np.float64), mask.astype(np.float64),
                         p.astype(np.float64))
        gn = g.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(gn - go) <= 1e-5 * np.linalg.norm(go)
        assert abs(loss.item() - Lo) <= 1e-5 * Lo


def test_grad_noalloc_rejects_batches_beyond_the_workspace(gpu):
    from fmmgen_amd import _lib
    L = _lib.load()
    fake = ctypes.c_void_p(16)
    code = L.fmm_mlp_grad_noalloc(fake, fake, fake, _lib.MLP_NOALLOC_MAX_BATCH + 1, fake, fake, fake, None)
    assert code == 1
    assert b"exceeds the fixed workspace" in L.fmm_last_error()
