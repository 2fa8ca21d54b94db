"""The C ABI without a GPU: the library loads, exports every entry point
include/fmm_hip.h declares, and its host-side argument checks return the
documented codes and messages.  No call here reaches a kernel."""
import ctypes
import os
import re

import pytest

from fmmgen_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fmm_hip.h")).read()
    return sorted(set(re.findall(r"\b(fmm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in _declared():
        assert hasattr(L, name), name


def test_argument_errors_are_reported():
    L = _lib.load()
    mn, mx = ctypes.c_int(), ctypes.c_int()
    assert L.fmm_variant_orders(0, 4, ctypes.byref(mn), ctypes.byref(mx)) == _lib.FMM_OK
    assert mn.value == 1 and mx.value > 8
    assert L.fmm_variant_orders(1, 3, ctypes.byref(mn), ctypes.byref(mx)) == _lib.FMM_OK
    assert mn.value == 2
    assert L.fmm_variant_orders(2, 4, ctypes.byref(mn), ctypes.byref(mx)) == _lib.FMM_OK
    assert (mn.value, mx.value) == (3, 9)
    assert L.fmm_variant_orders(0, 1, ctypes.byref(mn), ctypes.byref(mx)) == _lib.FMM_OK
    assert mn.value == 0
    assert L.fmm_variant_orders(3, 4, ctypes.byref(mn), ctypes.byref(mx)) == 2
    assert b"no compiled variant" in L.fmm_last_error()
    h = ctypes.c_void_p()
    assert L.fmm_tree_build(None, None, 10, 64, 4, 0.5, 0, 4, None, ctypes.byref(h)) == 1
    assert b"required" in L.fmm_last_error()
    fake = ctypes.c_void_p(16)
    assert L.fmm_tree_build(fake, fake, 0, 64, 4, 0.5, 0, 4, None, ctypes.byref(h)) == 1
    assert L.fmm_tree_build(fake, fake, 10, 0, 4, 0.5, 0, 4, None, ctypes.byref(h)) == 1
    assert L.fmm_tree_build(fake, fake, 10, 64, 4, -1.0, 0, 4, None, ctypes.byref(h)) == 1
    assert L.fmm_tree_build(fake, fake, 10, 64, 40, 0.5, 0, 4, None, ctypes.byref(h)) == 2
    assert b"order" in L.fmm_last_error()
    assert L.fmm_evaluate(None, 0, fake, None) == 1
    assert L.fmm_tree_info(None, None) == 1
    assert L.fmm_op_batch(99, 4, 0, 4, 1, fake, fake, 1, fake, 1, None) == 1
    assert L.fmm_op_batch(1, 4, 0, 4, 0, None, None, 1, None, 1, None) == _lib.FMM_OK   # empty batch
    assert L.fmm_tree_destroy(None) == _lib.FMM_OK
    big = _lib.MLP_NOALLOC_MAX_BATCH + 1
    assert L.fmm_mlp_grad_noalloc(fake, fake, fake, big, fake, fake, fake, None) == 1
    assert b"exceeds the fixed workspace" in L.fmm_last_error()
    assert L.fmm_mlp_grad_noalloc(None, fake, fake, 8, fake, fake, fake, None) == 1


def test_python_layer_rejects_host_tensors():
    torch = pytest.importorskip("torch")
    from fmmgen_amd import build_tree
    pos = torch.zeros(4, 3, dtype=torch.float64)
    with pytest.raises(TypeError):
        build_tree(pos, torch.zeros(4, 1, dtype=torch.float64), 4, 64, 4, 0.5)
