"""The operator-level Python surface (fmmgen_amd.operators) and the module
that generate_code(..., cython=True) writes in place of the reference's
Cython wrapper (fmmgen/writer.py:346-432).

CPU: the generated module's text, names and constants (the library loads
without a GPU; no operator runs).  GPU: the reference's own multipole
checks (tests/test_multipoles.py:10-203: a linear dipole, a linear
quadrupole and a quadrupole of two dipoles, exact moments), and every
operator of the module in both forms against the oracle's operators at
orders beyond the whole-body order.
"""
import importlib
import os
import sys

import numpy as np
import pytest

from fmmgen_amd import generate_code, Nterms
from oracle import oracle as O
from tests import util

TOTALORDER = 4


def _generate(tmp_path, name, order, so=0, potential=True, field=True):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        generate_code(order, name, precision="double", CSE=True, cython=True, potential=potential,
                      field=field, source_order=so, atomic=True, minpow=5, harmonic_derivs=True,
                      language="c")
    finally:
        os.chdir(cwd)
    sys.path.insert(0, str(tmp_path))
    try:
        sys.modules.pop(f"{name}_wrap", None)
        return importlib.import_module(f"{name}_wrap")
    finally:
        sys.path.remove(str(tmp_path))


def test_generated_module_names_and_constants(tmp_path):
    fmm = _generate(tmp_path, "Flavour", 5, so=0, potential=False, field=True)
    assert (tmp_path / "Flavour.h").exists() and (tmp_path / "Flavour.c").exists()
    assert fmm.FMMGEN_MINORDER == 1 and fmm.FMMGEN_MAXORDER == 5
    assert fmm.FMMGEN_SOURCEORDER == 0 and fmm.FMMGEN_SOURCESIZE == 1 and fmm.FMMGEN_OUTPUTSIZE == 3
    for op in ("P2M", "M2M", "M2L", "L2L", "L2P", "M2P", "P2P"):
        assert callable(getattr(fmm, op))
    for p in range(1, 5):
        for op in ("P2M", "M2M", "M2L", "L2L", "L2P", "M2P"):
            assert callable(getattr(fmm, f"{op}_{p}"))
    assert not hasattr(fmm, "M2M_5")
    # orders outside the generated range do nothing (the reference's switch default)
    M = np.zeros(Nterms(6))
    fmm.M2M(1.0, 0.0, 0.0, np.ones(Nterms(6)), M, 6)
    assert not M.any()


def test_generated_module_refused_for_gpu_code(tmp_path):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(Warning):
            generate_code(4, "Dev", cython=True, gpu=True)
    finally:
        os.chdir(cwd)


@pytest.mark.gpu
def test_linear_dipole(gpu, tmp_path):
    fmm = _generate(tmp_path, "LinearDipole", TOTALORDER)
    d, q = 1.0, 5.0
    x1, x2, O_ = np.array([d, 0.0, 0.0]), np.array([-d, 0.0, 0.0]), np.zeros(3)
    for order in range(1, fmm.FMMGEN_MAXORDER):
        S_px = np.zeros(Nterms(order))
        S_px[0] = q
        S_mx = np.zeros(Nterms(order))
        S_mx[0] = -q
        M = np.zeros(Nterms(order))
        fmm.M2M(*(O_ - x1), S_px, M, order)
        fmm.M2M(*(O_ - x2), S_mx, M, order)
        assert M[0] == 0.0 and M[1] == -2 * q * d and M[2] == 0.0 and M[3] == 0.0


@pytest.mark.gpu
def test_linear_quadrupole(gpu, tmp_path):
    fmm = _generate(tmp_path, "LinearQuadrupole", TOTALORDER)
    d, q = 2.0, 5.0
    x1, x2, O_ = np.array([-d, 0.0, 0.0]), np.array([d, 0.0, 0.0]), np.zeros(3)
    for order in range(2, fmm.FMMGEN_MAXORDER):
        S_mx = np.zeros(Nterms(order))
        S_mx[0] = -q
        S_px = np.zeros(Nterms(order))
        S_px[0] = -q
        S_O = np.zeros(Nterms(order))
        S_O[0] = 2 * q
        M = np.zeros(Nterms(order))
        fmm.M2M(*(O_ - x1), S_mx, M, order)
        fmm.M2M(*(O_ - x2), S_px, M, order)
        fmm.M2M(*(O_ - O_), S_O, M, order)
        assert M[0] == 0.0 and M[1] == 0.0 and M[2] == 0.0 and M[3] == 0.0
        assert M[4] == -2 * d * q


@pytest.mark.gpu
def test_quadrupole_two_dipoles(gpu, tmp_path):
    fmm = _generate(tmp_path, "QuadrupoleTwoDipoles", TOTALORDER)
    d, mux = 2.0, -5.0
    xp, xm, O_ = np.array([d, 0.0, 0.0]), np.array([-d, 0.0, 0.0]), np.zeros(3)
    for order in range(2, fmm.FMMGEN_MAXORDER):
        S_px = np.zeros(Nterms(order))
        S_px[1] = -mux
        S_mx = np.zeros(Nterms(order))
        S_mx[1] = mux
        M = np.zeros(Nterms(order))
        fmm.M2M(*(O_ - xp), S_px, M, order)
        fmm.M2M(*(O_ - xm), S_mx, M, order)
        assert M[0] == 0.0 and M[1] == 0.0 and M[2] == 0.0 and M[3] == 0.0
        assert M[4] == 2 * mux * d


@pytest.mark.gpu
@pytest.mark.parametrize("order", [3, 9, 12])
def test_module_operators_match_oracle(gpu, order):
    """Reference form (one item, numpy, in place) and batch form (CUDA
    tensors) of every operator against the oracle, charges with potential
    and field, and the field-only flavour (a subset of the same variant)."""
    import torch
    from fmmgen_amd.operators import Library
    from fmmgen_amd.codegen.layout import Msize, Lsize
    lib = O.lib(0, 4)
    ms, ls = Msize(order, 0), Lsize(order, 0)
    rng = np.random.default_rng(order)
    B = 8
    xyz = rng.uniform(-1, 1, (B, 3))
    far = xyz * 3 + np.sign(xyz) * 2.0
    ops = {"P2M": (xyz, rng.uniform(-1, 1, (B, 1)), ms), "M2M": (xyz, rng.uniform(-1, 1, (B, ms)), ms),
           "M2L": (far, rng.uniform(-1, 1, (B, ms)), ls), "L2L": (xyz, rng.uniform(-1, 1, (B, ls)), ls),
           "L2P": (xyz, rng.uniform(-1, 1, (B, ls)), 4), "M2P": (far, rng.uniform(-1, 1, (B, ms)), 4),
           "P2P": (far, rng.uniform(-1, 1, (B, 1)), 4)}
    for flavour in ("pf", "f"):
        L = Library(0, potential=(flavour == "pf"), field=True)
        for op, (x, inp, osz) in ops.items():
            ref = np.zeros((B, osz))
            for i in range(B):
                a = np.ascontiguousarray(inp[i])
                o = np.zeros(osz)
                fn = getattr(lib, f"oracle_op_{op}") if op != "P2P" else None
                if op == "P2M":
                    fn(x[i, 0], x[i, 1], x[i, 2], a[0], o.ctypes.data_as(O.c_double_p), order)
                elif op == "P2P":
                    r = x[i]
                    R = np.linalg.norm(r)
                    o[:] = [a[0] / R, *(a[0] * r / R ** 3)]
                else:
                    fn(x[i, 0], x[i, 1], x[i, 2], a.ctypes.data_as(O.c_double_p), o.ctypes.data_as(O.c_double_p),
                       order)
                ref[i] = o
            if flavour == "f" and osz == 4:
                ref = ref[:, 1:]
            got = np.zeros_like(ref)
            for i in range(B):                 # reference form
                args = (x[i, 0], x[i, 1], x[i, 2], inp[i] if op != "P2M" else inp[i, 0], got[i])
                if op == "P2P":
                    L.P2P(*args)
                else:
                    getattr(L, op)(*args, order)
            tol = 1e-14 if op == "P2P" else 1e-12
            assert util.rel_norm(got, ref) < tol, (flavour, op, "reference form")
            gb = torch.zeros(ref.shape, dtype=torch.float64, device="cuda")  # batch form
            xt = [torch.from_numpy(np.ascontiguousarray(x[:, k])).cuda() for k in range(3)]
            it = torch.from_numpy(np.ascontiguousarray(inp)).cuda()
            if op == "P2P":
                L.P2P(*xt, it, gb)
            else:
                getattr(L, f"{op}_{order}")(*xt, it, gb)
            assert util.rel_norm(gb.cpu().numpy(), ref) < tol, (flavour, op, "batch form")
