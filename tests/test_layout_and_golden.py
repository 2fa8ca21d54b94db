"""Index layout and the known-answer values of the reference's own tests.

The layout is checked against its definition (grevlex over [z, y, x]:
degree, then descending n_x, then descending n_y; fmmgen/utils.py:65-117)
and the slots the reference's generated code uses
(example/operators.cpp M2M_2: slots 4..9 = xx, xy, xz, yy, yz, zz).
The golden values are those of tests/test_multipoles.py and
tests/test_field.py of the reference, run against the oracle's operators
(the generated host C++ the CPU baseline also uses).
"""
import numpy as np
import pytest

from fmmgen_amd.codegen.layout import (Nterms, Msize, Lsize, generate_mappings, idx, monomials,
                                       mac_count_m2l, sourcesize, outputsize)
from oracle import oracle as O


def test_nterms_and_sizes():
    # fmmgen/utils.py:25-34 and example/utils.cpp:7-26
    assert [Nterms(p) for p in range(-1, 6)] == [0, 1, 4, 10, 20, 35, 56]
    assert Nterms(8) == 165 and Nterms(12) == 455
    assert Msize(6, 1) == 83 and Lsize(6, 1) == 56
    assert sourcesize(0) == 1 and sourcesize(1) == 3
    assert outputsize(True, True) == 4 and outputsize(False, True) == 3 and outputsize(True, False) == 1
    # useful multiply-adds per M2L pair: C(p+6, 6) for charges
    from math import comb
    for p in range(0, 13):
        assert mac_count_m2l(p, 0) == comb(p + 6, 6)
    assert mac_count_m2l(6, 1) == comb(12, 6) - Nterms(6)


def test_mapping_is_graded_order():
    for p in range(0, 9):
        fwd, rev = generate_mappings(p)
        keys = sorted(fwd, key=lambda n: (sum(n), -n[0], -n[1]))
        assert [fwd[k] for k in keys] == list(range(len(keys)))
        assert all(idx(n) == i for n, i in fwd.items())
        assert len(fwd) == Nterms(p)
    fwd, _ = generate_mappings(2)
    assert [fwd[n] for n in [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]] == [4, 5, 6, 7, 8, 9]
    # source order 1 starts at (1, 0, 0) (tests/test_M_shift.py:106-116)
    fwd1, _ = generate_mappings(2, source_order=1)
    assert fwd1[(1, 0, 0)] == 0 and fwd1[(0, 0, 2)] == 8
    with pytest.raises(ValueError):
        generate_mappings(0, source_order=1)
    # prefix property: the order-q layout is the first Nterms(q) slots of order p
    for q in range(0, 6):
        assert list(monomials(8))[:Nterms(q)] == list(monomials(q))


def _m2m(lib, x, M, Ms, order):
    lib.oracle_op_M2M(x[0], x[1], x[2], M.ctypes.data_as(O.c_double_p), Ms.ctypes.data_as(O.c_double_p), order)


def test_linear_dipole():
    """tests/test_multipoles.py:42-65: +q at (d,0,0), -q at (-d,0,0)."""
    lib = O.lib(0, 4)
    d, q = 1.0, 5.0
    x1, x2, Or = np.array([d, 0, 0.0]), np.array([-d, 0, 0.0]), np.zeros(3)
    for order in range(1, 5):
        ms = Nterms(order)
        Sp, Sm, M = np.zeros(ms), np.zeros(ms), np.zeros(ms)
        Sp[0], Sm[0] = q, -q
        _m2m(lib, Or - x1, Sp, M, order)
        _m2m(lib, Or - x2, Sm, M, order)
        assert M[0] == 0.0 and M[1] == -2 * q * d and M[2] == 0.0 and M[3] == 0.0


def test_linear_quadrupole():
    """tests/test_multipoles.py:106-133: -q at +-d, +2q at the origin."""
    lib = O.lib(0, 4)
    d, q = 2.0, 5.0
    x1, x2, Or = np.array([-d, 0, 0.0]), np.array([d, 0, 0.0]), np.zeros(3)
    for order in range(2, 5):
        ms = Nterms(order)
        S1, S2, S0, M = np.zeros(ms), np.zeros(ms), np.zeros(ms), np.zeros(ms)
        S1[0], S2[0], S0[0] = -q, -q, 2 * q
        _m2m(lib, Or - x1, S1, M, order)
        _m2m(lib, Or - x2, S2, M, order)
        _m2m(lib, Or - Or, S0, M, order)
        assert list(M[:4]) == [0.0, 0.0, 0.0, 0.0] and M[4] == -2 * d * q


def test_quadrupole_two_dipoles():
    """tests/test_multipoles.py:173-197: dipoles -+mux in slot 1 at +-d."""
    lib = O.lib(0, 4)
    d, mux = 2.0, -5.0
    xp, xm, Or = np.array([d, 0, 0.0]), np.array([-d, 0, 0.0]), np.zeros(3)
    for order in range(2, 5):
        ms = Nterms(order)
        Sp, Sm, M = np.zeros(ms), np.zeros(ms), np.zeros(ms)
        Sp[1], Sm[1] = -mux, mux
        _m2m(lib, Or - xp, Sp, M, order)
        _m2m(lib, Or - xm, Sm, M, order)
        assert list(M[:4]) == [0.0, 0.0, 0.0, 0.0] and M[4] == 2 * mux * d


def test_monopole_m2p_field():
    """tests/test_field.py:42-77, with the bound the reference meant to
    check: q at (2,0,0) expanded at the origin, evaluated at (20,0,0).  The y
    and z field are exactly zero, and on the axis the order-p expansion is
    the truncated series q sum_{k<=p} (k+1) d^k / X^(k+2) of q / (X - d)^2."""
    lib = O.lib(0, 4)
    d, q, X = 2.0, 5.0, 20.0
    x1, x2 = np.array([d, 0, 0.0]), np.array([X, 0, 0])
    exact = q / (X - d) ** 2
    last = np.inf
    for order in range(1, 12):
        ms = Nterms(order)
        S, M, F = np.zeros(ms), np.zeros(ms), np.zeros(4)
        S[0] = q
        _m2m(lib, -x1, S, M, order)
        lib.oracle_op_M2P(x2[0], x2[1], x2[2], M.ctypes.data_as(O.c_double_p), F.ctypes.data_as(O.c_double_p), order)
        assert F[2] == 0.0 and F[3] == 0.0
        series = q * sum((k + 1) * d ** k / X ** (k + 2) for k in range(order + 1))
        assert abs(F[1] - series) <= 1e-15 * abs(series)
        potential = q * sum(d ** k / X ** (k + 1) for k in range(order + 1))
        assert abs(F[0] - potential) <= 1e-15 * abs(potential)
        e = abs(F[1] - exact) / exact
        assert e < last
        last = e


def test_out_of_range_order_is_a_silent_noop():
    """Generated wrappers ignore orders outside the generated range
    (fmmgen/writer.py:291-309)."""
    lib = O.lib(0, 4)
    M = np.zeros(455)
    Ms = np.zeros(455)
    M[0] = 1.0
    _m2m(lib, np.array([0.1, 0.2, 0.3]), M, Ms, 40)
    assert not Ms.any()
