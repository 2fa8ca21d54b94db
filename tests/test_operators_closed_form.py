"""Pins the generated operators (host C++ flavour, the oracle's operators)
against closed forms in 50-digit mpmath, and against invariants a dropped
term, a wrong sign or a transposed index would break.  The HIP flavour is
printed from the same programs and compared with these on the GPU."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import closed_form as CF
from fmmgen_amd.codegen.layout import Msize, Lsize

TOL = 2e-14


def _call(fn, x, inp, nout, p=None):
    a = np.ascontiguousarray(inp, dtype=np.float64)
    o = np.zeros(nout)
    if p is None:
        fn(x[0], x[1], x[2], a.ctypes.data_as(O.c_double_p), o.ctypes.data_as(O.c_double_p))
    else:
        fn(x[0], x[1], x[2], a.ctypes.data_as(O.c_double_p), o.ctypes.data_as(O.c_double_p), p)
    return o


def _rel(a, b):
    b = np.array([float(v) for v in b])
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# every compiled order of the operator variants with their own operator
# programs (oracle and HIP library): (source order, output size, potential,
# field, orders).  The charge variants with potential only or field only
# share every program but L2P / M2P / P2P with the first; they are pinned
# against it below.
VARIANTS = [(0, 4, True, True, list(range(1, 13))), (1, 3, False, True, list(range(2, 9))),
            (2, 4, True, True, list(range(3, 9)))]


@pytest.mark.parametrize("so,outsz,pot,field,orders", VARIANTS)
def test_shift_and_p2m_operators(so, outsz, pot, field, orders):
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(1 + so)
    for p in orders:
        ms, ls = Msize(p, so), Lsize(p, so)
        x = rng.uniform(-0.6, 0.6, 3)
        M = rng.uniform(-1, 1, ms)
        L = rng.uniform(-1, 1, ls)
        assert _rel(_call(lib.oracle_op_M2M, x, M, ms, p), CF.M2M(x, M, p, so)) < TOL, ("M2M", p)
        assert _rel(_call(lib.oracle_op_L2L, x, L, ls, p), CF.L2L(x, L, p, so)) < TOL, ("L2L", p)
        assert _rel(_call(lib.oracle_op_L2P, x, L, outsz, p), CF.L2P(x, L, p, so, pot, field)) < TOL, ("L2P", p)
        # the reference P2M is the charge formula in the so-offset slots
        o = np.zeros(ms)
        lib.oracle_op_P2M(x[0], x[1], x[2], 0.7, o.ctypes.data_as(O.c_double_p), p)
        assert _rel(o, CF.P2M(x, 0.7, p, so)) < TOL, ("P2M", p)


@pytest.mark.parametrize("so,outsz,pot,field,orders", VARIANTS)
def test_translation_operators(so, outsz, pot, field, orders):
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(11 + so)
    for p in orders:
        ms, ls = Msize(p, so), Lsize(p, so)
        x = rng.uniform(1.0, 2.0, 3) * rng.choice([-1, 1], 3)
        M = rng.uniform(-1, 1, ms)
        assert _rel(_call(lib.oracle_op_M2L, x, M, ls, p), CF.M2L(tuple(x), M, p, so)) < 1e-13, ("M2L", p)
        assert _rel(_call(lib.oracle_op_M2P, x, M, outsz, p),
                    CF.M2P(tuple(x), M, p, so, pot, field)) < 1e-13, ("M2P", p)


@pytest.mark.parametrize("so,outsz,pot,field", [(0, 4, True, True), (1, 3, False, True), (2, 4, True, True),
                                                (0, 1, True, False), (0, 3, False, True)])
def test_p2p(so, outsz, pot, field):
    from fmmgen_amd.codegen.layout import sourcesize
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = rng.uniform(-1, 1, 3)
        S = rng.uniform(-1, 1, sourcesize(so))
        assert _rel(_call(lib.oracle_op_P2P, x, S, outsz), CF.P2P(x, S, so, pot, field)) < TOL


@pytest.mark.parametrize("outsz,keep,orders", [(1, [0], list(range(0, 13))), (3, [1, 2, 3], list(range(1, 13)))])
def test_output_variants_agree_with_potential_and_field(outsz, keep, orders):
    """The charge variants with potential only / field only: the expansion
    operators are the pinned variant's exactly, and L2P / M2P give its
    components; order 0 (potential only) against the closed forms."""
    lib, ref = O.lib(0, outsz), O.lib(0, 4)
    rng = np.random.default_rng(21)
    for p in orders:
        ms, ls = Msize(p, 0), Lsize(p, 0)
        x = rng.uniform(-0.6, 0.6, 3)
        far = rng.uniform(1.0, 2.0, 3) * rng.choice([-1, 1], 3)
        M = rng.uniform(-1, 1, ms)
        L = rng.uniform(-1, 1, ls)
        if p == 0:
            assert _rel(_call(lib.oracle_op_L2P, x, L, 1, 0), CF.L2P(x, L, 0, 0, True, False)) < TOL
            assert _rel(_call(lib.oracle_op_M2P, far, M, 1, 0), CF.M2P(tuple(far), M, 0, 0, True, False)) < TOL
            assert _rel(_call(lib.oracle_op_M2L, far, M, 1, 0), CF.M2L(tuple(far), M, 0, 0)) < TOL
            continue
        for op, xx, inp, n in (("M2M", x, M, ms), ("L2L", x, L, ls), ("M2L", far, M, ls)):
            a = _call(getattr(lib, f"oracle_op_{op}"), xx, inp, n, p)
            b = _call(getattr(ref, f"oracle_op_{op}"), xx, inp, n, p)
            assert np.array_equal(a, b), (op, p)
        for op, xx, inp in (("L2P", x, L), ("M2P", far, M)):
            a = _call(getattr(lib, f"oracle_op_{op}"), xx, inp, outsz, p)
            b = _call(getattr(ref, f"oracle_op_{op}"), xx, inp, 4, p)[keep]
            assert np.allclose(a, b, rtol=1e-15, atol=0), (op, p)


def test_shift_composition_is_exact():
    """M2M(a) then M2M(b) == M2M(a + b); the same for L2L (polynomial shifts)."""
    lib = O.lib(0, 4)
    rng = np.random.default_rng(3)
    for p in (3, 6, 8):
        ms = Msize(p, 0)
        a, b = rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3)
        M = rng.uniform(-1, 1, ms)
        two = _call(lib.oracle_op_M2M, b, _call(lib.oracle_op_M2M, a, M, ms, p), ms, p)
        one = _call(lib.oracle_op_M2M, a + b, M, ms, p)
        assert np.linalg.norm(two - one) / np.linalg.norm(one) < 1e-13
        L = rng.uniform(-1, 1, ms)
        two = _call(lib.oracle_op_L2L, b, _call(lib.oracle_op_L2L, a, L, ms, p), ms, p)
        one = _call(lib.oracle_op_L2L, a + b, L, ms, p)
        assert np.linalg.norm(two - one) / np.linalg.norm(one) < 1e-13


def test_l2p_of_m2l_at_centre_is_m2p():
    """L2P(0) o M2L(x) equals M2P(x): both evaluate the multipole at the target."""
    lib = O.lib(0, 4)
    rng = np.random.default_rng(8)
    for p in (2, 5, 8):
        ms = Msize(p, 0)
        x = np.array([1.5, -0.7, 0.9])
        M = rng.uniform(-1, 1, ms)
        L = _call(lib.oracle_op_M2L, x, M, ms, p)
        F1 = _call(lib.oracle_op_L2P, np.zeros(3), L, 4, p)
        F2 = _call(lib.oracle_op_M2P, x, M, 4, p)
        # the potential agrees exactly; the field uses one degree less in L2P
        assert abs(F1[0] - F2[0]) <= 1e-13 * abs(F2[0])


def test_field_is_minus_gradient_of_potential():
    """F = -grad(phi) for L2P, M2P and P2P, by a central difference."""
    lib = O.lib(0, 4)
    rng = np.random.default_rng(9)
    p = 6
    ms = Msize(p, 0)
    h = 1e-5
    for fn, arr, nargs in ((lib.oracle_op_L2P, rng.uniform(-1, 1, ms), p),
                           (lib.oracle_op_M2P, rng.uniform(-1, 1, ms), p),
                           (lib.oracle_op_P2P, np.array([0.8]), None)):
        x = np.array([1.3, -0.4, 0.8])
        F = _call(fn, x, arr, 4, nargs)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            d = (_call(fn, x + e, arr, 4, nargs)[0] - _call(fn, x - e, arr, 4, nargs)[0]) / (2 * h)
            assert abs(-d - F[1 + a]) < 1e-7 * max(1.0, abs(F[1 + a]))


def test_multipole_expansion_converges_to_direct_potential():
    """P2M of a charge cloud, then M2P far away: error falls with order like (r/R)^(p+1)."""
    lib = O.lib(0, 4)
    rng = np.random.default_rng(10)
    pts = rng.uniform(-0.2, 0.2, (30, 3))
    q = rng.uniform(-1, 1, 30)
    tgt = np.array([1.1, 0.9, -0.7])
    exact = sum(q[i] / np.linalg.norm(tgt - pts[i]) for i in range(30))
    errs = []
    for p in range(1, 13):
        M = np.zeros(Msize(p, 0))
        for i in range(30):
            S = np.zeros(Msize(p, 0))
            S[0] = q[i]
            M += _call(lib.oracle_op_M2M, -pts[i], S, Msize(p, 0), p)
        errs.append(abs(_call(lib.oracle_op_M2P, tgt, M, 4, p)[0] - exact))
    # (0.35 / 1.6)^13 ~ 3e-9: the order-12 error sits far below order 8's
    assert errs[-1] < 1e-8 * abs(exact) + 1e-12
    assert all(errs[i + 2] < errs[i] for i in range(len(errs) - 2))


def test_derivative_series_matches_numerical_differentiation():
    """The binomial series the closed forms use for D_n = d^n(1/R) equals
    mpmath's numerical differentiation (an unrelated method) to 40 digits."""
    for x in ((1.3, -0.7, 0.9), (-2.0, 0.25, 1.5), (0.0, 0.0, 1.75)):
        a = CF.derivs_series(x, 6)
        b = CF.derivs_numdiff(x, 6)
        for k in b:
            assert abs(a[k] - b[k]) <= 1e-40 * max(abs(b[k]), 1), (x, k)
