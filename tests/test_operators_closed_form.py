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


# every compiled order of the two operator variants (oracle and HIP library)
VARIANTS = [(0, 4, list(range(1, 13))), (1, 3, list(range(2, 9)))]


@pytest.mark.parametrize("so,outsz,orders", VARIANTS)
def test_shift_and_p2m_operators(so, outsz, orders):
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(1 + so)
    for p in orders:
        ms, ls = Msize(p, so), Lsize(p, so)
        x = rng.uniform(-0.6, 0.6, 3)
        M = rng.uniform(-1, 1, ms)
        L = rng.uniform(-1, 1, ls)
        assert _rel(_call(lib.oracle_op_M2M, x, M, ms, p), CF.M2M(x, M, p, so)) < TOL, ("M2M", p)
        assert _rel(_call(lib.oracle_op_L2L, x, L, ls, p), CF.L2L(x, L, p, so)) < TOL, ("L2L", p)
        assert _rel(_call(lib.oracle_op_L2P, x, L, outsz, p), CF.L2P(x, L, p, so, so == 0, True)) < TOL, ("L2P", p)
        # the reference P2M is the charge formula in the so-offset slots
        o = np.zeros(ms)
        lib.oracle_op_P2M(x[0], x[1], x[2], 0.7, o.ctypes.data_as(O.c_double_p), p)
        assert _rel(o, CF.P2M(x, 0.7, p, so)) < TOL, ("P2M", p)


@pytest.mark.parametrize("so,outsz,orders", VARIANTS)
def test_translation_operators(so, outsz, orders):
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(11 + so)
    for p in orders:
        ms, ls = Msize(p, so), Lsize(p, so)
        x = rng.uniform(1.0, 2.0, 3) * rng.choice([-1, 1], 3)
        M = rng.uniform(-1, 1, ms)
        assert _rel(_call(lib.oracle_op_M2L, x, M, ls, p), CF.M2L(tuple(x), M, p, so)) < 1e-13, ("M2L", p)
        assert _rel(_call(lib.oracle_op_M2P, x, M, outsz, p),
                    CF.M2P(tuple(x), M, p, so, so == 0, True)) < 1e-13, ("M2P", p)


@pytest.mark.parametrize("so,outsz", [(0, 4), (1, 3)])
def test_p2p(so, outsz):
    lib = O.lib(so, outsz)
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = rng.uniform(-1, 1, 3)
        S = rng.uniform(-1, 1, 1 if so == 0 else 3)
        assert _rel(_call(lib.oracle_op_P2P, x, S, outsz), CF.P2P(x, S, so, so == 0, True)) < TOL


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
