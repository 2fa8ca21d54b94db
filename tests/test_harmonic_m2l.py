"""The harmonic form of the M2L translation (CPU only).

The local expansion of a multipole, L_n = sum_m D_{n+m} M_m over
|n| + |m| <= p (fmmgen/expansions.py:144-172), with D_k = d^k (1/R), is
computed by the GPU GEMM in a reduced form:

* the fold: every multipole term with m_z >= 2 is moved onto m_z <= 1 by
  D_{k+2z} = -D_{k+2x} - D_{k+2y} (1/R is harmonic), so M_(a,b,c+2k) adds
  (-1)^k C(k,i) M to slot (a + 2i, b + 2(k - i), c);
* only the rows L_n with n_z <= 1 are translated, and the others follow from
  L_(a,b,c+2k) = (-1)^k sum_i C(k,i) L_(a+2i, b+2(k-i), c) (the unfold);
* the (p + 1)^2 harmonic slots are the multi-indices with n_z <= 1, degree
  d at d^2 .. (d+1)^2 - 1, in the graded order restricted to n_z <= 1.

These checks restate the layout and the two maps and pin them with the
oracle's generated M2L: M2L(fold(M)) = M2L(M) in every component, and
unfold of the n_z <= 1 rows of M2L(M) = M2L(M), at orders 1-12 and several
displacements; a sign mistake in either map fails them.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def nterms(p):
    return (p + 1) * (p + 2) * (p + 3) // 6 if p >= 0 else 0


def slot(x, y, z):
    d = x + y + z
    a = d - x
    return nterms(d - 1) + a * (a + 1) // 2 + (d - x - y)


def multis(p):
    return [(d - a, a - b, b) for d in range(p + 1) for a in range(d + 1) for b in range(a + 1)]


def harm_multi(j):
    d = math.isqrt(j)
    r = j - d * d
    a = (r + 1) // 2
    z = 0 if r == 0 else (r + 1) % 2
    return (d - a, a - z, z)


def harm_slot(x, y, z):
    d, a = x + y + z, y + z
    return d * d + (0 if a == 0 else 2 * a - 1 + z)


def fold(M, p, sign=-1):
    """Graded multipole -> harmonic slots ((p+1)^2)."""
    out = np.zeros((p + 1) ** 2)
    for j in range((p + 1) ** 2):
        x, y, z = harm_multi(j)
        for k in range((x + y) // 2 + 1):
            for i in range(k + 1):
                sx, sy = x - 2 * i, y - 2 * (k - i)
                if sx >= 0 and sy >= 0:
                    out[j] += sign ** k * math.comb(k, i) * M[slot(sx, sy, z + 2 * k)]
    return out


def unfold(Lh, p, sign=-1):
    """Harmonic rows -> graded local expansion."""
    out = np.zeros(nterms(p))
    for (x, y, z) in multis(p):
        k, c = z // 2, z % 2
        out[slot(x, y, z)] = sign ** k * sum(math.comb(k, i) * Lh[harm_slot(x + 2 * i, y + 2 * (k - i), c)]
                                             for i in range(k + 1))
    return out


def embed(Mh, p):
    """Harmonic slots -> a graded array (zero where n_z >= 2)."""
    out = np.zeros(nterms(p))
    for j in range((p + 1) ** 2):
        out[slot(*harm_multi(j))] = Mh[j]
    return out


def _m2l(lib, x, M, p):
    L = np.zeros(nterms(p))
    lib.oracle_op_M2L(x[0], x[1], x[2], O._p(M), O._p(L), p)
    return L


def _multipole(lib, p, rng):
    M = np.zeros(nterms(p))
    for q, r in zip(rng.uniform(-1, 1, 12), rng.uniform(-0.5, 0.5, (12, 3))):
        lib.oracle_op_P2M(-r[0], -r[1], -r[2], q, O._p(M), p)
    return M


def _failures(p, fold_sign=-1, unfold_sign=-1):
    lib = O.lib(0, 4)
    rng = np.random.default_rng(p)
    M = _multipole(lib, p, rng)
    fails = []
    for x in ([2.1, -1.3, 1.7], [0.0, 0.0, -3.0], [-2.5, 2.0, 0.5]):
        L = _m2l(lib, x, M, p)
        scale = np.abs(L).max()
        Lf = _m2l(lib, x, embed(fold(M, p, fold_sign), p), p)
        if np.abs(Lf - L).max() > 1e-13 * scale:
            fails.append(f"fold at {x}")
        rows = np.array([L[slot(*harm_multi(j))] for j in range((p + 1) ** 2)])
        if np.abs(unfold(rows, p, unfold_sign) - L).max() > 1e-13 * scale:
            fails.append(f"unfold at {x}")
    return fails


def test_harmonic_layout_is_a_graded_prefix_bijection():
    for p in range(16):
        hs = [harm_multi(j) for j in range((p + 1) ** 2)]
        assert len(set(hs)) == len(hs)
        assert set(hs) == {m for m in multis(p) if m[2] <= 1}
        assert all(harm_slot(*m) == j for j, m in enumerate(hs))
        assert all(sum(m) == math.isqrt(j) for j, m in enumerate(hs))
        # graded order restricted to n_z <= 1
        assert [slot(*m) for m in hs] == sorted(slot(*m) for m in hs)


@pytest.mark.parametrize("p", range(1, 13))
def test_fold_and_unfold_reproduce_the_generated_m2l(p):
    assert _failures(p) == []


@pytest.mark.parametrize("fold_sign,unfold_sign", [(1, -1), (-1, 1)])
def test_sign_mistakes_are_detected(fold_sign, unfold_sign):
    assert _failures(6, fold_sign, unfold_sign)
