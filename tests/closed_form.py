"""Closed-form operator values in mpmath (50 digits) -- test infrastructure.

Each formula is the textbook statement of the operator the reference
generates (citations to the reference's files), evaluated term by term with
no code shared with the project's generator.  In particular the
derivatives D_n = d^n(1/R) come from mpmath's numerical differentiation at
raised precision, not from the recurrence the generator uses.
"""
from functools import lru_cache
from math import factorial

import mpmath as mp

from fmmgen_amd.codegen.layout import monomials, Nterms

mp.mp.dps = 50


def _f(n):
    return factorial(n[0]) * factorial(n[1]) * factorial(n[2])


def _pow(x, n):
    return mp.mpf(x[0]) ** n[0] * mp.mpf(x[1]) ** n[1] * mp.mpf(x[2]) ** n[2]


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _deg(n):
    return n[0] + n[1] + n[2]


def mslot(so):
    return {n: i for i, n in enumerate(monomials(40, so)) if _deg(n) <= 40}


@lru_cache(maxsize=None)
def derivs_numdiff(x, maxdeg):
    """D_n = d^n (1/R) at x for |n| <= maxdeg (fmmgen/expansions.py:108-141),
    by mpmath's numerical differentiation (slow above degree 8)."""
    X = tuple(mp.mpf(v) for v in x)
    f = lambda a, b, c: 1 / mp.sqrt(a * a + b * b + c * c)  # noqa: E731
    out = {}
    for n in monomials(maxdeg, 0):
        out[n] = mp.diff(f, X, n)
    return out


def _polymul(a, b, maxdeg):
    out = {}
    for ka, va in a.items():
        da = _deg(ka)
        for kb, vb in b.items():
            if da + _deg(kb) > maxdeg:
                continue
            k = _add(ka, kb)
            out[k] = out.get(k, 0) + va * vb
    return out


@lru_cache(maxsize=None)
def derivs_series(x, maxdeg):
    """D_n = d^n (1/R) at x for |n| <= maxdeg from the binomial series of
    1/|x + h| = (1/r) (1 + u)^(-1/2), u = (2 x.h + h.h) / r^2:
    D_n = n! [h^n] sum_k C(-1/2, k) u^k / r.  Exact rational polynomial
    arithmetic in h, one mpmath square root at the end; shares nothing with
    numerical differentiation or the generator's recurrence."""
    from fractions import Fraction
    X = [Fraction(v) for v in x]
    r2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    u = {}
    for a in range(3):
        e = [0, 0, 0]
        e[a] = 1
        u[tuple(e)] = 2 * X[a] / r2
        e[a] = 2
        u[tuple(e)] = Fraction(1) / r2
    total = {(0, 0, 0): Fraction(1)}
    uk = {(0, 0, 0): Fraction(1)}
    binom = Fraction(1)
    for k in range(1, maxdeg + 1):  # u^k has lowest degree k
        uk = _polymul(uk, u, maxdeg)
        binom = binom * (Fraction(-1, 2) - (k - 1)) / k
        for key, v in uk.items():
            total[key] = total.get(key, 0) + binom * v
    rinv = 1 / mp.sqrt(mp.mpf(r2.numerator) / r2.denominator)
    out = {}
    for n in monomials(maxdeg, 0):
        c = total.get(n, Fraction(0)) * _f(n)
        out[n] = mp.mpf(c.numerator) / c.denominator * rinv
    return out


def derivs(x, maxdeg):
    """D_n = d^n (1/R) at x for |n| <= maxdeg: the binomial series (checked
    against numerical differentiation in the tests)."""
    return derivs_series(tuple(float(v) for v in x), maxdeg)


def P2M(x, q, p, so):
    """M_n = (-1)^|n| q x^n / n! over the so-offset slots (fmmgen/expansions.py:22-39)."""
    return [(-1) ** _deg(n) * mp.mpf(q) * _pow(x, n) / _f(n) for n in monomials(p, so)]


def M2M(x, M, p, so):
    """Ms_n = sum_{k <= n, |n-k| >= so} M_{n-k} x^k / k! (fmmgen/expansions.py:42-78)."""
    idx = {n: i for i, n in enumerate(monomials(p, so))}
    out = []
    for n in monomials(p, so):
        s = mp.mpf(0)
        for k in monomials(_deg(n), 0):
            r = _sub(n, k)
            if min(r) < 0 or _deg(r) < so:
                continue
            s += mp.mpf(M[idx[r]]) * _pow(x, k) / _f(k)
        out.append(s)
    return out


def M2L(x, M, p, so):
    """L_n = sum_{so <= |m| <= p - |n|} D_{n+m} M_m (fmmgen/expansions.py:144-172)."""
    D = derivs(tuple(x), p)
    idx = {n: i for i, n in enumerate(monomials(p, so))}
    out = []
    for n in monomials(p - so, 0):
        s = mp.mpf(0)
        for m in monomials(p - _deg(n), so):
            s += D[_add(n, m)] * mp.mpf(M[idx[m]])
        out.append(s)
    return out


def L2L(x, L, p, so):
    """Ls_n = sum_{|n+k| <= p - so} L_{n+k} x^k / k! (fmmgen/expansions.py:175-194)."""
    q = p - so
    idx = {n: i for i, n in enumerate(monomials(q, 0))}
    out = []
    for n in monomials(q, 0):
        s = mp.mpf(0)
        for k in monomials(q - _deg(n), 0):
            s += mp.mpf(L[idx[_add(n, k)]]) * _pow(x, k) / _f(k)
        out.append(s)
    return out


def _fields(phi_terms, grad_terms, potential, field):
    out = []
    if potential:
        out.append(phi_terms)
    if field:
        out += grad_terms
    return out


def L2P(x, L, p, so, potential=True, field=True):
    """phi = sum_n L_n x^n/n!, F = -grad phi (fmmgen/expansions.py:197-207)."""
    q = p - so
    ns = monomials(q, 0)
    phi = sum((mp.mpf(L[i]) * _pow(x, n) / _f(n) for i, n in enumerate(ns)), mp.mpf(0))
    grads = []
    for a in range(3):
        s = mp.mpf(0)
        for i, n in enumerate(ns):
            if n[a] >= 1:
                nm = list(n)
                nm[a] -= 1
                s += mp.mpf(L[i]) * _pow(x, tuple(nm)) / _f(tuple(nm))
        grads.append(-s)
    return _fields(phi, grads, potential, field)


def M2P(x, M, p, so, potential=True, field=True):
    """phi = sum_m M_m D_m(x), F = -grad phi (fmmgen/generator.py:174-206)."""
    D = derivs(tuple(x), p + 1)
    ms = monomials(p, so)
    phi = sum((mp.mpf(M[i]) * D[m] for i, m in enumerate(ms)), mp.mpf(0))
    grads = []
    for a in range(3):
        e = [0, 0, 0]
        e[a] = 1
        grads.append(-sum((mp.mpf(M[i]) * D[_add(m, tuple(e))] for i, m in enumerate(ms)), mp.mpf(0)))
    return _fields(phi, grads, potential, field)


def P2P(x, S, so, potential=True, field=True):
    """so=0: phi = q/R, F = q x/R^3; so=1: phi = S . grad(1/R); so=2:
    phi = sum_s S_s d^s (1/R); F = -grad phi (fmmgen/generator.py:257-288)."""
    X = [mp.mpf(v) for v in x]
    R = mp.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    if so == 0:
        q = mp.mpf(S[0])
        return _fields(q / R, [q * X[a] / R ** 3 for a in range(3)], potential, field)
    if so == 2:
        # quadrupole source S_s, |s| = 2, s = e_a + e_b in graded slot order:
        # phi = sum_s S_s d_a d_b (1/R),  d_a d_b (1/R) = 3 x_a x_b/R^5 - delta_ab/R^3,
        # d_c d_a d_b (1/R) = 3 (delta_ac x_b + delta_bc x_a + delta_ab x_c)/R^5 - 15 x_a x_b x_c/R^7
        pairs = []
        for n in monomials(2, 2):
            ab = [a for a in range(3) for _ in range(n[a])]
            pairs.append((ab[0], ab[1]))
        d = lambda i, j: 1 if i == j else 0  # noqa: E731
        phi = sum(mp.mpf(S[i]) * (3 * X[a] * X[b] / R ** 5 - d(a, b) / R ** 3) for i, (a, b) in enumerate(pairs))
        grads = []
        for c in range(3):
            g = sum(mp.mpf(S[i]) * (3 * (d(a, c) * X[b] + d(b, c) * X[a] + d(a, b) * X[c]) / R ** 5
                                    - 15 * X[a] * X[b] * X[c] / R ** 7) for i, (a, b) in enumerate(pairs))
            grads.append(-g)
        return _fields(phi, grads, potential, field)
    # dipole: grad(1/R) = -x/R^3, d_a d_b (1/R) = 3 x_a x_b / R^5 - delta_ab / R^3
    phi = -sum(mp.mpf(S[i]) * X[i] for i in range(3)) / R ** 3
    grads = []
    for a in range(3):
        s = mp.mpf(0)
        for i in range(3):
            s += mp.mpf(S[i]) * (3 * X[i] * X[a] / R ** 5 - (1 if i == a else 0) / R ** 3)
        grads.append(-s)
    return _fields(phi, grads, potential, field)
