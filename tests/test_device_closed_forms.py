"""Closed forms behind the wave-per-cell kernels, pinned to the generated bodies.

The FP32 P2M of charges (k_p2m_wave) and the FP64 top-level M2M / L2L of
charges (k_m2m_top_wave, k_l2l_top_wave) do not run the generated operator
bodies; they evaluate the closed forms
    P2M_S: M[a,b,c] += q x^a y^b z^c / (a! b! c!)
    M2M:   Ms[n] += sum over k <= n (componentwise) of M[n - k] x^k / k!
    L2L:   Ls[n] += sum over |n + k| <= p of L[n + k] x^k / k!
in the multi-index layout.  This test evaluates the straight-line HIP bodies
that the generator emits (fmmgen_amd.codegen.writer.device_source, the same
text the library compiles) with Python floats and compares them with those
closed forms at orders 2-5.  A wrong sign convention, a missing factorial or
a transposed shift direction fails it.  The kernels themselves are checked
on the GPU by the mixed-precision parity tests against the plain FP64 path.
"""
import math
import re

import numpy as np
import pytest

from fmmgen_amd.codegen.writer import device_source

SRC = device_source(6)  # operators of orders 1-5


def _body(name):
    i = SRC.index(f"void {name}(")
    return SRC[i:SRC.index("\n}\n", i)]


def _num(expr):
    return re.sub(r"T\(\(?([-0-9.e]+)\)?\)", r"(\1)", expr)


def _run(name, x, y, z, inp, out_name, in_name, size):
    env = {"x": x, "y": y, "z": z}
    out = [0.0] * size
    for line in _body(name).split("\n")[1:]:
        line = line.strip()
        m = re.match(r"const T (\w+) = (.*);", line)
        if m:
            env[m.group(1)] = eval(_num(m.group(2)), {}, env)
            continue
        m = re.match(out_name + r"\[(\d+)\] \+= (.*);", line)
        if m:
            e = re.sub(in_name + r"\[(\d+)\]", r"IN[\1]", _num(m.group(2)))
            out[int(m.group(1))] += eval(e, {}, dict(env, IN=inp))
    return np.array(out)


def _nterms(d):
    return (d + 1) * (d + 2) * (d + 3) // 6 if d >= 0 else 0


def _multi(j):
    d = 0
    while _nterms(d) <= j:
        d += 1
    r = j - _nterms(d - 1)
    a = 0
    while (a + 1) * (a + 2) // 2 <= r:
        a += 1
    b = r - a * (a + 1) // 2
    return d - a, a - b, b


def _slot(x, y, z):
    d = x + y + z
    a = d - x
    return _nterms(d - 1) + a * (a + 1) // 2 + (d - x - y)


def _fact3(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c)


X, Y, Z = 0.3, -0.7, 0.55


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_p2m_s_closed_form(p):
    ms, q = _nterms(p), 1.7
    # P2M_S reads the source strength as S[0]
    body_out = [0.0] * ms
    env = {"x": X, "y": Y, "z": Z}
    for line in _body(f"P2M_S_{p}").split("\n")[1:]:
        line = line.strip()
        m = re.match(r"const T (\w+) = (.*);", line)
        if m:
            env[m.group(1)] = eval(_num(m.group(2)), {}, env)
            continue
        m = re.match(r"M\[(\d+)\] \+= (.*);", line)
        if m:
            body_out[int(m.group(1))] += eval(_num(m.group(2)).replace("S[0]", "q"), {}, dict(env, q=q))
    ref = [q * X ** a * Y ** b * Z ** c / _fact3(a, b, c) for a, b, c in map(_multi, range(ms))]
    np.testing.assert_allclose(body_out, ref, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_m2m_closed_form(p):
    ms = _nterms(p)
    inp = list(np.random.default_rng(p).uniform(-1, 1, ms))
    got = _run(f"M2M_{p}", X, Y, Z, inp, "Ms", "M", ms)
    ref = []
    for n in range(ms):
        a, b, c = _multi(n)
        s = 0.0
        for i in range(a + 1):
            for j in range(b + 1):
                for k in range(c + 1):
                    s += inp[_slot(a - i, b - j, c - k)] * X ** i * Y ** j * Z ** k / _fact3(i, j, k)
        ref.append(s)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_l2l_closed_form(p):
    ls = _nterms(p)
    inp = list(np.random.default_rng(10 + p).uniform(-1, 1, ls))
    got = _run(f"L2L_{p}", X, Y, Z, inp, "Ls", "L", ls)
    ref = []
    for n in range(ls):
        a, b, c = _multi(n)
        rest = p - (a + b + c)
        s = 0.0
        for i in range(rest + 1):
            for j in range(rest + 1 - i):
                for k in range(rest + 1 - i - j):
                    s += inp[_slot(a + i, b + j, c + k)] * X ** i * Y ** j * Z ** k / _fact3(i, j, k)
        ref.append(s)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)
