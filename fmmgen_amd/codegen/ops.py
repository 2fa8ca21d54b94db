"""Straight-line programs for the Cartesian FMM operators.

This is the project's own operator generator.  It derives the same
mathematics as the reference's SymPy pipeline and emits each operator as a
straight-line program whose common subexpressions are shared by
construction:

* monomials  t_k = x^k / k!  are built one multiply each from a parent
  monomial (t_k = t_{k - e_a} * (a / k_a));
* derivatives D_k = d^k (1/R) come from the three-term recurrence

      |k| R^2 D_k = -(2|k|-1) sum_a k_a x_a D_{k-e_a}
                    - (|k|-1) sum_a k_a (k_a - 1) D_{k-2e_a}

  and, with ``harmonic_derivs``, D_k = -D_{k-2z+2x} - D_{k-2z+2y} for
  k_z >= 2 (the harmonic identity the reference uses,
  fmmgen/generator.py:82-93, fmmgen/expansions.py:130-141).

Operator mathematics (reference file:line):
  P2M  M_n += (-1)^|n| q x^n / n!          fmmgen/expansions.py:22-39,
                                            fmmgen/generator.py:9-43
  M2M  Ms_n += sum_k M_{n-k} x^k / k!      fmmgen/expansions.py:42-78
  M2L  L_n += sum_m D_{n+m} M_m            fmmgen/expansions.py:144-172,
                                            fmmgen/writer.py:168-182
  L2L  Ls_n += sum_k L_{n+k} x^k / k!      fmmgen/expansions.py:175-194
  L2P  phi = sum_n L_n x^n/n!, F = -grad   fmmgen/expansions.py:197-207,
                                            fmmgen/generator.py:209-254
  M2P  phi = sum_m M_m D_m, F = -grad      fmmgen/generator.py:174-206
  P2P  phi = sum_s S_s D_s (|s| = so), F = -grad
                                            fmmgen/generator.py:257-288

A program is a list of statements:
  ('tmp', name, expr)      T name = expr;
  ('arr', name, size)      T name[size];
  ('set', name, k, expr)   name[k] = expr;
  ('out', name, k, expr)   name[k] += expr;
Expressions are C expression strings in which literal rational constants
appear as ``@num/den@`` and are printed per precision by the printers.
"""
import re
from fractions import Fraction
from math import factorial

from .layout import Nterms, idx, monomials, degree, fact3

AXES = ("x", "y", "z")


def lit(fr) -> str:
    fr = Fraction(fr)
    return f"@{fr.numerator}/{fr.denominator}@"


class Program:
    def __init__(self, name, args, out_name):
        self.name = name
        self.args = args          # list of (name, kind), kind in scalar/in/out
        self.out_name = out_name
        self.stmts = []
        self.flops = 0

    def tmp(self, name, expr, flops=1):
        self.stmts.append(("tmp", name, expr))
        self.flops += flops

    def arr(self, name, size):
        self.stmts.append(("arr", name, size))

    def set(self, name, k, expr, flops=1):
        self.stmts.append(("set", name, k, expr))
        self.flops += flops

    def out(self, k, expr, flops=1):
        self.stmts.append(("out", self.out_name, k, expr))
        self.flops += flops


def _parent(k):
    """Parent multi-index and the axis/exponent that builds t_k from it."""
    kx, ky, kz = k
    if kz > 0:
        return (kx, ky, kz - 1), 2, kz
    if ky > 0:
        return (kx, ky - 1, kz), 1, ky
    return (kx - 1, ky, kz), 0, kx


def _emit_monomials(prog, maxdeg, prefix="t", sign=1, seed=None):
    """t_k = (sign*x)^k / k! for |k| <= maxdeg (t_0 = seed or 1).

    Returns a dict multi-index -> expression string.  Scaled coordinates
    a/j are emitted once per (axis, j).
    """
    names = {}
    scaled = {}
    for a in range(3):
        for j in range(1, maxdeg + 1):
            c = Fraction(sign, j)
            if c == 1:
                scaled[(a, j)] = AXES[a]
            elif c == -1:
                nm = f"{AXES[a]}m1"
                prog.tmp(nm, f"-{AXES[a]}")
                scaled[(a, j)] = nm
            else:
                nm = f"{AXES[a]}{'m' if sign < 0 else ''}{j}"
                prog.tmp(nm, f"{AXES[a]}*{lit(c)}")
                scaled[(a, j)] = nm
    zero = (0, 0, 0)
    names[zero] = seed if seed is not None else None   # None == 1
    for k in monomials(maxdeg, 1):
        par, a, j = _parent(k)
        pn = names[par]
        nm = f"{prefix}{idx(k)}"
        if pn is None:
            names[k] = scaled[(a, j)]
        else:
            prog.tmp(nm, f"{pn}*{scaled[(a, j)]}")
            names[k] = nm
    return names


def _mono(names, k):
    v = names[k]
    return v


def _term(coef_name, mono):
    """coef * mono where mono may be None (== 1)."""
    if mono is None:
        return coef_name
    return f"{coef_name}*{mono}"


def _emit_derivs(prog, maxdeg, harmonic, prefix="D", need=None):
    """D_k = d^k(1/R) at (x, y, z) for |k| <= maxdeg, as temporaries.

    Returns dict multi-index -> temporary name.  Uses R2 / Rinv / Rinv2.
    """
    prog.tmp("R2", "x*x + y*y + z*z", 5)
    prog.tmp("Rinv", "FMM_RSQRT(R2)", 4)
    names = {(0, 0, 0): "Rinv"}
    if maxdeg >= 1:
        prog.tmp("Rinv2", "Rinv*Rinv")
    for d in range(1, maxdeg + 1):
        for k in monomials(d, d):
            kx, ky, kz = k
            nm = f"{prefix}{idx(k)}"
            if harmonic and kz >= 2:
                k1 = (kx + 2, ky, kz - 2)
                k2 = (kx, ky + 2, kz - 2)
                prog.tmp(nm, f"-{names[k1]} - {names[k2]}", 2)
                names[k] = nm
                continue
            terms = []
            for a in range(3):
                if k[a] >= 1:
                    km = list(k)
                    km[a] -= 1
                    c = Fraction(-(2 * d - 1) * k[a], d)
                    terms.append(f"{lit(c)}*{AXES[a]}*{names[tuple(km)]}")
            for a in range(3):
                if k[a] >= 2:
                    km = list(k)
                    km[a] -= 2
                    c = Fraction(-(d - 1) * k[a] * (k[a] - 1), d)
                    terms.append(f"{lit(c)}*{names[tuple(km)]}")
            prog.tmp(nm, f"Rinv2*({' + '.join(terms)})", 2 * len(terms) + 1)
            names[k] = nm
    return names


def _mslot(n, so):
    return idx(n) - Nterms(so - 1)


def _sum(terms):
    return " + ".join(terms) if terms else "0"


# ---------------------------------------------------------------- operators

def gen_P2M(p, so):
    """P2M_p(x, y, z, q, M): charge multipole, outputs over M_dict slots.

    The reference builds it from expansions.M without the source order
    (fmmgen/generator.py:41-42), so for so > 0 it is still the charge
    formula written into the so-offset slots; kept as is for API parity.
    """
    prog = Program(f"P2M_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("q", "scalar"), ("M", "out")], "M")
    names = _emit_monomials(prog, p, prefix="w", sign=-1, seed="q")
    for i, n in enumerate(monomials(p, so)):
        prog.out(i, names[n], 0)
    return prog


def gen_P2M_S(p, so):
    """P2M_S_p(x, y, z, S, M): one particle's multipole as the example forms it.

    example/calculate.cpp:135-161 writes the source strengths into the
    leading slots of a zero multipole and applies M2M with x = centre -
    particle.  This is that composition with the zero slots folded away:
    M_n += sum_s S_s x^{n-s}/(n-s)!,  |s| = so.
    """
    prog = Program(f"P2M_S_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                  ("S", "in"), ("M", "out")], "M")
    names = _emit_monomials(prog, p - so, prefix="t")
    srcs = monomials(so, so)
    for i, n in enumerate(monomials(p, so)):
        terms = []
        for j, s in enumerate(srcs):
            k = (n[0] - s[0], n[1] - s[1], n[2] - s[2])
            if min(k) < 0:
                continue
            terms.append(_term(f"S[{j}]", names[k]))
        if terms:
            prog.out(i, _sum(terms), 2 * len(terms))
    return prog


def gen_M2M(p, so):
    prog = Program(f"M2M_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("M", "in"), ("Ms", "out")], "Ms")
    names = _emit_monomials(prog, p - so, prefix="t")
    for n in monomials(p, so):
        terms = []
        for k in monomials(degree(n) - so, 0):
            nmk = (n[0] - k[0], n[1] - k[1], n[2] - k[2])
            if min(nmk) < 0 or degree(nmk) < so:
                continue
            terms.append(_term(f"M[{_mslot(nmk, so)}]", names[k]))
        prog.out(_mslot(n, so), _sum(terms), 2 * len(terms))
    return prog


def gen_Dblock(p, so, harmonic, prog=None, arr="D"):
    """Internal derivative array D[Msize] (indexed by M slots), as in M2L."""
    own = prog is None
    if own:
        prog = Program(f"Dgen_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                     ("D", "out")], "D")
    names = _emit_derivs(prog, p, harmonic, prefix="d")
    if own:
        for n in monomials(p, so):
            prog.stmts.append(("put", "D", _mslot(n, so), names[n]))
            prog.flops += 0
        return prog
    prog.arr(arr, Nterms(p) - Nterms(so - 1))
    for n in monomials(p, so):
        prog.set(arr, _mslot(n, so), names[n], 0)
    return prog


def gen_M2L(p, so, harmonic=True):
    prog = Program(f"M2L_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("M", "in"), ("L", "out")], "L")
    gen_Dblock(p, so, harmonic, prog=prog)
    for n in monomials(p - so, 0):
        terms = []
        for m in monomials(p - degree(n), so):
            npm = (n[0] + m[0], n[1] + m[1], n[2] + m[2])
            terms.append(f"D[{_mslot(npm, so)}]*M[{_mslot(m, so)}]")
        prog.out(idx(n), _sum(terms), 2 * len(terms))
    return prog


def gen_L2L(p, so):
    prog = Program(f"L2L_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("L", "in"), ("Ls", "out")], "Ls")
    q = p - so
    names = _emit_monomials(prog, q, prefix="t")
    for n in monomials(q, 0):
        terms = []
        for k in monomials(q - degree(n), 0):
            npk = (n[0] + k[0], n[1] + k[1], n[2] + k[2])
            terms.append(_term(f"L[{idx(npk)}]", names[k]))
        prog.out(idx(n), _sum(terms), 2 * len(terms))
    return prog


def _outputs(potential, field):
    outs = []
    if potential:
        outs.append(None)
    if field:
        outs += [0, 1, 2]
    return outs


def gen_L2P(p, so, potential=True, field=True):
    prog = Program(f"L2P_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("L", "in"), ("F", "out")], "F")
    q = p - so
    names = _emit_monomials(prog, q, prefix="t")
    for j, a in enumerate(_outputs(potential, field)):
        terms = []
        for n in monomials(q, 0):
            if a is None:
                terms.append(_term(f"L[{idx(n)}]", names[n]))
            else:
                if n[a] < 1:
                    continue
                nm = list(n)
                nm[a] -= 1
                terms.append(_term(f"L[{idx(n)}]", names[tuple(nm)]))
        if a is None:
            prog.out(j, _sum(terms), 2 * len(terms))
        else:
            prog.out(j, f"-({_sum(terms)})" if terms else "0", 2 * len(terms))
    return prog


def gen_M2P(p, so, potential=True, field=True, harmonic=True):
    prog = Program(f"M2P_{p}", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                                ("M", "in"), ("F", "out")], "F")
    dn = _emit_derivs(prog, p + (1 if field else 0), harmonic, prefix="d")
    for j, a in enumerate(_outputs(potential, field)):
        terms = []
        for m in monomials(p, so):
            k = list(m)
            if a is not None:
                k[a] += 1
            terms.append(f"M[{_mslot(m, so)}]*{dn[tuple(k)]}")
        if a is None:
            prog.out(j, _sum(terms), 2 * len(terms))
        else:
            prog.out(j, f"-({_sum(terms)})", 2 * len(terms))
    return prog


def gen_P2P(so, potential=True, field=True, harmonic=True):
    prog = Program("P2P", [("x", "scalar"), ("y", "scalar"), ("z", "scalar"),
                           ("S", "in"), ("F", "out")], "F")
    if so == 0:
        # phi = S/R, F = S x / R^3 (fmmgen/generator.py:279-287)
        prog.tmp("R2", "x*x + y*y + z*z", 5)
        prog.tmp("Rinv", "FMM_RSQRT(R2)", 4)
        prog.tmp("SR", "S[0]*Rinv")
        if field:
            prog.tmp("SR3", "SR*Rinv*Rinv", 2)
        j = 0
        if potential:
            prog.out(0, "SR")
            j = 1
        if field:
            for a in range(3):
                prog.out(j + a, f"SR3*{AXES[a]}", 2)
        return prog
    dn = _emit_derivs(prog, so + (1 if field else 0), harmonic, prefix="d")
    srcs = monomials(so, so)
    for j, a in enumerate(_outputs(potential, field)):
        terms = []
        for i, s in enumerate(srcs):
            k = list(s)
            if a is not None:
                k[a] += 1
            terms.append(f"S[{i}]*{dn[tuple(k)]}")
        if a is None:
            prog.out(j, _sum(terms), 2 * len(terms))
        else:
            prog.out(j, f"-({_sum(terms)})", 2 * len(terms))
    return prog


_NAME = re.compile(r"[A-Za-z_][A-Za-z0-9_]*")


def prune(prog, lo, hi, name=None):
    """Copy of ``prog`` that keeps only the outputs with slot in [lo, hi) and
    the temporaries they need (backward liveness over the straight-line
    program).  Used to emit one small body per output chunk."""
    keep = []
    live = set()
    for st in reversed(prog.stmts):
        kind = st[0]
        if kind == "out":
            if lo <= st[2] < hi:
                keep.append(st)
                live.update(_NAME.findall(st[3]))
        elif kind == "tmp":
            if st[1] in live:
                keep.append(st)
                live.update(_NAME.findall(st[2]))
        elif kind in ("set", "put"):
            if st[1] in live:
                keep.append(st)
                live.update(_NAME.findall(st[3]))
        elif kind == "arr":
            if st[1] in live:
                keep.append(st)
        else:
            keep.append(st)
    out = Program(name or f"{prog.name}_c{lo}", prog.args, prog.out_name)
    out.stmts = list(reversed(keep))
    return out
