"""Printers for the operator straight-line programs.

Two flavours share one program (``ops.Program``):

* ``HostPrinter``  -- plain C / C++ functions with the reference operator
  signatures, e.g. ``void M2M_4(double x, double y, double z, double * M,
  double * Ms)`` (fmmgen/printers.py:296-313), optional
  ``#pragma omp atomic`` on every output line (fmmgen/printers.py:271-275).
* ``HipPrinter``   -- ``__device__ __forceinline__`` templates for gfx950.
  The scalar type ``T`` is float or double and each array argument is its
  own template parameter, so a kernel can pass a plain pointer, an LDS
  pointer or an accumulator proxy (a wave reduction, a scaled store) and
  the operator body is inlined into the kernel.  Called with double
  arguments and ``double *`` arrays it has exactly the reference
  signature and ``+=`` semantics.
"""
import re
from fractions import Fraction

import numpy as np

_LIT = re.compile(r"@(-?\d+)/(\d+)@")


def _fmt_double(fr: Fraction) -> str:
    v = float(fr)
    if v == int(v) and abs(v) < 1e15:
        s = f"{int(v)}.0"
    else:
        s = repr(v)
    return f"({s})" if v < 0 else s


def _fmt_float(fr: Fraction) -> str:
    v = np.float32(float(fr))
    if float(v) == int(v) and abs(v) < 1e7:
        s = f"{int(v)}.0f"
    else:
        s = f"{float(v):.9g}f"
        if "e" not in s and "." not in s:
            s = s[:-1] + ".0f"
    return f"({s})" if v < 0 else s


class _Base:
    def literal(self, fr: Fraction) -> str:
        raise NotImplementedError

    def expr(self, e: str) -> str:
        def rep(m):
            return self.literal(Fraction(int(m.group(1)), int(m.group(2))))
        e = _LIT.sub(rep, e)
        return e.replace("FMM_RSQRT(", self.rsqrt_name + "(")


class HostPrinter(_Base):
    def __init__(self, precision="double", atomic=False, language="c++"):
        assert precision in ("double", "float")
        self.T = precision
        self.atomic = atomic
        self.language = language
        self.rsqrt_name = "FMMGEN_RSQRT"

    def literal(self, fr):
        return _fmt_double(fr) if self.T == "double" else _fmt_float(fr)

    def signature(self, prog):
        parts = []
        for name, kind in prog.args:
            parts.append(f"{self.T} {name}" if kind == "scalar" else f"{self.T} * {name}")
        return f"void {prog.name}({', '.join(parts)})"

    def body(self, prog):
        T = self.T
        lines = []
        for st in prog.stmts:
            kind = st[0]
            if kind == "tmp":
                lines.append(f"{T} {st[1]} = {self.expr(st[2])};")
            elif kind == "arr":
                lines.append(f"{T} {st[1]}[{st[2]}];")
            elif kind in ("set", "put"):
                lines.append(f"{st[1]}[{st[2]}] = {self.expr(st[3])};")
            elif kind == "out":
                if self.atomic:
                    lines.append("#pragma omp atomic")
                lines.append(f"{st[1]}[{st[2]}] += {self.expr(st[3])};")
            else:
                raise ValueError(kind)
        return "\n".join(lines)

    def function(self, prog):
        return f"{self.signature(prog)} {{\n{self.body(prog)}\n}}\n"

    def prelude(self):
        if self.T == "double":
            return "#define FMMGEN_RSQRT(v) (1.0/sqrt(v))\n"
        return "#define FMMGEN_RSQRT(v) (1.0f/sqrtf(v))\n"


class HipPrinter(_Base):
    """Device templates; literals are ``T(c)`` so one body serves f32/f64."""

    def __init__(self):
        self.rsqrt_name = "fmmgen_rsqrt"

    def literal(self, fr):
        s = _fmt_double(fr)
        return f"T({s})"

    def signature(self, prog, name=None):
        tparams = ["typename T"]
        parts = []
        for aname, kind in prog.args:
            if kind == "scalar":
                parts.append(f"T {aname}")
            else:
                tparams.append(f"typename A_{aname}")
                parts.append(f"A_{aname}&& {aname}")
        return (f"template <{', '.join(tparams)}>\n__device__ __forceinline__ void "
                f"{name or prog.name}({', '.join(parts)})")

    def body(self, prog):
        lines = []
        for st in prog.stmts:
            kind = st[0]
            if kind == "tmp":
                lines.append(f"  const T {st[1]} = {self.expr(st[2])};")
            elif kind == "arr":
                lines.append(f"  T {st[1]}[{st[2]}];")
            elif kind in ("set", "put"):
                lines.append(f"  {st[1]}[{st[2]}] = {self.expr(st[3])};")
            elif kind == "out":
                lines.append(f"  {st[1]}[{st[2]}] += {self.expr(st[3])};")
            else:
                raise ValueError(kind)
        return "\n".join(lines)

    def function(self, prog):
        return f"{self.signature(prog)} {{\n{self.body(prog)}\n}}\n"

    @staticmethod
    def prelude():
        return (
            "#include <hip/hip_runtime.h>\n"
            "__device__ __forceinline__ double fmmgen_rsqrt(double v) { return rsqrt(v); }\n"
            "// FP32: the hardware reciprocal square root (v_rsq_f32, ~1 ulp); inputs\n"
            "// are squared distances, never denormal\n"
            "__device__ __forceinline__ float fmmgen_rsqrt(float v) { return __builtin_amdgcn_rsqf(v); }\n"
        )
