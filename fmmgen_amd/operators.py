"""Operator-level Python surface of the compiled HIP operators.

The reference reaches its generated operators from Python through the
Cython module that ``generate_code(..., cython=True)`` writes
(fmmgen/writer.py:346-432): one function per generated C function, single
items, arrays updated in place, plus the ``FMMGEN_*`` constants:

    P2M(x, y, z, q, M, order)     M2M(x, y, z, M, Ms, order)
    M2L(x, y, z, M, L, order)     L2L(x, y, z, L, Ls, order)
    L2P(x, y, z, L, F, order)     M2P(x, y, z, M, F, order)
    P2P(x, y, z, S, F)            P2M_<p>(x, y, z, q, M), M2M_<p>(...), ...

``Library`` offers the same names over libfmm_hip.so (fmm_op_batch: the
device operators of the evaluator, every compiled order).  Two forms:

* reference form: ``x, y, z`` numbers and 1-D numpy float64 arrays, one
  item; the output array accumulates the result in place, as the generated
  C functions do;
* batch form: ``x, y, z`` 1-D float64 CUDA tensors of length B and 2-D
  CUDA tensors [B, k]: B independent items in one launch; the output
  tensor accumulates in place.

As in the reference's ``X(..., int order)`` wrappers, an order outside the
library's range [FMMGEN_MINORDER, FMMGEN_MAXORDER) does nothing.  A
library whose outputs (potential / field) are a subset of a compiled
variant's runs that variant and keeps the requested components.  There is
no CPU path: the library has to be built (``python -m fmmgen_amd.build``).
"""
import ctypes
import re

import numpy as np

from . import _lib
from .codegen.layout import Msize, Lsize, sourcesize, outputsize

_ORDERED = ("P2M", "M2M", "M2L", "L2L", "L2P", "M2P")
_PER_ORDER = re.compile(r"^(P2M|M2M|M2L|L2L|L2P|M2P)_(\d+)$")


def _compiled_range(so, outsz):
    mn, mx = ctypes.c_int(), ctypes.c_int()
    code = _lib.load().fmm_variant_orders(so, outsz, ctypes.byref(mn), ctypes.byref(mx))
    return (mn.value, mx.value) if code == _lib.FMM_OK else None


class Library:
    """The operators of one (source order, potential, field) flavour.

    ``min_order`` / ``max_order`` narrow the range to that of a generated
    file (``generate_code(order, ...)`` makes orders start..order-1).
    """

    def __init__(self, source_order=0, potential=True, field=True, min_order=None, max_order=None):
        if not (potential or field):
            raise ValueError("at least one of potential and field")
        self.source_order = so = int(source_order)
        self.potential, self.field = bool(potential), bool(field)
        # the compiled variant with exactly these outputs, else the one with
        # potential and field
        want = outputsize(potential, field)
        self._outsz = rng = None
        for vsz in (want, 4) if want != 4 else (4,):
            rng = _compiled_range(so, vsz)
            if rng is not None:
                self._outsz = vsz
                break
        if self._outsz is None:
            raise _lib.FmmError(2, f"no compiled operator variant for source order {so}, "
                                   f"potential={potential}, field={field}")
        lo, hi = rng
        start = so + (1 if field else 0)
        self.FMMGEN_MINORDER = max(lo, start, min_order if min_order is not None else lo)
        self.FMMGEN_MAXORDER = min(hi, max_order if max_order is not None else hi)
        self.FMMGEN_SOURCEORDER = so
        self.FMMGEN_SOURCESIZE = sourcesize(so)
        self.FMMGEN_OUTPUTSIZE = outputsize(potential, field)
        # output components of the variant kept by this flavour
        if self._outsz == self.FMMGEN_OUTPUTSIZE:
            self._keep = list(range(self._outsz))
        else:
            self._keep = ([0] if potential else []) + ([1, 2, 3] if field else [])

    # ------------------------------------------------------------ sizes

    def _sizes(self, op, order):
        so = self.source_order
        ms, ls, ss, osz = Msize(order, so), Lsize(order, so), sourcesize(so), self.FMMGEN_OUTPUTSIZE
        return {"P2M": (1, ms), "M2M": (ms, ms), "M2L": (ms, ls), "L2L": (ls, ls),
                "L2P": (ls, osz), "M2P": (ms, osz), "P2P": (ss, osz)}[op]

    # ------------------------------------------------------------ calls

    def _run(self, op, order, x, y, z, a, out):
        import torch
        nin, nout = self._sizes(op, order)
        vout = self._outsz if op in ("L2P", "M2P", "P2P") else nout
        batch = isinstance(out, torch.Tensor)
        if batch:
            if not out.is_cuda or out.dtype != torch.float64:
                raise TypeError("batch form: float64 CUDA tensors")
            xyz = torch.stack([torch.as_tensor(v, dtype=torch.float64, device=out.device).reshape(-1)
                               for v in (x, y, z)], 1).contiguous()
            B = xyz.shape[0]
            a2 = a.reshape(B, -1) if a.dim() != 2 else a
            if a2.shape[1] < nin or out.reshape(B, -1).shape[1] < nout:
                raise ValueError(f"{op} at order {order}: inputs need {nin} and outputs {nout} entries per item")
            inp = a2[:, :nin].to(torch.float64).contiguous()
        else:
            out = np.asarray(out)
            if out.dtype != np.float64:
                raise TypeError("reference form: float64 numpy arrays")
            a1 = np.atleast_1d(np.asarray(a, dtype=np.float64))
            if a1.size < nin or out.size < nout:
                raise ValueError(f"{op} at order {order}: inputs need {nin} and outputs {nout} entries")
            B = 1
            xyz = torch.tensor([[float(x), float(y), float(z)]], dtype=torch.float64, device="cuda")
            inp = torch.from_numpy(np.ascontiguousarray(a1[:nin]).reshape(1, nin)).cuda()
        res = torch.zeros(B, vout, dtype=torch.float64, device=xyz.device)
        code = _lib.load().fmm_op_batch(_lib.OPS[op], int(order), self.source_order, self._outsz, B,
                                        ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(inp.data_ptr()), nin,
                                        ctypes.c_void_p(res.data_ptr()), vout,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        _lib.check(code)
        if op in ("L2P", "M2P", "P2P"):
            res = res[:, self._keep]
        if batch:
            o2 = out.reshape(B, -1)
            o2[:, :nout] += res
        else:
            flat = out.reshape(-1)
            flat[:nout] += res.cpu().numpy()[0]
        return None

    def _ordered(self, op, x, y, z, a, out, order):
        order = int(order)
        if order < self.FMMGEN_MINORDER or order >= self.FMMGEN_MAXORDER:
            return None  # the reference's switch wrappers ignore other orders
        return self._run(op, order, x, y, z, a, out)

    def P2M(self, x, y, z, q, M, order):
        """M += the multipole of charge q at displacement (x, y, z) (fmmgen/expansions.py:22-39)."""
        return self._ordered("P2M", x, y, z, q, M, order)

    def M2M(self, x, y, z, M, Ms, order):
        """Ms += M shifted by (x, y, z) (fmmgen/expansions.py:42-78)."""
        return self._ordered("M2M", x, y, z, M, Ms, order)

    def M2L(self, x, y, z, M, L, order):
        """L += the local expansion of M at displacement (x, y, z) (fmmgen/expansions.py:144-172)."""
        return self._ordered("M2L", x, y, z, M, L, order)

    def L2L(self, x, y, z, L, Ls, order):
        """Ls += L shifted by (x, y, z) (fmmgen/expansions.py:175-194)."""
        return self._ordered("L2L", x, y, z, L, Ls, order)

    def L2P(self, x, y, z, L, F, order):
        """F += potential / field of L at (x, y, z) (fmmgen/generator.py:209-254)."""
        return self._ordered("L2P", x, y, z, L, F, order)

    def M2P(self, x, y, z, M, F, order):
        """F += potential / field of M at displacement (x, y, z) (fmmgen/generator.py:174-206)."""
        return self._ordered("M2P", x, y, z, M, F, order)

    def P2P(self, x, y, z, S, F):
        """F += potential / field of source S at displacement (x, y, z) (fmmgen/generator.py:257-288)."""
        return self._run("P2P", self.FMMGEN_MINORDER, x, y, z, S, F)

    def __getattr__(self, name):
        m = _PER_ORDER.match(name)
        if not m:
            raise AttributeError(name)
        op, p = m.group(1), int(m.group(2))
        if p < self.FMMGEN_MINORDER or p >= self.FMMGEN_MAXORDER:
            raise AttributeError(f"{name}: order {p} outside [{self.FMMGEN_MINORDER}, {self.FMMGEN_MAXORDER})")

        def fn(x, y, z, a, out):
            return self._run(op, p, x, y, z, a, out)
        fn.__name__ = name
        return fn

    def functions(self):
        """Name -> callable of every operator, the per-order names included
        (the names the reference's generated module exports)."""
        d = {op: getattr(self, op) for op in _ORDERED + ("P2P",)}
        for p in range(self.FMMGEN_MINORDER, self.FMMGEN_MAXORDER):
            for op in _ORDERED:
                d[f"{op}_{p}"] = getattr(self, f"{op}_{p}")
        d.update(FMMGEN_MINORDER=self.FMMGEN_MINORDER, FMMGEN_MAXORDER=self.FMMGEN_MAXORDER,
                 FMMGEN_SOURCEORDER=self.FMMGEN_SOURCEORDER, FMMGEN_SOURCESIZE=self.FMMGEN_SOURCESIZE,
                 FMMGEN_OUTPUTSIZE=self.FMMGEN_OUTPUTSIZE)
        return d


_default = None


def __getattr__(name):
    # module-level names: the charge library with potential and field
    global _default
    if _default is None:
        _default = Library(0, True, True)
    if name.startswith("FMMGEN_") or name in _ORDERED or name == "P2P" or _PER_ORDER.match(name):
        return getattr(_default, name)
    raise AttributeError(name)
