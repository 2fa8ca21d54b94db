"""``generate_code()``: the operator-library entry point.

Same keyword interface and file contract as fmmgen/writer.py:33-344:
per order ``start..order-1`` (``start = source_order + 1`` when the field is
requested, fmmgen/writer.py:121-140) it emits P2M, M2M, M2L (with an
internal ``D`` array), L2L, L2P, M2P, plus ``P2P`` once; then the
``X(..., int order)`` switch wrappers (fmmgen/writer.py:266-311) and the
``FMMGEN_{MINORDER,MAXORDER,SOURCEORDER,SOURCESIZE,OUTPUTSIZE}`` macros
(fmmgen/writer.py:317-328).  As in the reference, an ``order`` outside the
generated range makes a wrapper a silent no-op.

Differences, by design:
* ``gpu=True`` selects the HIP printer (the reference accepts the flag but
  never forwards it, fmmgen/writer.py:107-112).  The result is one
  header-only file ``<name>.h`` of ``__device__`` templates, since device
  bodies must be visible to the kernels that inline them.
* The bodies are straight-line programs whose common subexpressions are
  shared by construction (see ``ops``), so ``CSE`` and ``minpow`` do not
  change them: no ``pow()`` call is ever printed.
* ``cython=True`` writes ``<name>_wrap.py`` instead of the Cython files
  (fmmgen/writer.py:346-432): a Python module with the same names (the
  ``FMMGEN_*`` constants, ``P2M(..., order)`` ... ``P2P``, ``P2M_<p>`` ...)
  bound to the compiled HIP operators (``fmmgen_amd.operators``), for the
  generated order range.  With ``gpu=True`` it raises ``Warning`` as the
  reference does.
* The extra device operators ``P2M_S_<p>`` (the example's P2M through M2M,
  example/calculate.cpp:135-161) and ``Dgen_<p>`` (the derivative block of
  M2L) are emitted with ``gpu=True``; they have no counterpart in the
  reference's generated files.
"""
import logging
import os

from . import ops
from .layout import Nterms, sourcesize, outputsize
from .printers import HostPrinter, HipPrinter

logger = logging.getLogger("fmmgen_amd")


def build_programs(order, source_order=0, potential=True, field=True,
                   harmonic_derivs=False, gpu_extras=False):
    """All operator programs, in the reference's emission order."""
    start = source_order + (1 if field else 0)
    progs = []
    for i in range(start, order):
        progs.append(ops.gen_P2M(i, source_order))
        progs.append(ops.gen_M2M(i, source_order))
        progs.append(ops.gen_M2L(i, source_order, harmonic=harmonic_derivs))
        progs.append(ops.gen_L2L(i, source_order))
        progs.append(ops.gen_L2P(i, source_order, potential, field))
        progs.append(ops.gen_M2P(i, source_order, potential, field, harmonic=harmonic_derivs))
        if gpu_extras:
            progs.append(ops.gen_P2M_S(i, source_order))
            progs.append(ops.gen_Dblock(i, source_order, harmonic_derivs))
        if i == start:
            progs.append(ops.gen_P2P(source_order, potential, field, harmonic=harmonic_derivs))
    if gpu_extras:
        # one derivative block beyond the last order: the field of M2P at
        # the last order needs derivatives of order + 1
        progs.append(ops.gen_Dblock(order, source_order, harmonic_derivs))
    return start, progs


def _macros(start, order, source_order, potential, field):
    return (
        "#pragma once\n"
        f"#define FMMGEN_MINORDER {start}\n"
        f"#define FMMGEN_MAXORDER {order}\n"
        f"#define FMMGEN_SOURCEORDER {source_order}\n"
        f"#define FMMGEN_SOURCESIZE {sourcesize(source_order)}\n"
        f"#define FMMGEN_OUTPUTSIZE {outputsize(potential, field)}\n"
    )


def _base_name(prog_name):
    """'M2M_7' -> ('M2M', 7); 'P2P' -> ('P2P', None)."""
    head, _, tail = prog_name.rpartition("_")
    if head and tail.isdigit():
        return head, int(tail)
    return prog_name, None


def host_source(order, name, precision="double", potential=True, field=True,
                source_order=0, atomic=False, harmonic_derivs=False, language="c++"):
    """(header_text, source_text, opscounts) for the host C / C++ flavour."""
    start, progs = build_programs(order, source_order, potential, field, harmonic_derivs)
    pr = HostPrinter(precision=precision, atomic=atomic, language=language)
    header = _macros(start, order, source_order, potential, field)
    header += pr.prelude()
    body = ""
    ops_counts = []
    wrappers = {}
    for prog in progs:
        header += pr.signature(prog) + ";\n"
        body += pr.function(prog) + "\n"
        ops_counts.append((prog.name, prog.flops))
        base, o = _base_name(prog.name)
        if o is not None:
            wrappers.setdefault(base, (prog, []))[1].append(o)
    for base, (prog, orders) in wrappers.items():
        args = [a for a, _ in prog.args]
        sig = pr.signature(prog).replace(f"{prog.name}(", f"{base}(").rstrip(")") + ", int order)"
        header += sig + ";\n"
        code = sig + " {\nswitch (order) {\n"
        for o in orders:
            code += f"  case {o}:\n    {base}_{o}({', '.join(args)});\n    break;\n"
        code += "  }\n}\n"
        body += code
    inc = '#include "math.h"\n' if language == "c" else "#include<cmath>\n"
    source = f'#include "{name}.h"\n{inc}' + body
    return header, source, ops_counts


CHUNK = 32  # output slots per chunk body (matches the kernels' chunk size)
CHUNKED = ("P2M_S", "M2M", "L2L")


def device_source(order, potential=True, field=True, source_order=0,
                  harmonic_derivs=False, full_bodies_max_order=8):
    """Header-only HIP flavour: __device__ templates plus order wrappers.

    Besides the full operator bodies, P2M_S, M2M and L2L are also emitted as
    one pruned body per chunk of CHUNK output slots (``<op>_<p>_c<k>``), so a
    kernel that computes one chunk per thread compiles only what it uses.
    Full bodies of orders above ``full_bodies_max_order`` are emitted only
    for the small operators (L2P, Dgen, P2P and the chunked ones' pieces).
    """
    start, progs = build_programs(order, source_order, potential, field,
                                  harmonic_derivs, gpu_extras=True)
    pr = HipPrinter()
    text = _macros(start, order, source_order, potential, field)
    text += f"#define FMMGEN_FULL_BODIES_MAXORDER {full_bodies_max_order}\n"
    text += f"#define FMMGEN_DGEN_MAXORDER {order}\n"
    text += f"#define FMMGEN_CHUNK {CHUNK}\n" + pr.prelude()
    text += "namespace fmmgen_ops {\n"
    wrappers = {}
    chunk_fns = {}
    for prog in progs:
        base, o = _base_name(prog.name)
        big = base in ("M2M", "M2L", "L2L", "M2P", "P2M", "P2M_S")
        if base in CHUNKED:
            nout = max(st[2] for st in prog.stmts if st[0] == "out") + 1
            names = []
            for c0 in range(0, nout, CHUNK):
                cp = ops.prune(prog, c0, c0 + CHUNK, name=f"{prog.name}_c{c0 // CHUNK}")
                text += pr.function(cp) + "\n"
                names.append(cp.name)
            chunk_fns[(base, o)] = (prog, names)
        if big and o is not None and o > full_bodies_max_order:
            continue
        text += pr.function(prog) + "\n"
        if o is not None:
            wrappers.setdefault(base, (prog, []))[1].append(o)
    # chunk dispatch: <op>_chunk<P, C>(...)
    for base in CHUNKED:
        items = [(o, v) for (b, o), v in chunk_fns.items() if b == base]
        if not items:
            continue
        prog0 = items[0][1][0]
        fwd = [f"static_cast<A_{a}&&>({a})" if k != "scalar" else a for a, k in prog0.args]
        tparams = ", ".join(["typename T"] + [f"typename A_{a}" for a, k in prog0.args if k != "scalar"])
        params = ", ".join(f"T {a}" if k == "scalar" else f"A_{a}&& {a}" for a, k in prog0.args)
        code = f"template <int P, int C, {tparams}>\n__device__ __forceinline__ void {base}_chunk({params}) {{\n"
        first = True
        for o, (prog, names) in sorted(items):
            for ci, nm in enumerate(names):
                kw = "if constexpr" if first else "else if constexpr"
                code += f"  {kw} (P == {o} && C == {ci}) {nm}<T>({', '.join(fwd)});\n"
                first = False
        code += "}\n"
        text += code
    for base, (prog, orders) in wrappers.items():
        args = [a for a, _ in prog.args]
        fwd = [f"static_cast<A_{a}&&>({a})" if k != "scalar" else a for a, k in prog.args]
        big = base in ("M2M", "M2L", "L2L", "M2P", "P2M", "P2M_S")
        if big:
            orders = [o for o in orders if o <= full_bodies_max_order]
        sig = pr.signature(prog, name=base).rstrip(")") + ", int order)"
        code = sig + " {\n  switch (order) {\n"
        for o in orders:
            code += f"    case {o}: {base}_{o}({', '.join(fwd)}); break;\n"
        code += "    default: break;\n  }\n}\n"
        text += code
        # compile-time order dispatch: only the requested order is instantiated
        tparams = ", ".join(["typename T"] + [f"typename A_{a}" for a, k in prog.args if k != "scalar"])
        params = ", ".join(f"T {a}" if k == "scalar" else f"A_{a}&& {a}" for a, k in prog.args)
        code = f"template <int P, {tparams}>\n__device__ __forceinline__ void {base}_o({params}) {{\n"
        for j, o in enumerate(orders):
            kw = "if constexpr" if j == 0 else "else if constexpr"
            code += f"  {kw} (P == {o}) {base}_{o}<T>({', '.join(fwd)});\n"
        code += "}\n"
        text += code
    text += "}  // namespace fmmgen_ops\n"
    return text


def generate_code(order, name, precision="double", cython=False, CSE=False,
                  harmonic_derivs=False, include_dir=None, src_dir=None,
                  potential=True, field=True, source_order=0, atomic=False,
                  gpu=False, minpow=0, language="c", save_opscounts=None):
    """Write the operator library for orders ``start..order-1``.

    Mirrors fmmgen.generate_code (fmmgen/writer.py:33-50).  Host flavour:
    ``<name>.h`` + ``<name>.c``/``.cpp``.  ``gpu=True``: ``<name>.h`` only,
    holding ``__device__`` templates.
    """
    if language not in ("c", "c++"):
        raise AssertionError("Language must be 'c' or 'c++'")
    if precision not in ("double", "float"):
        raise AssertionError("Precision must be float or double")
    if order <= source_order + (1 if field else 0):
        raise ValueError("order must exceed the first generated order")
    logger.info("Generating FMM operators to order %d (CSE shared by construction: %s, minpow %s)",
                order, CSE, minpow)
    inc_dir = include_dir.rstrip("/") if include_dir else "."
    src = src_dir.rstrip("/") if src_dir else "."
    if gpu:
        text = device_source(order, potential, field, source_order, harmonic_derivs)
        with open(os.path.join(inc_dir, f"{name}.h"), "w") as f:
            f.write(text)
        if cython:
            raise Warning("Cannot write a Cython wrapper for GPU code; skipping")
        return
    header, source, counts = host_source(order, name, precision, potential, field,
                                         source_order, atomic, harmonic_derivs, language)
    with open(os.path.join(inc_dir, f"{name}.h"), "w") as f:
        f.write(header)
    fext = "c" if language == "c" else "cpp"
    with open(os.path.join(src, f"{name}.{fext}"), "w") as f:
        f.write(source)
    if save_opscounts:
        with open(save_opscounts, "w") as f:
            for nm, c in counts:
                f.write(f"{nm},{c}\n")
    if cython:
        start = source_order + (1 if field else 0)
        logger.info("Generating Python operator module: %s_wrap.py", name)
        with open(f"{name}_wrap.py", "w") as f:
            f.write(wrapper_module(name, start, order, source_order, potential, field))


def wrapper_module(name, start, order, source_order, potential, field):
    """Text of ``<name>_wrap.py``: the reference's Cython module names bound
    to the compiled HIP operators for orders start..order-1."""
    return (f'"""Operators {name}: orders {start}..{order - 1}, source order {source_order}, '
            f'potential={potential}, field={field}.\n\n'
            "Written by fmmgen_amd generate_code(..., cython=True); the functions run the\n"
            'compiled HIP operators of libfmm_hip.so (fmmgen_amd.operators)."""\n'
            "from fmmgen_amd.operators import Library as _Library\n\n"
            f"_lib = _Library(source_order={source_order}, potential={potential}, field={field}, "
            f"min_order={start}, max_order={order})\n"
            "globals().update(_lib.functions())\n")
