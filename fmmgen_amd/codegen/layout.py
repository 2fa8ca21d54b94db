"""Multi-index layout of the Cartesian multipole / local expansion arrays.

Reference: fmmgen/utils.py:25-34 (Nterms), fmmgen/utils.py:65-117
(generate_mappings, graded ordering by grevlex over [z, y, x]),
example/utils.cpp:18-26 (Msize, Lsize).

Slots are ordered by total degree; inside one degree by descending n_x,
then descending n_y.  The closed form used here is

    idx(n) = Nterms(d - 1) + T(d - n_x) + (d - n_x - n_y),   d = |n|,
    T(k)   = k (k + 1) / 2,

which reproduces that ordering (degree 2: xx, xy, xz, yy, yz, zz).  The
order-q expansion is the first Nterms(q) slots of any order-p >= q array
(prefix property), which is what lets one storage stride serve several
per-cell orders.  For source_order so > 0 the multipole array starts at the
first degree-so slot: index = idx(n) - Nterms(so - 1).
"""
from functools import lru_cache


def triangle(n: int) -> int:
    return n * (n + 1) // 2


@lru_cache(maxsize=None)
def Nterms(p: int) -> int:
    """Number of multi-indices with total degree <= p (0 for p < 0)."""
    if p < 0:
        return 0
    return (p + 1) * (p + 2) * (p + 3) // 6


def Msize(order: int, source_order: int) -> int:
    """example/utils.cpp:18-23."""
    if source_order == 0:
        return Nterms(order)
    return Nterms(order) - Nterms(source_order - 1)


def Lsize(order: int, source_order: int) -> int:
    """example/utils.cpp:25-26."""
    return Nterms(order - source_order)


def sourcesize(source_order: int) -> int:
    """FMMGEN_SOURCESIZE, fmmgen/writer.py:321."""
    return Nterms(source_order) - Nterms(source_order - 1)


def outputsize(potential: bool, field: bool) -> int:
    """FMMGEN_OUTPUTSIZE, fmmgen/writer.py:322-328."""
    if potential and not field:
        return 1
    if field and not potential:
        return 3
    if field and potential:
        return 4
    raise ValueError("at least one of potential / field must be requested")


def idx(n) -> int:
    """Array slot of multi-index n = (nx, ny, nz) in a source_order-0 array."""
    nx, ny, nz = n
    d = nx + ny + nz
    return Nterms(d - 1) + triangle(d - nx) + (d - nx - ny)


@lru_cache(maxsize=None)
def monomials(order: int, source_order: int = 0):
    """Multi-indices of degree source_order..order in slot order."""
    out = []
    for d in range(source_order, order + 1):
        for nx in range(d, -1, -1):
            for ny in range(d - nx, -1, -1):
                out.append((nx, ny, d - nx - ny))
    return tuple(out)


def generate_mappings(order: int, source_order: int = 0):
    """Forward and reverse maps between multi-indices and array slots.

    Same contract as fmmgen/utils.py:65-117 (without the sympy symbols
    argument): raises ValueError when order < source_order.
    """
    if order < source_order:
        raise ValueError("source_order must be <= order for meaningful calculations to occur")
    index_dict = {}
    rindex_dict = {}
    for i, n in enumerate(monomials(order, source_order)):
        index_dict[n] = i
        rindex_dict[i] = n
    return index_dict, rindex_dict


def degree(n) -> int:
    return n[0] + n[1] + n[2]


def fact3(n) -> int:
    from math import factorial
    return factorial(n[0]) * factorial(n[1]) * factorial(n[2])


def mac_count_m2l(order: int, source_order: int = 0) -> int:
    """Useful multiply-adds of one M2L pair: #{(n, m): |n|+|m| <= p, |m| >= so}."""
    total = 0
    for dn in range(0, order - source_order + 1):
        for dm in range(source_order, order - dn + 1):
            total += triangle(dn + 1) * triangle(dm + 1)
    return total
