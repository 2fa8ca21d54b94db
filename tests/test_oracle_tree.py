"""Pins the oracle's tree and interaction lists.

* A literal pure-Python restatement of build_tree / split_cell / add_child
  (example/tree.cpp:91-233) and interact_dehnen_lazy
  (example/calculate.cpp:81-133) must give the C oracle's tree -- same cell
  numbering (creation order), centres bit for bit, counts and leaf lists --
  and the same lists as multisets.
* Independent invariants: every particle in exactly one leaf, leaves hold
  fewer than ncrit particles, every ordered pair of leaves is covered by
  exactly one interaction (P2P of the pair, or M2L between ancestors), M2L
  pairs satisfy the opening criterion and P2P pairs do not.
"""
import math

import numpy as np
import pytest

from fmmgen_amd.codegen.layout import Msize, Lsize
from oracle import oracle as O
from tests import util


class PyCell:
    __slots__ = ("x", "y", "z", "r", "rmax", "parent", "level", "nleaf", "nchild", "child", "leaf")

    def __init__(self, x, y, z, r, parent, level):
        self.x, self.y, self.z, self.r = x, y, z, r
        self.rmax = math.sqrt(3 * r * r)
        self.parent, self.level = parent, level
        self.nleaf, self.nchild = 0, 0
        self.child = [0] * 8
        self.leaf = []


def py_build_tree(pos, ncrit):
    n = len(pos)
    xa = ya = za = 0.0
    for i in range(n):
        xa += pos[i][0]
        ya += pos[i][1]
        za += pos[i][2]
    xa /= n
    ya /= n
    za /= n
    xm = ym = zm = 0.0
    for i in range(n):
        xm = max(xm, abs(pos[i][0] - xa)) if abs(pos[i][0] - xa) > xm else xm
        ym = max(ym, abs(pos[i][1] - ya)) if abs(pos[i][1] - ya) > ym else ym
        zm = max(zm, abs(pos[i][2] - za)) if abs(pos[i][2] - za) > zm else zm
    r = (xm if xm > zm else zm) if xm > ym else (ym if ym > zm else zm)
    r *= 1.001
    cells = [PyCell(xa, ya, za, r, 0, 0)]

    def octant(i, c):
        return int(pos[i][0] > c.x) + (int(pos[i][1] > c.y) << 1) + (int(pos[i][2] > c.z) << 2)

    def add_child(o, p):
        rr = cells[p].r / 2.0
        c = PyCell(cells[p].x + rr * ((o & 1) * 2 - 1), cells[p].y + rr * ((o & 2) - 1),
                   cells[p].z + rr * ((o & 4) // 2 - 1), rr, p, cells[p].level + 1)
        cells.append(c)
        cells[p].child[o] = len(cells) - 1
        cells[p].nchild |= 1 << o

    def split(p):
        for l in list(cells[p].leaf):
            o = octant(l, cells[p])
            if not cells[p].nchild & (1 << o):
                add_child(o, p)
            c = cells[p].child[o]
            cells[c].leaf.append(l)
            cells[c].nleaf += 1
            if cells[c].nleaf >= ncrit:
                split(c)

    for i in range(n):
        cur = 0
        while cells[cur].nleaf >= ncrit:
            cells[cur].nleaf += 1
            o = octant(i, cells[cur])
            if not cells[cur].nchild & (1 << o):
                add_child(o, cur)
            cur = cells[cur].child[o]
        cells[cur].leaf.append(i)
        cells[cur].nleaf += 1
        if cells[cur].nleaf >= ncrit:
            split(cur)
    return cells


def py_lists(cells, theta, ncrit):
    m2l, p2p = [], []

    def rec(A, B):
        a, b = cells[A], cells[B]
        dx, dy, dz = a.x - b.x, a.y - b.y, a.z - b.z
        R = math.sqrt(dx * dx + dy * dy + dz * dz)
        if R * theta > a.rmax + b.rmax:
            m2l.append((B, A))
        elif a.nchild == 0 and b.nchild == 0:
            (m2l if b.nleaf >= ncrit else p2p).append((B, A) if b.nleaf >= ncrit else (A, B))
        elif b.nchild == 0 or (a.rmax >= b.rmax and a.nchild != 0):
            for o in range(8):
                if a.nchild & (1 << o):
                    rec(a.child[o], B)
        else:
            for o in range(8):
                if b.nchild & (1 << o):
                    rec(A, b.child[o])

    rec(0, 0)
    return m2l, p2p


CASES = [("uniform", 800, 8, 0.5), ("plummer", 1500, 16, 0.4), ("blobs", 1200, 32, 0.5), ("uniform", 300, 4, 0.3)]


def _pos(kind, n):
    if kind == "uniform":
        return util.uniform(n, 1)[0]
    if kind == "plummer":
        return util.plummer(n, 2)
    return util.blobs(n, 3)


@pytest.mark.parametrize("kind,n,ncrit,theta", CASES)
def test_oracle_matches_literal_restatement(kind, n, ncrit, theta):
    pos = _pos(kind, n)
    q = np.ones((n, 1))
    T = O.OracleTree(pos, q, ncrit, 3, theta)
    geo, meta = T.cells()
    cells = py_build_tree([tuple(p) for p in pos], ncrit)
    assert len(cells) == geo.shape[0]
    for c, pc in enumerate(cells):
        assert (geo[c, 0], geo[c, 1], geo[c, 2], geo[c, 3], geo[c, 4]) == (pc.x, pc.y, pc.z, pc.r, pc.rmax)
        assert (meta[c, 0], meta[c, 1], meta[c, 2], meta[c, 3]) == (pc.parent, pc.level, pc.nleaf, pc.nchild)
    leaf_of = T.leaf_of()
    for c, pc in enumerate(cells):
        if pc.nchild == 0:
            assert list(np.nonzero(leaf_of == c)[0]) == pc.leaf     # increasing particle order
    m2l, p2p = T.lists()
    pm2l, pp2p = py_lists(cells, theta, ncrit)
    assert sorted(map(tuple, m2l.tolist())) == sorted(pm2l)
    assert sorted(map(tuple, p2p.tolist())) == sorted(pp2p)
    # the M2L list is sorted by source (example/tree.cpp:218-222)
    assert np.all(np.diff(m2l[:, 0]) >= 0)


@pytest.mark.parametrize("kind,n,ncrit,theta", CASES)
def test_tree_and_list_invariants(kind, n, ncrit, theta):
    pos = _pos(kind, n)
    T = O.OracleTree(pos, np.ones((n, 1)), ncrit, 3, theta)
    geo, meta = T.cells()
    nc = geo.shape[0]
    leaf_of = T.leaf_of()
    assert np.all(leaf_of >= 0)
    leaves = [c for c in range(nc) if meta[c, 3] == 0]
    for c in leaves:
        assert meta[c, 2] < ncrit
        assert np.count_nonzero(leaf_of == c) == meta[c, 2]
    for c in range(nc):
        if meta[c, 3]:
            assert meta[c, 2] >= ncrit
            assert meta[c, 2] == sum(meta[k, 2] for k in meta[c, 4:12] if k >= 0)
    # particles lie inside their leaf
    for i in range(n):
        c = leaf_of[i]
        assert np.all(np.abs(pos[i] - geo[c, :3]) <= geo[c, 3] * (1 + 1e-12))
    # ancestors
    anc = {}
    for c in range(nc):
        a, chain = c, [c]
        while a != 0:
            a = meta[a, 0]
            chain.append(a)
        anc[c] = set(chain)
    m2l, p2p = T.lists()
    cover = {}
    for A, B in p2p:
        cover[(A, B)] = cover.get((A, B), 0) + 1

    def dist(A, B):
        # the reference's operation order: exact ties of the criterion occur
        # on the dyadic lattice of cell centres, so rounding decides them
        dx, dy, dz = (float(geo[A, k] - geo[B, k]) for k in range(3))
        return math.sqrt(dx * dx + dy * dy + dz * dz)

    for B, A in m2l:
        assert dist(A, B) * theta > geo[A, 4] + geo[B, 4]
    m2l_set = {(int(B), int(A)) for B, A in m2l}
    for A in leaves:
        for B in leaves:
            k = cover.get((A, B), 0)
            k += sum(1 for a in anc[A] for b in anc[B] if (b, a) in m2l_set)
            assert k == 1, (A, B, k)
    for A, B in p2p:
        assert not (dist(A, B) * theta > geo[A, 4] + geo[B, 4])


def test_fmm_error_falls_with_order_and_charge_is_conserved():
    pos, q = util.uniform(4000, 9)
    tg = np.arange(0, 4000, 13)
    Fd = O.direct_targets(pos, q, tg)
    errs = []
    for p in (1, 3, 5, 7):
        T = O.OracleTree(pos, q, 32, p, 0.5)
        F = T.compute_field_fmm()
        errs.append(util.rel_norm(F[tg, 1:], Fd[:, 1:]))
        M, _ = T.ML(Msize(p, 0), Lsize(p, 0))
        geo, meta = T.cells()
        leaves = meta[:, 3] == 0
        assert abs(M[0, 0] - M[leaves, 0].sum()) <= 1e-10 * np.abs(q).sum()   # M_sanity_check
        assert abs(M[0, 0] - q.sum()) <= 1e-10 * np.abs(q).sum()
    assert all(errs[i + 1] < errs[i] / 3 for i in range(len(errs) - 1))
    assert errs[-1] < 1e-4


@pytest.mark.parametrize("so,outsz,orders", [(1, 3, (2, 4, 6, 8)), (2, 4, (3, 5, 7)), (0, 1, (0, 2, 4, 6)),
                                             (0, 3, (1, 3, 5, 7))])
def test_fmm_variants_converge_to_direct_sum(so, outsz, orders):
    """The oracle FMM of the other operator variants (dipole and quadrupole
    sources, potential only, field only) approaches the direct sum as the
    order grows."""
    from fmmgen_amd.codegen.layout import sourcesize
    pos, _ = util.uniform(3000, 19)
    S = np.random.default_rng(20).uniform(-1, 1, (3000, sourcesize(so)))
    tg = np.arange(0, 3000, 11)
    Fd = O.direct_targets(pos, S, tg, source_order=so, outputsize=outsz)
    errs = [util.rel_norm(O.OracleTree(pos, S, 32, p, 0.5, source_order=so, outputsize=outsz)
                          .compute_field_fmm()[tg], Fd) for p in orders]
    assert all(b < a / 3 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-3, errs
