"""Pins the oracle's tree and list rules with known answers worked out by
hand and with consequences of the rules, and shows that these checks reject
each of eight planted mistakes (the oracle built with ORACLE_MUTANT = k).

The rules (example/tree.cpp:91-233, example/calculate.cpp:81-133):
  * root centre = mean of the positions, half-width = 1.001 * the largest
    per-axis |x - mean| (tree.cpp:145-188);
  * a cell holding ncrit particles is split; children only for non-empty
    octants; octant bit 0/1/2 = x/y/z strictly greater than the centre;
    child centre = parent +- r/2, children numbered in creation order
    (tree.cpp:91-128, 189-204);
  * dual traversal from (root, root): M2L (B, A) when R theta > rmax_A +
    rmax_B; P2P (A, B) for two leaves; otherwise split A when B is a leaf or
    (rmax_A >= rmax_B and A is not a leaf), else split B (calculate.cpp:81-133).

Consequences checked on realistic trees: an M2L source finer than its
target only occurs for a leaf target (the rule splits A first on ties); a
target two or more levels finer than its source only for a leaf source;
every ordered pair of leaves covered exactly once; the criterion holds for
M2L pairs and fails for P2P pairs; every particle inside its leaf; leaves
below ncrit, internal cells at or above it; the M2L list sorted by source
(tree.cpp:218-222).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import util

R1001 = 3.0 * 1.001            # 3.0029999999999997: the (a) root half-width in FP64


def _tree(pos, ncrit, theta, library):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    return O.OracleTree(pos, np.ones((len(pos), 1)), ncrit, 3, theta, library=library)


def _pairs(a):
    return sorted(map(tuple, np.asarray(a).tolist()))


def known_answer_failures(library):
    f = []
    # (a) root geometry of four particles: mean (0.25, 0.5, 1.0), largest
    # deviation 3.0 (the z of (0, 0, 4) from 1.0), half-width 3.0 * 1.001
    T = _tree([(0, 0, 0), (1, 0, 0), (0, 2, 0), (0, 0, 4)], 8, 0.5, library)
    geo, meta = T.cells()
    if T.ncells != 1 or tuple(geo[0, :4]) != (0.25, 0.5, 1.0, R1001):
        f.append(("root", T.ncells, tuple(geo[0, :4])))
    if geo[0, 4] != math.sqrt(3 * R1001 * R1001):
        f.append(("rmax", geo[0, 4]))
    # (b) eight particles on the corners of a cube: ncrit = 8 splits the
    # root into eight one-particle leaves, ncrit = 9 keeps it a leaf
    corners = [(x, y, z) for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)]
    if _tree(corners, 8, 0.5, library).ncells != 9:
        f.append(("split at ncrit",))
    if _tree(corners, 9, 0.5, library).ncells != 1:
        f.append(("leaf below ncrit",))
    # (c) a particle exactly at the centre goes to octant 0 (strict >):
    # mean (0, 0, 0), half-width 1.001; particles 0 and 2 in octant 0 (cell
    # 1, created first), particle 1 in octant 1 (cell 2)
    T = _tree([(-1, 0, 0), (1, 0, 0), (0, 0, 0)], 3, 0.5, library)
    geo, meta = T.cells()
    h = 1.001 / 2.0
    want_geo = [(0.0, 0.0, 0.0, 1.001), (-h, -h, -h, h), (h, -h, -h, h)]
    if T.ncells != 3 or [tuple(g[:4]) for g in geo] != want_geo:
        f.append(("octant tie geometry", [tuple(g[:4]) for g in geo]))
    elif list(T.leaf_of()) != [1, 2, 1] or list(meta[0, 4:12]) != [1, 2, -1, -1, -1, -1, -1, -1]:
        f.append(("octant tie membership", list(T.leaf_of()), list(meta[0, 4:12])))
    # (d) creation order: the octant-7 particle comes first, so cell 1 is
    # the octant-7 child and cell 2 the octant-0 child
    T = _tree([(1, 1, 1), (-1, -1, -1)], 2, 0.5, library)
    geo, meta = T.cells()
    if T.ncells != 3 or meta[0, 4 + 7] != 1 or meta[0, 4] != 2 or list(meta[1:, 1]) != [1, 1]:
        f.append(("creation order", T.ncells, list(meta[0, 4:12])))
    # (e) two particles (0, 0, 0) and (1, 1, 1), ncrit 2: root centre 0.5,
    # half-width 0.5005; children 1 (octant 0) and 2 (octant 7) with
    # half-width 0.25025.  (root, root) splits A (equal rmax, A internal):
    # (1, 0) has R = 0.25025 sqrt(3) = 0.4335 against rmax_1 + rmax_0 =
    # 1.3003: theta = 3.5 accepts it (1.517), theta = 0.5 does not and
    # splits the root as B, giving the four leaf pairs.
    T = _tree([(0, 0, 0), (1, 1, 1)], 2, 3.5, library)
    m2l, p2p = T.lists()
    if _pairs(m2l) != [(0, 1), (0, 2)] or len(p2p) != 0:
        f.append(("lists theta 3.5", _pairs(m2l), _pairs(p2p)))
    T = _tree([(0, 0, 0), (1, 1, 1)], 2, 0.5, library)
    m2l, p2p = T.lists()
    if len(m2l) != 0 or _pairs(p2p) != [(1, 1), (1, 2), (2, 1), (2, 2)]:
        f.append(("lists theta 0.5", _pairs(m2l), _pairs(p2p)))
    # (f) a leaf beside an internal cell of the same size: (0,0,0) alone in
    # octant 0, two particles in octant 7 (ncrit 2).  (leaf, internal) must
    # split the internal cell, so the octant-0 leaf meets both octant-7
    # grandchildren by P2P
    T = _tree([(0, 0, 0), (0.9, 0.9, 0.9), (1.0, 1.0, 1.0)], 2, 0.5, library)
    geo, meta = T.cells()
    m2l, p2p = T.lists()
    leaves = [c for c in range(T.ncells) if meta[c, 3] == 0]
    if len(leaves) != 3 or sorted(map(tuple, p2p.tolist())) != sorted((a, b) for a in leaves for b in leaves) \
            or len(m2l) != 0:
        f.append(("leaf beside internal", leaves, _pairs(m2l), _pairs(p2p)))
    return f


RULE_CASES = [("uniform", 2000, 16, 0.5), ("plummer", 3000, 16, 0.4), ("blobs", 3000, 32, 0.5),
              ("uniform", 1500, 8, 0.3)]


def _positions(kind, n):
    if kind == "uniform":
        return util.uniform(n, 5)[0]
    if kind == "plummer":
        return util.plummer(n, 6)
    return util.blobs(n, 7)


def rule_failures(library, kind, n, ncrit, theta):
    f = []
    pos = _positions(kind, n)
    T = O.OracleTree(pos, np.ones((n, 1)), ncrit, 3, theta, library=library)
    geo, meta = T.cells()
    nc = geo.shape[0]
    level = meta[:, 1]
    is_leaf = meta[:, 3] == 0
    leaf_of = T.leaf_of()
    # tree: counts and boxes
    for c in range(nc):
        if is_leaf[c] and meta[c, 2] >= ncrit:
            f.append(("leaf at ncrit", c))
        if not is_leaf[c] and meta[c, 2] < ncrit:
            f.append(("internal below ncrit", c))
    if np.any(leaf_of < 0):
        f.append(("particle without leaf",))
        return f
    inside = np.abs(pos - geo[leaf_of, :3]) <= geo[leaf_of, 3:4] * (1 + 1e-12)
    if not np.all(inside):
        f.append(("particle outside its leaf", int(np.argmin(inside.all(axis=1)))))
    m2l, p2p = T.lists()
    if len(m2l) and np.any(np.diff(m2l[:, 0]) < 0):
        f.append(("M2L list not sorted by source",))

    def dist(A, B):
        dx, dy, dz = (float(geo[A, k] - geo[B, k]) for k in range(3))
        return math.sqrt(dx * dx + dy * dy + dz * dz)

    for B, A in m2l:
        if level[B] > level[A] and not is_leaf[A]:
            f.append(("source finer than an internal target", int(B), int(A)))
            break
        if level[A] >= level[B] + 2 and not is_leaf[B]:
            f.append(("target two levels finer than an internal source", int(B), int(A)))
            break
        if not dist(A, B) * theta > geo[A, 4] + geo[B, 4]:
            f.append(("M2L pair fails the criterion", int(B), int(A)))
            break
    for A, B in p2p:
        if not (is_leaf[A] and is_leaf[B]) or dist(A, B) * theta > geo[A, 4] + geo[B, 4]:
            f.append(("bad P2P pair", int(A), int(B)))
            break
    # every ordered pair of leaves covered exactly once
    anc = [None] * nc
    for c in range(nc):
        a, chain = c, [c]
        while a != 0:
            a = int(meta[a, 0])
            chain.append(a)
        anc[c] = chain
    leaves = np.nonzero(is_leaf)[0]
    cover = {}
    for A, B in p2p:
        cover[(int(A), int(B))] = cover.get((int(A), int(B)), 0) + 1
    m2l_set = {}
    for B, A in m2l:
        m2l_set[(int(B), int(A))] = m2l_set.get((int(B), int(A)), 0) + 1
    bad = 0
    for A in leaves:
        for B in leaves:
            k = cover.get((int(A), int(B)), 0)
            k += sum(m2l_set.get((b, a), 0) for a in anc[A] for b in anc[B])
            bad += k != 1
    if bad:
        f.append(("leaf pairs not covered exactly once", bad))
    return f


def test_oracle_meets_known_answers():
    assert known_answer_failures(O.lib(0, 4)) == []


@pytest.mark.parametrize("kind,n,ncrit,theta", RULE_CASES)
def test_oracle_meets_rule_consequences(kind, n, ncrit, theta):
    assert rule_failures(O.lib(0, 4), kind, n, ncrit, theta) == []


@pytest.mark.parametrize("k", range(1, 9))
def test_checks_reject_planted_mistake(k):
    lib = O.mutant_lib(k)
    found = known_answer_failures(lib)
    for case in RULE_CASES[:2]:
        found += rule_failures(lib, *case)
    assert found, f"mutant {k} passes every tree and list check"
