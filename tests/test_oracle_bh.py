"""Pins of the oracle's Barnes-Hut evaluation (Tree::compute_field_bh,
example/tree.cpp:273-293; evaluate_M2P_and_P2P,
example/calculate.cpp:266-314) -- CPU only.

* theta = 0 opens every cell: the result is the direct sum;
* at theta = 0.2 the error against the direct sum falls geometrically with
  the order (about 3x per order: the farthest particle of an applied cell
  sits within sqrt(3) r_c of its centre, at least r_c / theta away), to
  below 1e-9 at order 12;
* known answers on a two-particle tree: the opening criterion
  r_child > theta |x_i - centre_child| switches a cell from P2P (exact) to
  M2P of its multipole at x_i - centre_child exactly at theta = r_c / d,
  with the half-width r (not rmax) and a strict comparison; the M2P value
  is the pinned operators' (oracle_op_M2M, oracle_op_M2P);
* each of three planted mistakes (criterion on rmax, >= instead of >,
  displacement centre - particle) fails one of these checks.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import util


def _direct(pos, q):
    return O.direct_targets(pos, q, np.arange(pos.shape[0]))


def _two_particle_failures(library):
    """Known answers of the two-particle tree; a list of failed checks."""
    fails = []
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.5, 0.25]])
    q = np.array([[1.5], [-0.75]])
    order = 1
    base = O.OracleTree(pos, q, 2, order, 0.5, library=library)
    geo, meta = base.cells()
    leaf = base.leaf_of()
    # the two leaves: one particle each, children of the root
    if base.ncells != 3 or set(leaf) != {1, 2}:
        return ["tree shape"]
    exact = _direct(pos, q)
    c_of = {i: int(leaf[i]) for i in range(2)}
    for i in range(2):
        j = 1 - i
        cj, ci = c_of[j], c_of[i]
        d = math.sqrt(sum((pos[i, k] - geo[cj, k]) ** 2 for k in range(3)))
        rc = geo[cj, 3]
        theta_star = rc / d
        # own cell stays open below 1/sqrt(3); other cell switches at theta_star
        assert theta_star < 0.5
        # M2P value of the other particle's cell at order 1
        M = np.zeros(4)
        S = np.zeros(4)
        S[0] = q[j, 0]
        library.oracle_op_M2M(geo[cj, 0] - pos[j, 0], geo[cj, 1] - pos[j, 1], geo[cj, 2] - pos[j, 2],
                              O._p(S), O._p(M), order)
        m2p = np.zeros(4)
        library.oracle_op_M2P(pos[i, 0] - geo[cj, 0], pos[i, 1] - geo[cj, 1], pos[i, 2] - geo[cj, 2],
                              O._p(M), O._p(m2p), order)
        # a theta with theta * d == r_c exactly: the cell is not opened
        th = theta_star
        for _ in range(8):
            if th * d == rc:
                break
            th = math.nextafter(th, 1.0 if th * d < rc else 0.0)
        cases = [(0.97 * theta_star, exact[i]), (1.03 * theta_star, m2p)]
        if th * d == rc:
            cases.append((th, m2p))
        for theta, want in cases:
            F = O.OracleTree(pos, q, 2, order, theta, library=library).compute_field_bh()[i]
            if not np.allclose(F, want, rtol=1e-13, atol=0):
                fails.append(f"target {i} theta={theta:.6f}")
        if np.allclose(m2p, exact[i], rtol=1e-3):
            fails.append("order-1 M2P indistinguishable from P2P")
    return fails


def _theta0_failures(library):
    pos, q = util.uniform(1500, 5)
    F = O.OracleTree(pos, q, 16, 4, 0.0, library=library).compute_field_bh()
    return [] if util.rel_norm(F, _direct(pos, q)) < 1e-14 else ["theta=0 is not the direct sum"]


def _convergence_failures(library, orders=(2, 5, 8, 12)):
    pos = util.blobs(3000, 6)
    q = np.random.default_rng(7).uniform(-1, 1, (3000, 1))
    Fd = _direct(pos, q)
    errs = [util.rel_norm(O.OracleTree(pos, q, 32, p, 0.2, library=library).compute_field_bh()[:, 1:],
                          Fd[:, 1:]) for p in orders]
    fails = []
    if not all(b < 0.1 * a for a, b in zip(errs, errs[1:])):
        fails.append(f"error does not fall with the order: {errs}")
    if errs[-1] > 1e-9:
        fails.append(f"order {orders[-1]} error {errs[-1]:.2e}")
    return fails


def test_bh_theta_zero_is_direct_sum():
    assert _theta0_failures(O.lib(0, 4)) == []


def test_bh_converges_with_order():
    assert _convergence_failures(O.lib(0, 4)) == []


def test_bh_two_particle_known_answers():
    assert _two_particle_failures(O.lib(0, 4)) == []


@pytest.mark.parametrize("k", [9, 10, 11])
def test_bh_checks_reject_planted_mistake(k):
    L = O.mutant_lib(k)
    fails = _two_particle_failures(L) + _theta0_failures(L) + _convergence_failures(L, orders=(2, 5, 8))
    assert fails, f"mutant {k} passes every Barnes-Hut check"
