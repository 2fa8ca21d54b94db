// Learned per-cell expansion order (no counterpart in the reference).
//
// Per leaf, twelve features describe how strongly the far field reaches
// the leaf: with rho = (rmax_A + rmax_B) / R, w = sum|S|_B / R^2 and
// v = sqrt(sum |S|^2)_B / R^2 over every M2L pair (B -> A) of the leaf and
// of its ancestors (v is the size of a multipole of random-sign sources),
//   f0..f3   log10(sum w rho^k / F_s), k = 3, 6, 9, 12
//   f4       largest rho
//   f5       log10(1 + number of those pairs)
//   f6       log2(particles in the leaf)
//   f7       log10(sum w / F_s)
//   f8..f11  log10(sum v rho^k / F_s), k = 3, 6, 9, 12
// where F_s is the field scale of the evaluation's judge sample: the rms
// |F| of the FP64 direct sums at the judge's stratified targets without
// the largest 0.1% of them (at least the largest; the field of nearly
// coincident pairs would dominate a plain rms).  The network maps the
// features to the predicted rms far-field error of the leaf per order,
// relative to F_s; the leaf orders are the cheapest set whose predicted
// norm-wise error over all targets meets the tolerance; a cell's order is
// the largest order of the leaves below it, and an M2L tile runs at the
// largest order of its target cells.
#pragma once
#include "fmm_common.h"
#include "mlp.h"

namespace fmm {

// Features of every leaf (canonical leaf order) into t.leaf_feat, from the
// sorted source strengths t.pS (ss per particle) and t.judge_scale.
void leaf_features(Tree& t, int ss, hipStream_t st);
// Root-mean-square |F - Fref| (field components) over the particles of
// each leaf; F and Fref in caller order.
void leaf_error(Tree& t, const double* F, const double* Fref, int outsz, int fo, double* err, hipStream_t st);
// Leaf orders into t.leaf_q from the network outputs Y [nleaves][O]: the
// cheapest orders (M2L multiply-adds) whose predicted norm-wise field
// error relative to F_s, sqrt(sum_l n_l e_l^2 / n) with
// e_l = 10^(Y_l + margin), meets tol (a Lagrange multiplier found by
// bisection); returns that predicted error.
double allocate_orders(Tree& t, const float* Y, double tol, double margin, int qmin, int qmax, hipStream_t st);
// Cell and tile orders from t.leaf_q; returns the largest order (the
// root's), read back to the host.
int orders_from_leaves(Tree& t, hipStream_t st);
// Leaves holding a failing judge target (flag set in t.judge_fail[1 + c])
// move to order qfull.
void raise_failing_leaves(Tree& t, int qfull, hipStream_t st);
// M2L chunks sorted by their tile's order into t.sched.class_chunks;
// start[q] .. start[q + 1] is order q's range (start has
// M2LSchedule::kMaxOrder + 1 entries, host).
void chunk_classes(Tree& t, hipStream_t st, int* start);
// Predicted norm-wise error relative to F_s of the current tile orders: each
// leaf at its effective order, the lowest tile order on its ancestor chain
// (at least its own: a tile runs at the largest order of its cells), so the
// allocation can spend the accuracy the tile maxima add.
double effective_error(Tree& t, const float* Y, double margin, int qmax, hipStream_t st);
// Sum over M2L pairs of C(q + 6, 6) at the order of the pair's target
// cell (what the truncated M2L would cost without the tile maximum).
double m2l_cell_macs(Tree& t, hipStream_t st);
// Every leaf moves to order qfull.
void raise_all_leaves(Tree& t, int qfull, hipStream_t st);
// Sum over M2L pairs of C(q + 6, 6) at their tile's order (the useful
// multiply-adds of the truncated M2L, into *macs_out), divided by the same
// sum at order pref; host result.
double m2l_mac_fraction(Tree& t, int pref, hipStream_t st, double* macs_out = nullptr);

}  // namespace fmm
