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

// Device scalars of the learned order (t.lw, doubles): every decision of
// the order selection stays on the device, so an evaluation synchronises
// the host only to launch the far field at its order.
enum LearnedSlot : int {
  kLwScale = 0,     // F_s, the field scale of the judge sample
  kLwTol,           // tolerance
  kLwTol2,          // tolerance of the second allocation
  kLwRealloc,       // 1: the second allocation is made
  kLwTarget,        // multiplier search: target sum n_l e_l^2
  kLwLo, kLwStep,   //   grid of log10 multipliers
  kLwSkip,          //   1: the coarse grid decided, no fine grid
  kLwLogLam,        //   chosen log10 multiplier
  kLwErr,           // scratch sum
  kLwPred, kLwPred2,  // predicted errors of the two allocations
  kLwEff1, kLwEff2,   // their effective errors (tile maxima included)
  kLwEffFinal,        // effective error of the kept allocation
  kLwCount
};

// Features of every leaf (canonical leaf order) into t.leaf_feat, from the
// sorted source strengths t.pS (ss per particle) and F_s (t.lw[kLwScale]).
void leaf_features(Tree& t, int ss, hipStream_t st);
// Root-mean-square |F - Fref| (field components) over the particles of
// each leaf; F and Fref in caller order.
void leaf_error(Tree& t, const double* F, const double* Fref, int outsz, int fo, double* err, hipStream_t st);
// t.lw allocated (zero-filled on the first call).
void learned_scalars(Tree& t);
// F_s into t.lw[kLwScale] from the judge's FP64 direct sums t.judge_Fd
// [J][outsz] (field components fo .. outsz-1).
void field_scale_from_samples(Tree& t, int J, int outsz, int fo, hipStream_t st);
// Leaf orders into q_out from the network outputs Y [nleaves][O]: the
// cheapest orders (M2L multiply-adds) whose predicted norm-wise field
// error relative to F_s, sqrt(sum_l n_l e_l^2 / n) with
// e_l = 10^(Y_l + margin), meets the tolerance lw[tol_slot] (a Lagrange
// multiplier found on two grids); that predicted error into lw[pred_slot].
// gated: nothing is done unless lw[kLwRealloc] is set.
void allocate_orders(Tree& t, const float* Y, int tol_slot, double margin, int qmin, int qmax, int* q_out,
                     int pred_slot, bool gated, hipStream_t st);
// Cell orders (the largest leaf order below) and supertile orders (the
// largest cell order of the supertile) from leaf orders.
void orders_from_leaves(Tree& t, const int* leaf_q, int* cell_q, signed char* tile_order, hipStream_t st);
// The learned evaluation's orders into t.leaf_q / t.cell_q / t.tile_order:
// an allocation for the tolerance, then (refine) a second one for the
// slack the tile maxima leave, kept when its effective error meets the
// tolerance.  Stream ordered, no host synchronisation.
void select_orders(Tree& t, const float* Y, double margin, int qmin, int qmax, bool refine, hipStream_t st);
// Leaves holding a failing judge target (flag set in t.judge_fail[1 + c])
// move to order qfull.
void raise_failing_leaves(Tree& t, int qfull, hipStream_t st);
// Work items sorted by their supertile's order into t.sched.class_chunks,
// the start of every order's range into t.sched.class_start (device).
void chunk_classes_async(Tree& t, hipStream_t st);
// Synchronises: the largest order (the root's, t.cell_q[0]) and, from
// chunk_classes_async, start[q] .. start[q + 1] = order q's range of work
// items (start: M2LSchedule::kMaxOrder + 1 entries, host).
int read_orders(Tree& t, hipStream_t st, int* start);
// Predicted norm-wise error relative to F_s of the given tile orders: each
// leaf at its effective order, the lowest tile order on its ancestor chain
// (qmax above the first tile), into lw[slot].
void effective_error(Tree& t, const float* Y, double margin, int qmax, const signed char* tile_order, int slot,
                     hipStream_t st);
// Sum over M2L pairs of C(q + 6, 6) at the order of the pair's target
// cell (what the truncated M2L would cost without the supertile maximum).
double m2l_cell_macs(Tree& t, hipStream_t st);
// Every leaf moves to order qfull.
void raise_all_leaves(Tree& t, int qfull, hipStream_t st);
// Sum over M2L pairs of C(q + 6, 6) at their supertile's order (the useful
// multiply-adds of the truncated M2L, into *macs_out), divided by the same
// sum at order pref; host result.
double m2l_mac_fraction(Tree& t, int pref, hipStream_t st, double* macs_out = nullptr, double* hmacs_out = nullptr);

}  // namespace fmm
