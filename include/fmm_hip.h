/* C ABI of the MI355X FMM evaluator (libfmm_hip.so).
 *
 * Plain pointers and sizes only.  Pointers named d_* are device memory on
 * the current HIP device; h_* are host memory.  `stream` is a hipStream_t
 * (NULL: the default stream).  Every call returns FMM_OK or an error code;
 * fmm_last_error() describes the last failure of the calling thread.
 * Calls are stream ordered.  The tree build, the exports, the statistics
 * queries and fmm_tree_set_model synchronise; a learned evaluation
 * synchronises once, to launch its far field at the largest selected order
 * (and once more per judge check: the judge's verdict decides a rerun); the
 * first evaluation of a tree at each order synchronises once, uploading
 * that order's M2L index table.  A learned evaluation runs half of its M2L
 * order classes on a second stream of the tree, forked from and joined to
 * the evaluation stream by events: the work stays ordered on that stream.
 *
 * Reference counterparts (paths in the rpep/fmmgen tree):
 *   fmm_tree_build        build_tree(pos, S, n, ncrit, order, theta)
 *                         example/tree.hpp:103, example/tree.cpp:130-233,
 *                         including interact_dehnen_lazy + sort
 *                         (example/tree.cpp:217-222)
 *   fmm_evaluate          Tree::compute_field_fmm(F)
 *                         example/tree.hpp:89, example/tree.cpp:243-271
 *   fmm_evaluate_bh       Tree::compute_field_bh(F)
 *                         example/tree.hpp:90, example/tree.cpp:273-293,
 *                         example/calculate.cpp:266-314
 *   fmm_direct_sampled    Tree::compute_field_exact(F) on a target subset
 *                         example/tree.hpp:91, example/calculate.cpp:252-264
 *   fmm_op_batch          the generated operators P2M/M2M/M2L/L2L/L2P/M2P
 *                         (..., int order) and P2P (...), applied to a
 *                         batch of independent inputs
 *                         example/operators.cpp:1376-1489, fmmgen/writer.py:266-311
 *   fmm_export_*          the Tree members cells, M2L_list, P2P_list, M, L
 *                         example/tree.hpp:78-95, in canonical numbering
 */
#ifndef FMM_HIP_H
#define FMM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM_OK 0
#define FMM_ERR_INVALID_ARGUMENT 1
#define FMM_ERR_UNSUPPORTED 2
#define FMM_ERR_HIP 3
#define FMM_ERR_TREE_TOO_DEEP 4
#define FMM_ERR_OUT_OF_MEMORY 5

/* evaluation modes */
#define FMM_MODE_PLAIN_F64 0    /* every phase in FP64 (the reference arithmetic) */
#define FMM_MODE_MIXED_F32 1    /* FP32 M2L on MFMA, FP32 + compensated P2P */
#define FMM_MODE_MIXED_BF16X3 2 /* BF16x3 split M2L on MFMA, FP32 + compensated P2P */
#define FMM_MODE_LEARNED_F32 3  /* FP32 mixed + learned per-cell order + sampled judge */
#define FMM_MODE_LEARNED_BF16X3 4 /* as 3 with the BF16x3 M2L */

/* phase mask bits for fmm_evaluate_phases */
#define FMM_PHASE_P2M 1
#define FMM_PHASE_M2M 2
#define FMM_PHASE_M2L 4
#define FMM_PHASE_P2P 8
#define FMM_PHASE_L2L 16
#define FMM_PHASE_L2P 32
#define FMM_PHASE_ALL 63

/* operator ids for fmm_op_batch */
#define FMM_OP_P2M 0
#define FMM_OP_M2M 1
#define FMM_OP_M2L 2
#define FMM_OP_L2L 3
#define FMM_OP_L2P 4
#define FMM_OP_M2P 5
#define FMM_OP_P2P 6
#define FMM_OP_P2M_S 7 /* the example's P2M: M2M of the strengths, x = centre - particle */

typedef struct fmm_tree_s* fmm_tree_t;

const char* fmm_last_error(void);
int fmm_abi_version(void);

/* Compiled operator variant for (source_order, output_size): orders
 * [*min_order, *max_order_exclusive). */
int fmm_variant_orders(int source_order, int output_size, int* min_order, int* max_order_exclusive);

/* Builds the octree and the M2L / P2P lists.  d_pos [n*3] interleaved x,y,z;
 * d_S [n*SOURCESIZE] is borrowed (read at every evaluation) and must stay
 * valid for the life of the tree, as build_tree keeps pointers into the
 * caller's arrays (example/tree.cpp:133-137).  theta >= 0 (0: every pair of
 * leaves interacts by P2P, as in the reference). */
int fmm_tree_build(const double* d_pos, const double* d_S, size_t n, size_t ncrit, int order, double theta,
                   int source_order, int output_size, void* stream, fmm_tree_t* out);
int fmm_tree_destroy(fmm_tree_t tree);

/* info must hold FMM_TREE_INFO_LEN entries:
 * info[0..12] = ncells, nleaves, levels, n_m2l, n_p2p, n, order,
 * source_order, P2P particle pairs (self pairs excluded), ncrit,
 * M2L target tiles, M2L (tile, offset) groups, distinct M2L offsets;
 * info[13] = P2P work items (source segments of the target leaves);
 * info[14] = M2L work items (runs of groups of one tile);
 * info[15] = 0 (reserved) */
#define FMM_TREE_INFO_LEN 16
int fmm_tree_info(fmm_tree_t tree, long long* info);

/* Rebinds the borrowed source strengths (same geometry, new sources). */
int fmm_tree_set_sources(fmm_tree_t tree, const double* d_S);

/* Potential / field at every particle, d_F [n*OUTPUTSIZE] in caller order,
 * overwritten (compute_field_fmm zeroes F first). */
int fmm_evaluate(fmm_tree_t tree, int mode, double* d_F, void* stream);
int fmm_evaluate_phases(fmm_tree_t tree, int mode, int phase_mask, double* d_F, void* stream);

/* Stream-ordered phase timing: with timing on, every evaluation records a
 * hipEvent pair around each phase on the evaluation stream.  read returns
 * the totals since the last enable in h_ms[FMM_PHASE_TIMERS] /
 * h_count[FMM_PHASE_TIMERS], indexed 0 prep (gather, zeroing), 1 P2M,
 * 2 M2M, 3 M2L, 4 P2P, 5 L2L, 6 L2P + final sum, 7 learned order
 * selection, 8 judge and reruns, 9 Barnes-Hut walk (synchronises).  No reference counterpart:
 * the example times the whole evaluation with std::chrono
 * (example/main.cpp:154-162). */
#define FMM_PHASE_TIMERS 10
int fmm_timing_enable(fmm_tree_t tree, int enable);
int fmm_timing_read(fmm_tree_t tree, double* h_ms, long long* h_count);

/* Barnes-Hut evaluation in FP64 (Tree::compute_field_bh): multipoles by
 * P2M and M2M at the tree's order, then per particle a walk from the root
 * that opens a child cell when its half-width exceeds theta times the
 * distance from the particle to its centre, applies M2P otherwise, and P2P
 * in the leaves reached.  d_F [n*OUTPUTSIZE] in caller order, overwritten.
 * Stream ordered. */
int fmm_evaluate_bh(fmm_tree_t tree, double* d_F, void* stream);

/* FP64 direct sum at d_targets[nt] over all n sources, d_F [nt*OUTPUTSIZE]. */
int fmm_direct_sampled(const double* d_pos, const double* d_S, size_t n, const long long* d_targets, size_t nt,
                       int source_order, int output_size, double* d_F, void* stream);

/* Item i: op(d_xyz[3i], d_xyz[3i+1], d_xyz[3i+2], d_in + i*in_stride,
 * d_out + i*out_stride, order), accumulating into d_out like the generated
 * operators.  FMM_OP_P2M takes its charge from d_in[i*in_stride]. */
int fmm_op_batch(int op, int order, int source_order, int output_size, size_t n, const double* d_xyz,
                 const double* d_in, int in_stride, double* d_out, int out_stride, void* stream);

/* Canonical exports into host memory.
 * cells: h_geo [ncells*5] = x, y, z, r, rmax; h_meta [ncells*8] = parent,
 *        level, pstart, nleaf, nchild, first_child, octant mask, 0.
 * particles: h_perm [n] sorted slot -> caller index, h_leaf [n] caller
 *        index -> leaf cell.
 * lists: h_m2l [n_m2l*2] (source, target) as the reference stores them,
 *        h_p2p [n_p2p*2] (target, source).
 * expansions of the last evaluation in FP64: h_M [ncells*Msize],
 *        h_L [ncells*Lsize] (NULL skips). */
int fmm_export_cells(fmm_tree_t tree, double* h_geo, long long* h_meta);
int fmm_export_particles(fmm_tree_t tree, long long* h_perm, long long* h_leaf);
int fmm_export_lists(fmm_tree_t tree, long long* h_m2l, long long* h_p2p);
int fmm_export_expansions(fmm_tree_t tree, double* h_M, double* h_L);

/* ---------------------------------------------------------------------
 * Learned per-cell expansion order.  No counterpart in the reference; the
 * fixed-order evaluation above (fmm_evaluate, modes 0-2) stays the method
 * it falls back to and the judge it is checked against.
 *
 * Network: FMM_MLP_FEATURES inputs, two hidden ReLU layers of
 * FMM_MLP_HIDDEN units, FMM_MLP_OUTPUTS outputs; FP32 parameters, flat
 * [W1 (H x F), b1 (H), W2 (H x H), b2 (H), W3 (O x H), b3 (O)], row-major.
 * Output o predicts log10(rms |dF| over a leaf's particles / F_s) of the
 * far field at order FMM_MLP_ORDER0 + o, F_s the field scale
 * (fmm_tree_field_scale).  All pointers are device
 * pointers except where named h_. */
#define FMM_MLP_FEATURES 12
#define FMM_MLP_HIDDEN 64
#define FMM_MLP_OUTPUTS 11
#define FMM_MLP_ORDER0 2
#define FMM_MLP_PARAMS 5707

/* FP32 forward: d_Y [B*O]. */
int fmm_mlp_forward(const float* d_X, int B, const float* d_params, float* d_Y, void* stream);
/* FP16 inference (FP32 accumulation): d_Y [B*O] (NULL skips) and d_order [B]
 * (NULL skips) = the lowest order FMM_MLP_ORDER0 + o with
 * Y[o] + margin <= log10(tol), else max_order, clamped to
 * [min_order, max_order]. */
int fmm_mlp_infer_f16(const float* d_X, int B, const float* d_params, float* d_Y, int* d_order, double tol,
                      double margin, int min_order, int max_order, void* stream);
/* Masked mean-square loss over the outputs with d_mask != 0 and its
 * gradient: d_grad [FMM_MLP_PARAMS], d_loss [1]; d_work holds
 * fmm_mlp_work_floats(B) floats. */
size_t fmm_mlp_work_floats(int B);
int fmm_mlp_grad(const float* d_X, const float* d_T, const float* d_mask, int B, const float* d_params,
                 float* d_grad, float* d_loss, float* d_work, void* stream);
/* fmm_mlp_grad without the workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * training loops).  The workspace is a fixed array inside the library, one
 * per device, of fmm_mlp_work_floats(FMM_MLP_NOALLOC_MAX_BATCH) floats
 * (9.1 MB, held until the process ends); B above FMM_MLP_NOALLOC_MAX_BATCH
 * is FMM_ERR_INVALID_ARGUMENT.  Calls that overlap on one device share the
 * array and race; calls ordered on one stream do not. */
#define FMM_MLP_NOALLOC_MAX_BATCH 8192
int fmm_mlp_grad_noalloc(const float* d_X, const float* d_T, const float* d_mask, int B, const float* d_params,
                         float* d_grad, float* d_loss, void* stream);
/* One bias-corrected Adam step (step >= 1) on the parameters. */
int fmm_mlp_adam(const float* d_grad, float* d_params, float* d_m, float* d_v, float lr, float beta1, float beta2,
                 float eps, int step, void* stream);

/* Evaluation order of the tree (any compiled order: the tree and lists
 * do not depend on it); the learned modes use it as the full order. */
int fmm_tree_set_order(fmm_tree_t tree, int order);
/* Operand form of the M2L translation (no counterpart in the reference,
 * whose M2L is the generated operator, fmmgen/expansions.py:144-172):
 *   FMM_M2L_AUTO (default): harmonic for the mixed and learned modes, full
 *     for plain FP64;
 *   FMM_M2L_FULL: the generated operator's sums over every multi-index;
 *   FMM_M2L_HARMONIC: the same local expansions from folded multipoles and
 *     the rows with n_z <= 1 only ((p+1)^2 instead of C(p+3,3) per side,
 *     exact by the harmonicity of 1/R; results differ from the full form by
 *     rounding), in every mode.
 * Charges only (source order 0): other variants always use the full form. */
#define FMM_M2L_AUTO 0
#define FMM_M2L_FULL 1
#define FMM_M2L_HARMONIC 2
int fmm_tree_set_m2l_form(fmm_tree_t tree, int form);
/* Leaf features d_feat [nleaves*FMM_MLP_FEATURES], leaves in canonical
 * order (computes the field scale first; needs one previous evaluation). */
int fmm_tree_leaf_features(fmm_tree_t tree, float* d_feat, void* stream);
/* Field scale F_s of the learned evaluation, the scale of the network's
 * features and outputs: the rms |F| (field components) of FP64 direct sums
 * at 1024 stratified targets (sorted slots floor((2i + 1) n / 2048)),
 * without the largest 0.1% of them (at least the largest).  Needs one
 * previous evaluation (for the gathered sources): h_scale [1]. */
int fmm_tree_field_scale(fmm_tree_t tree, double* h_scale, void* stream);
/* Root-mean-square |F - Fref| (field components) over each leaf's
 * particles: d_err [nleaves]; d_F, d_Fref [n*OUTPUTSIZE] in caller order. */
int fmm_tree_leaf_error(fmm_tree_t tree, const double* d_F, const double* d_Fref, double* d_err, void* stream);
/* Leaf cell ids in canonical leaf order: h_leaves [nleaves]. */
int fmm_tree_leaves(fmm_tree_t tree, long long* h_leaves);
/* Network (copied, converted to FP16), tolerance and safety margin (in
 * log10 units) of the learned modes.  For charges it also re-forms the M2L
 * tiles with the cell orders the network predicts for the tree's current
 * sources as a key inside each tile class (a tile runs at the largest order
 * of its cells); synchronises the default stream. */
int fmm_tree_set_model(fmm_tree_t tree, const float* d_params, double tol, double margin);
/* Statistics of the last learned evaluation, h_stats [FMM_LEARNED_STATS_LEN]: largest order
 * used, judge targets above tol * rms |F| (a rerun happens when the
 * sample's rel. field error exceeds tol), reruns (0: none, 1: the
 * failing leaves (or all leaves) at the full order, 2: a partial rerun
 * that still missed tol, then every leaf at the full order), judge rel. field
 * error (after the rerun, if any), judge targets, M2L multiply-adds
 * relative to the tree's order, M2L useful multiply-adds (sum over pairs of
 * C(q + 6, 6) at the pair's order), predicted error of the chosen orders
 * relative to F_s, the field scale F_s, M2L multiply-adds at the target
 * cells' own orders (without the tile maximum), M2L useful multiply-adds
 * in the harmonic form (sum over pairs of sum_n (2n+1)(q-n+1)^2).  After
 * fmm_tree_set_model and before the next learned evaluation, entries 5, 6,
 * 9 and 10 describe the orders set_model predicted (fmm_tree_leaf_orders
 * likewise); the others keep the last evaluation's values. */
#define FMM_LEARNED_STATS_LEN 11
int fmm_tree_learned_stats(fmm_tree_t tree, double* h_stats);
/* Per-leaf orders of the last learned evaluation: h_order [nleaves]. */
int fmm_tree_leaf_orders(fmm_tree_t tree, int* h_order);

#ifdef __cplusplus
}
#endif

#endif /* FMM_HIP_H */
