// extern "C" entry points of libfmm_hip.so (declared in include/fmm_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/fmm_hip.h"
#include "evaluate.h"
#include "learned.h"
#include "mlp.h"
#include "fmm_common.h"

struct fmm_tree_s {
  fmm::Tree t;
  const fmm::Variant* v = nullptr;
};

namespace fmm {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }

PhaseTimer::PhaseTimer(Tree& tree, int ph, hipStream_t s) : t(tree), phase(ph), st(s) {
  if (!t.timing) return;
  auto take = [&]() {
    hipEvent_t e;
    if (!t.event_pool.empty()) {
      e = t.event_pool.back();
      t.event_pool.pop_back();
    } else {
      FMM_HIP(hipEventCreate(&e));
    }
    return e;
  };
  a = take();
  b = take();
  FMM_HIP(hipEventRecord(a, st));
}

PhaseTimer::~PhaseTimer() {
  if (!a) return;
  (void)hipEventRecord(b, st);
  t.events.push_back({phase, {a, b}});
}

// Folds the recorded event pairs into the per-phase totals (synchronises).
void timing_collect(Tree& t) {
  for (auto& e : t.events) {
    FMM_HIP(hipEventSynchronize(e.second.second));
    float ms = 0.f;
    FMM_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    t.phase_ms[e.first] += ms;
    t.phase_count[e.first] += 1;
    t.event_pool.push_back(e.second.first);
    t.event_pool.push_back(e.second.second);
  }
  t.events.clear();
}
const char* last_error() { return g_err.c_str(); }

const Variant* find_variant(int so, int outsz) {
  for (const Variant* v : {&variant_so0_pf, &variant_so0_p, &variant_so0_f, &variant_so1_f, &variant_so2_pf})
    if (so == v->so && outsz == v->outsz) return v;
  return nullptr;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return kOutOfMemory;
  } catch (const std::exception& e) {
    set_error(e.what());
    return kInvalidArgument;
  }
}

}  // namespace fmm

using namespace fmm;

extern "C" {

const char* fmm_last_error(void) { return fmm::last_error(); }
int fmm_abi_version(void) { return 1; }

int fmm_variant_orders(int source_order, int output_size, int* min_order, int* max_order_exclusive) {
  return guarded([&] {
    const Variant* v = find_variant(source_order, output_size);
    FMM_CHECK(v, kUnsupported, "no compiled variant for this source order / output size");
    if (min_order) *min_order = v->min_order;
    if (max_order_exclusive) *max_order_exclusive = v->max_order;
  });
}

int fmm_tree_build(const double* d_pos, const double* d_S, size_t n, size_t ncrit, int order, double theta,
                   int source_order, int output_size, void* stream, fmm_tree_t* out) {
  return guarded([&] {
    FMM_CHECK(out, kInvalidArgument, "out is NULL");
    *out = nullptr;
    FMM_CHECK(d_pos && d_S, kInvalidArgument, "positions and sources are required");
    FMM_CHECK(n >= 1 && n < (1u << 31), kInvalidArgument, "particle count must be in [1, 2^31)");
    FMM_CHECK(ncrit >= 1 && ncrit <= 4096, kInvalidArgument, "ncrit must be in [1, 4096]");
    FMM_CHECK(theta >= 0.0 && std::isfinite(theta), kInvalidArgument, "theta must be finite and non-negative");
    const Variant* v = find_variant(source_order, output_size);
    FMM_CHECK(v, kUnsupported, "no compiled variant for this source order / output size");
    FMM_CHECK(order >= v->min_order && order < v->max_order, kUnsupported,
              "expansion order outside the compiled range [" + std::to_string(v->min_order) + ", " +
                  std::to_string(v->max_order) + ")");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto* h = new fmm_tree_s();
    try {
      h->v = v;
      Tree& t = h->t;
      t.order = order;
      t.so = source_order;
      t.outsz = output_size;
      t.ncrit = ncrit;
      t.theta = theta;
      t.n = n;
      t.user_S = d_S;
      build_tree_device(t, d_pos, st);
      build_lists_device(t, st);
      build_m2l_schedule(t, st);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int fmm_tree_destroy(fmm_tree_t tree) {
  return guarded([&] {
    if (tree) {
      (void)hipDeviceSynchronize();
      for (auto& e : tree->t.events) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
      }
      for (auto e : tree->t.event_pool) (void)hipEventDestroy(e);
      if (tree->t.side) (void)hipStreamDestroy(tree->t.side);
      if (tree->t.side_fork) (void)hipEventDestroy(tree->t.side_fork);
      if (tree->t.side_join) (void)hipEventDestroy(tree->t.side_join);
    }
    delete tree;
  });
}

int fmm_timing_enable(fmm_tree_t tree, int enable) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    Tree& t = tree->t;
    timing_collect(t);
    for (int k = 0; k < kNumPhaseTimers; ++k) {
      t.phase_ms[k] = 0;
      t.phase_count[k] = 0;
    }
    t.timing = enable != 0;
  });
}

int fmm_timing_read(fmm_tree_t tree, double* h_ms, long long* h_count) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    Tree& t = tree->t;
    timing_collect(t);
    for (int k = 0; k < kNumPhaseTimers; ++k) {
      if (h_ms) h_ms[k] = t.phase_ms[k];
      if (h_count) h_count[k] = t.phase_count[k];
    }
  });
}

int fmm_tree_info(fmm_tree_t tree, long long* info) {
  return guarded([&] {
    FMM_CHECK(tree && info, kInvalidArgument, "NULL argument");
    const Tree& t = tree->t;
    info[0] = t.ncells;
    info[1] = t.nleaves;
    info[2] = t.depth;
    info[3] = t.n_m2l;
    info[4] = t.n_p2p;
    info[5] = (long long)t.n;
    info[6] = t.order;
    info[7] = t.so;
    info[8] = t.p2p_particle_pairs;
    info[9] = (long long)t.ncrit;
    info[10] = t.sched.ntiles;
    info[11] = t.sched.ngroups;
    info[12] = t.sched.nkeys;
    info[13] = t.n_p2p_items;
    info[14] = t.sched.nchunks;
    info[15] = 0;
  });
}

int fmm_tree_set_sources(fmm_tree_t tree, const double* d_S) {
  return guarded([&] {
    FMM_CHECK(tree && d_S, kInvalidArgument, "NULL argument");
    tree->t.user_S = d_S;
  });
}

int fmm_evaluate_phases(fmm_tree_t tree, int mode, int phase_mask, double* d_F, void* stream) {
  return guarded([&] {
    FMM_CHECK(tree && d_F, kInvalidArgument, "NULL argument");
    if (mode == FMM_MODE_LEARNED_F32 || mode == FMM_MODE_LEARNED_BF16X3) {
      FMM_CHECK(phase_mask == kPhaseAll, kInvalidArgument, "the learned modes run every phase");
      FMM_CHECK(tree->t.so == 0, kUnsupported, "the learned modes need source order 0");
      FMM_CHECK(tree->t.outsz != 1, kUnsupported, "the learned modes need the field (their error model)");
      tree->v->evaluate_learned(tree->t, mode, d_F, static_cast<hipStream_t>(stream));
      FMM_HIP(hipGetLastError());
      return;
    }
    FMM_CHECK(mode == FMM_MODE_PLAIN_F64 || mode == FMM_MODE_MIXED_F32 || mode == FMM_MODE_MIXED_BF16X3,
              kInvalidArgument, "unknown evaluation mode");
    FMM_CHECK((phase_mask & ~kPhaseAll) == 0, kInvalidArgument, "unknown phase bits");
    tree->v->evaluate(tree->t, mode, phase_mask, d_F, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_evaluate(fmm_tree_t tree, int mode, double* d_F, void* stream) {
  return fmm_evaluate_phases(tree, mode, FMM_PHASE_ALL, d_F, stream);
}

int fmm_evaluate_bh(fmm_tree_t tree, double* d_F, void* stream) {
  return guarded([&] {
    FMM_CHECK(tree && d_F, kInvalidArgument, "NULL argument");
    tree->v->evaluate_bh(tree->t, d_F, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_direct_sampled(const double* d_pos, const double* d_S, size_t n, const long long* d_targets, size_t nt,
                       int source_order, int output_size, double* d_F, void* stream) {
  return guarded([&] {
    FMM_CHECK(d_pos && d_S && d_F && (d_targets || nt == 0), kInvalidArgument, "NULL argument");
    FMM_CHECK(n < (1u << 31) && nt < (1u << 31), kInvalidArgument, "sizes must be below 2^31");
    const Variant* v = find_variant(source_order, output_size);
    FMM_CHECK(v, kUnsupported, "no compiled variant for this source order / output size");
    v->direct((int)n, d_pos, d_S, d_targets, (int)nt, d_F, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_op_batch(int op, int order, int source_order, int output_size, size_t n, const double* d_xyz,
                 const double* d_in, int in_stride, double* d_out, int out_stride, void* stream) {
  return guarded([&] {
    FMM_CHECK(op >= FMM_OP_P2M && op <= FMM_OP_P2M_S, kInvalidArgument, "unknown operator id");
    FMM_CHECK(n == 0 || (d_xyz && d_in && d_out), kInvalidArgument, "NULL argument");
    FMM_CHECK(n < (1u << 31), kInvalidArgument, "batch too large");
    FMM_CHECK(in_stride >= 1 && out_stride >= 1, kInvalidArgument, "strides must be positive");
    const Variant* v = find_variant(source_order, output_size);
    FMM_CHECK(v, kUnsupported, "no compiled variant for this source order / output size");
    FMM_CHECK(op == FMM_OP_P2P || (order >= v->min_order && order < v->max_order), kUnsupported,
              "expansion order outside the compiled range");
    if (n == 0) return;
    v->op_batch(op, order, (int)n, d_xyz, d_in, in_stride, d_out, out_stride, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_export_cells(fmm_tree_t tree, double* h_geo, long long* h_meta) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    const Tree& t = tree->t;
    const int nc = t.ncells;
    FMM_HIP(hipDeviceSynchronize());
    if (h_geo) {
      std::vector<double> a(nc);
      const double* src[5] = {t.cx.get(), t.cy.get(), t.cz.get(), t.cr.get(), t.crmax.get()};
      for (int k = 0; k < 5; ++k) {
        FMM_HIP(hipMemcpy(a.data(), src[k], sizeof(double) * nc, hipMemcpyDeviceToHost));
        for (int c = 0; c < nc; ++c) h_geo[5 * (size_t)c + k] = a[c];
      }
    }
    if (h_meta) {
      std::vector<int> a(nc);
      const int* src[7] = {t.parent.get(), t.level.get(), t.pstart.get(), t.pcount.get(),
                           t.nchild.get(), t.first_child.get(), t.octmask.get()};
      for (int k = 0; k < 7; ++k) {
        FMM_HIP(hipMemcpy(a.data(), src[k], sizeof(int) * nc, hipMemcpyDeviceToHost));
        for (int c = 0; c < nc; ++c) h_meta[8 * (size_t)c + k] = a[c];
      }
      for (int c = 0; c < nc; ++c) h_meta[8 * (size_t)c + 7] = 0;
    }
  });
}

int fmm_export_particles(fmm_tree_t tree, long long* h_perm, long long* h_leaf) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    const Tree& t = tree->t;
    const size_t n = t.n;
    FMM_HIP(hipDeviceSynchronize());
    std::vector<int> perm(n), lf(n);
    FMM_HIP(hipMemcpy(perm.data(), t.perm.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
    FMM_HIP(hipMemcpy(lf.data(), t.leaf_of.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < n; ++j) {
      if (h_perm) h_perm[j] = perm[j];
      if (h_leaf) h_leaf[perm[j]] = lf[j];
    }
  });
}

int fmm_export_lists(fmm_tree_t tree, long long* h_m2l, long long* h_p2p) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    const Tree& t = tree->t;
    FMM_HIP(hipDeviceSynchronize());
    auto dump = [&](const DevBuf<int>& off, const DevBuf<int>& src, long long m, long long* out, bool src_first) {
      if (!out) return;
      std::vector<int> o(t.ncells + 1), s(std::max<long long>(m, 1));
      FMM_HIP(hipMemcpy(o.data(), off.get(), sizeof(int) * (t.ncells + 1), hipMemcpyDeviceToHost));
      if (m > 0) FMM_HIP(hipMemcpy(s.data(), src.get(), sizeof(int) * m, hipMemcpyDeviceToHost));
      for (int c = 0; c < t.ncells; ++c)
        for (int i = o[c]; i < o[c + 1]; ++i) {
          out[2 * (size_t)i] = src_first ? s[i] : c;
          out[2 * (size_t)i + 1] = src_first ? c : s[i];
        }
    };
    dump(t.m2l_off, t.m2l_src, t.n_m2l, h_m2l, true);
    dump(t.p2p_off, t.p2p_src, t.n_p2p, h_p2p, false);
  });
}

int fmm_export_expansions(fmm_tree_t tree, double* h_M, double* h_L) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    const Tree& t = tree->t;
    FMM_HIP(hipDeviceSynchronize());
    const size_t ms = (size_t)msize(t.order, t.so) * t.ncells, ls = (size_t)lsize(t.order, t.so) * t.ncells;
    FMM_CHECK(t.ml_order == t.order && t.M.n >= ms && t.L.n >= ls, kInvalidArgument,
              "no FP64 expansions at the tree's order: the last evaluation was not a plain FP64 one at it");
    if (h_M) FMM_HIP(hipMemcpy(h_M, t.M.get(), sizeof(double) * ms, hipMemcpyDeviceToHost));
    if (h_L) FMM_HIP(hipMemcpy(h_L, t.L.get(), sizeof(double) * ls, hipMemcpyDeviceToHost));
  });
}

// ------------------------------------------------------------ learned order
int fmm_mlp_forward(const float* d_X, int B, const float* d_params, float* d_Y, void* stream) {
  return guarded([&] {
    FMM_CHECK(B >= 0 && (B == 0 || (d_X && d_params && d_Y)), kInvalidArgument, "NULL argument");
    mlp_forward(B, d_X, d_params, d_Y, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_mlp_infer_f16(const float* d_X, int B, const float* d_params, float* d_Y, int* d_order, double tol,
                      double margin, int min_order, int max_order, void* stream) {
  return guarded([&] {
    FMM_CHECK(B >= 0 && (B == 0 || (d_X && d_params)), kInvalidArgument, "NULL argument");
    FMM_CHECK(tol > 0 && min_order <= max_order, kInvalidArgument, "bad tolerance or order range");
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBuf<_Float16> h;
    h.alloc(kMlpParams);
    mlp_to_half(d_params, h.get(), st);
    mlp_infer_f16(B, d_X, h.get(), d_Y, d_order, (float)log10(tol), (float)margin, min_order, max_order, st);
    FMM_HIP(hipStreamSynchronize(st));
  });
}

size_t fmm_mlp_work_floats(int B) { return B > 0 ? mlp_work_floats(B) : 0; }

int fmm_mlp_grad(const float* d_X, const float* d_T, const float* d_mask, int B, const float* d_params,
                 float* d_grad, float* d_loss, float* d_work, void* stream) {
  return guarded([&] {
    FMM_CHECK(d_X && d_T && d_mask && d_params && d_grad && d_loss && d_work, kInvalidArgument, "NULL argument");
    mlp_grad(B, d_X, d_T, d_mask, d_params, d_grad, d_loss, d_work, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_mlp_grad_noalloc(const float* d_X, const float* d_T, const float* d_mask, int B, const float* d_params,
                         float* d_grad, float* d_loss, void* stream) {
  return guarded([&] {
    FMM_CHECK(d_X && d_T && d_mask && d_params && d_grad && d_loss, kInvalidArgument, "NULL argument");
    FMM_CHECK(B <= kMlpScratchBatch, kInvalidArgument,
              "batch of " + std::to_string(B) + " samples exceeds the fixed workspace (" +
                  std::to_string(kMlpScratchBatch) + " samples at most)");
    mlp_grad(B, d_X, d_T, d_mask, d_params, d_grad, d_loss, mlp_scratch(), static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_mlp_adam(const float* d_grad, float* d_params, float* d_m, float* d_v, float lr, float beta1, float beta2,
                 float eps, int step, void* stream) {
  return guarded([&] {
    FMM_CHECK(d_grad && d_params && d_m && d_v, kInvalidArgument, "NULL argument");
    mlp_adam(d_grad, d_params, d_m, d_v, lr, beta1, beta2, eps, step, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_tree_set_order(fmm_tree_t tree, int order) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    FMM_CHECK(order >= tree->v->min_order && order < tree->v->max_order, kUnsupported,
              "order " + std::to_string(order) + " not compiled for this variant");
    tree->t.order = order;
  });
}

int fmm_tree_set_m2l_form(fmm_tree_t tree, int form) {
  return guarded([&] {
    FMM_CHECK(tree, kInvalidArgument, "NULL tree");
    FMM_CHECK(form == kM2LAuto || form == kM2LFull || form == kM2LHarmonic, kInvalidArgument,
              "M2L form must be FMM_M2L_AUTO, FMM_M2L_FULL or FMM_M2L_HARMONIC");
    tree->t.m2l_form = form;
  });
}

int fmm_tree_leaf_features(fmm_tree_t tree, float* d_feat, void* stream) {
  return guarded([&] {
    FMM_CHECK(tree && d_feat, kInvalidArgument, "NULL argument");
    Tree& t = tree->t;
    hipStream_t st = static_cast<hipStream_t>(stream);
    tree->v->gather_sources(t, st);  // the current strengths (set_sources may have rebound them)
    tree->v->field_scale(t, st);
    leaf_features(t, sourcesize(t.so), st);
    if (t.nleaves > 0)
      FMM_HIP(hipMemcpyAsync(d_feat, t.leaf_feat.get(), sizeof(float) * t.nleaves * kMlpF,
                             hipMemcpyDeviceToDevice, st));
    FMM_HIP(hipStreamSynchronize(st));
  });
}

int fmm_tree_field_scale(fmm_tree_t tree, double* h_scale, void* stream) {
  return guarded([&] {
    FMM_CHECK(tree && h_scale, kInvalidArgument, "NULL argument");
    Tree& t = tree->t;
    hipStream_t st = static_cast<hipStream_t>(stream);
    tree->v->gather_sources(t, st);
    tree->v->field_scale(t, st);
    FMM_HIP(hipMemcpyAsync(h_scale, t.lw.get() + kLwScale, sizeof(double), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
  });
}

int fmm_tree_leaf_error(fmm_tree_t tree, const double* d_F, const double* d_Fref, double* d_err, void* stream) {
  return guarded([&] {
    FMM_CHECK(tree && d_F && d_Fref && d_err, kInvalidArgument, "NULL argument");
    Tree& t = tree->t;
    leaf_error(t, d_F, d_Fref, t.outsz, t.outsz == 4 ? 1 : 0, d_err, static_cast<hipStream_t>(stream));
    FMM_HIP(hipGetLastError());
  });
}

int fmm_tree_leaves(fmm_tree_t tree, long long* h_leaves) {
  return guarded([&] {
    FMM_CHECK(tree && h_leaves, kInvalidArgument, "NULL argument");
    const Tree& t = tree->t;
    std::vector<int> l(t.nleaves);
    if (t.nleaves > 0) FMM_HIP(hipMemcpy(l.data(), t.leaves.get(), sizeof(int) * t.nleaves, hipMemcpyDeviceToHost));
    for (int i = 0; i < t.nleaves; ++i) h_leaves[i] = l[i];
  });
}

int fmm_tree_set_model(fmm_tree_t tree, const float* d_params, double tol, double margin) {
  return guarded([&] {
    FMM_CHECK(tree && d_params, kInvalidArgument, "NULL argument");
    FMM_CHECK(tol > 0 && tol < 1, kInvalidArgument, "tolerance must be in (0, 1)");
    FMM_CHECK(tree->t.so == 0 && tree->t.outsz != 1, kUnsupported,
              "the order model needs charges and the field (its error model is the field's)");
    Tree& t = tree->t;
    // no evaluation on any stream may still read the buffers re-formed here
    FMM_HIP(hipDeviceSynchronize());
    t.mlp_half.alloc(kMlpParams);
    mlp_to_half(d_params, t.mlp_half.get(), nullptr);
    FMM_HIP(hipStreamSynchronize(nullptr));
    t.has_model = true;
    t.tol = tol;
    t.margin = margin;
    // re-form the M2L tiles with the predicted cell orders as a key inside
    // each class, so a tile's cells share their order
    if (t.so == 0 && t.sched.ntiles > 0 && t.n > 0) {
      tree->v->predict_cell_orders(t, nullptr);
      t.tile_hint.alloc(t.ncells);
      FMM_HIP(hipMemcpyAsync(t.tile_hint.get(), t.cell_q.get(), sizeof(int) * t.ncells, hipMemcpyDeviceToDevice,
                             nullptr));
      t.use_tile_hint = true;
      build_m2l_schedule(t, nullptr);
      FMM_HIP(hipStreamSynchronize(nullptr));
    }
    t.lstats_ready = false;  // the order statistics now describe the predicted orders
  });
}

int fmm_tree_learned_stats(fmm_tree_t tree, double* h_stats) {
  return guarded([&] {
    FMM_CHECK(tree && h_stats, kInvalidArgument, "NULL argument");
    Tree& t = tree->t;
    if (!t.lstats_ready && t.lw.get() && t.cell_q.get()) {  // order statistics of the last learned evaluation
      FMM_HIP(hipDeviceSynchronize());
      t.use_tile_order = true;
      t.lstats[5] = m2l_mac_fraction(t, t.order, nullptr, &t.lstats[6], &t.lstats[10]);
      t.use_tile_order = false;
      double lw[kLwCount];
      FMM_HIP(hipMemcpy(lw, t.lw.get(), sizeof(lw), hipMemcpyDeviceToHost));
      t.lstats[7] = lw[kLwEffFinal];
      t.lstats[8] = lw[kLwScale];
      t.lstats[9] = m2l_cell_macs(t, nullptr);
      t.lstats_ready = true;
    }
    for (int k = 0; k < FMM_LEARNED_STATS_LEN; ++k) h_stats[k] = t.lstats[k];
  });
}

int fmm_tree_leaf_orders(fmm_tree_t tree, int* h_order) {
  return guarded([&] {
    FMM_CHECK(tree && h_order, kInvalidArgument, "NULL argument");
    const Tree& t = tree->t;
    FMM_CHECK(t.leaf_q.get() != nullptr, kInvalidArgument, "no learned evaluation has run on this tree");
    if (t.nleaves > 0) FMM_HIP(hipMemcpy(h_order, t.leaf_q.get(), sizeof(int) * t.nleaves, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
