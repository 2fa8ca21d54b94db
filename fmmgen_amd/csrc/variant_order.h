// Per-order kernels of one operator variant: P2M, M2M, M2L, L2L and L2P
// with the generated operator bodies inlined.  The phase arithmetic follows
// example/calculate.cpp:135-250:
//   P2M  per leaf, the example's P2M through M2M (generated P2M_S_<p>),
//        x = leaf centre - particle
//   M2M  one level at a time, deepest first, x = parent - child
//   M2L  per target cell over its M2L sources, x = target - source
//   L2L  one level at a time, root first, x = child - parent
//   L2P  per particle, x = particle - leaf centre, added in FP64 to the
//        near-field sum and written in caller order.
#pragma once

#include "m2l_gemm.h"
#include "variant_decl.h"

namespace fmm {
namespace FMM_VARIANT_NS {

// Output proxy of a generated body that keeps slots [LO, HI) in registers.
// Inlined with constant slot indices, every output outside the chunk folds
// to dead code, so one instantiation per chunk computes only its slots.
template <typename T, int LO, int HI>
struct ChunkAcc {
  T v[HI - LO];
  struct Ref {
    T* p;
    __device__ __forceinline__ void operator+=(T x) {
      if (p) *p += x;
    }
  };
  __device__ __forceinline__ Ref operator[](int k) { return Ref{(k >= LO && k < HI) ? &v[k - LO] : nullptr}; }
};

constexpr int kChunk = FMMGEN_CHUNK;  // expansion slots per thread (one pruned generated body per chunk)

template <int NCH, typename F>
__device__ __forceinline__ void dispatch_chunk(int ch, F&& f) {
  static_for<NCH>([&](auto C) {
    if (ch == C) f(C);
  });
}

// P2M: the example's P2M through M2M (example/calculate.cpp:135-161),
// x = leaf centre - particle, one thread per (leaf, slot chunk).
template <int P>
__global__ __launch_bounds__(64) void k_p2m(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pstart,
                      const int* __restrict__ pcount, const double* __restrict__ cx, const double* __restrict__ cy,
                      const double* __restrict__ cz, const double* __restrict__ px, const double* __restrict__ py,
                      const double* __restrict__ pz, const double* __restrict__ pS, double* __restrict__ M) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk;
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  const double x0 = cx[c], y0 = cy[c], z0 = cz[c];
  const int b = pstart[c], e = b + pcount[c];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) acc.v[k] = 0.0;
    for (int j = b; j < e; ++j)
      fmmgen_ops::P2M_S_chunk<P, LO / kChunk, double>(x0 - px[j], y0 - py[j], z0 - pz[j], pS + (size_t)j * SS, acc);
    double* Mc = M + (size_t)c * MS;
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) Mc[LO + k] = acc.v[k];
  });
}

// M2M (example/calculate.cpp:163-187), x = parent - child, one thread per
// (parent, slot chunk) of one level; children are summed in registers.
template <int P>
__global__ __launch_bounds__(64) void k_m2m(int beg, int end, const int* __restrict__ first_child, const int* __restrict__ nchild,
                      const double* __restrict__ cx, const double* __restrict__ cy, const double* __restrict__ cz,
                      double* __restrict__ M) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk;
  const int p = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= end) return;
  const int k = nchild[p];
  if (k == 0) return;
  const int fc = first_child[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.0;
    for (int c = fc; c < fc + k; ++c)
      fmmgen_ops::M2M_chunk<P, LO / kChunk, double>(cx[p] - cx[c], cy[p] - cy[c], cz[p] - cz[c],
                                   static_cast<const double*>(M + (size_t)c * MS), acc);
    double* Mp = M + (size_t)p * MS;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) Mp[LO + j] = acc.v[j];
  });
}

// D vector of every M2L key (the derivative block of the generated M2L
// operator, fmmgen/writer.py:168-182), FP64, or normalised by the target
// half-width (Dhat_k = D_k r^(|k|+1)) and rounded to FP32.
template <int P, typename T, bool SCALED>
__global__ void k_dtab(int nkeys, const double* __restrict__ key_dx, const double* __restrict__ key_r,
                       T* __restrict__ Dtab) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  constexpr int MS = Op<P>::MS;
  double D[MS];
  fmmgen_ops::Dgen_o<P, double>(key_dx[3 * k], key_dx[3 * k + 1], key_dx[3 * k + 2], D);
  if constexpr (SCALED) {
    const double r = key_r[k];
    double pw[P + 2];
    pw[0] = r;
#pragma unroll
    for (int d = 1; d <= P + 1; ++d) pw[d] = pw[d - 1] * r;
#pragma unroll
    for (int j = 0; j < MS; ++j) D[j] *= pw[slot_degree(j + M2LPlan<P, SO>::OFF)];
  }
#pragma unroll
  for (int j = 0; j < MS; ++j) Dtab[(size_t)k * MS + j] = (T)D[j];
}

template <int P, int W, int LAY = kGraded>
const unsigned* ensure_index_table(Tree& t, hipStream_t st) {
  static_assert(W == 2 || W == 4 || W == 8, "blocks of 2, 4 or 8 waves");
  M2LSchedule& S = t.sched;
  const int wi = W == 8 ? 1 : (W == 2 ? 2 : 0);
  if (!S.idx_ready[LAY][P][wi]) {
    const std::vector<unsigned> tab = m2l_index_table<P, SO, W, LAY>();
    S.idxtab[LAY][P][wi].alloc(tab.size());
    FMM_HIP(hipMemcpyAsync(S.idxtab[LAY][P][wi].get(), tab.data(), sizeof(unsigned) * tab.size(),
                           hipMemcpyHostToDevice, st));
    FMM_HIP(hipStreamSynchronize(st));
    S.idx_ready[LAY][P][wi] = true;
  }
  return S.idxtab[LAY][P][wi].get();
}

// Harmonic-layout M2L around the GEMM: fold every cell's multipole (M,
// graded, stride ms) into Mh, clear Lh (stride of order ps), then, after
// the GEMM, store the unfolded rows into L (graded, stride ls).
template <int PS, typename T>
void harm_fold(Tree& t, const T* M, int ms, DevBuf<T>& Mh, DevBuf<T>& Lh, hipStream_t st) {
  const int nc = t.ncells, hs = harm_size(PS);
  Mh.alloc((size_t)nc * hs);
  Lh.alloc((size_t)nc * hs);
  if (nc == 0) return;
  constexpr int per = 4 * kHarmCellsPerWave;
  k_harm_fold<PS, T><<<(nc + per - 1) / per, 256, 0, st>>>(nc, M, ms, Mh.get(), hs);
  FMM_HIP(hipMemsetAsync(Lh.get(), 0, sizeof(T) * (size_t)nc * hs, st));
}
template <int PS, typename T>
void harm_unfold(Tree& t, const DevBuf<T>& Lh, T* L, int ls, hipStream_t st) {
  if (t.ncells == 0) return;
  constexpr int per = 4 * kHarmCellsPerWave;
  k_harm_unfold<PS, T><<<(t.ncells + per - 1) / per, 256, 0, st>>>(t.ncells, Lh.get(), harm_size(PS), L, ls);
}

template <int P, int W, int LAY = kGraded>
const unsigned* ensure_index_table32(Tree& t, hipStream_t st) {
  M2LSchedule& S = t.sched;
  if (!S.idx32_ready[LAY][P]) {
    const std::vector<unsigned> tab = m2l_index_table32<P, SO, W, LAY>();
    S.idxtab32[LAY][P].alloc(tab.size());
    FMM_HIP(hipMemcpyAsync(S.idxtab32[LAY][P].get(), tab.data(), sizeof(unsigned) * tab.size(),
                           hipMemcpyHostToDevice, st));
    FMM_HIP(hipStreamSynchronize(st));
    S.idx32_ready[LAY][P] = true;
  }
  return S.idxtab32[LAY][P].get();
}

template <int P>
void m2l_plain(Tree& t, hipStream_t st) {
  M2LSchedule& S = t.sched;
  if (S.ngroups == 0) return;
  S.Dtab64.alloc((size_t)S.nkeys * Op<P>::MS);
  k_dtab<P, double, false><<<nblk(S.nkeys, 64), 64, 0, st>>>(S.nkeys, S.key_dx.get(), S.key_r.get(), S.Dtab64.get());
  const int grid = ((S.nchunks + 7) / 8) * 8;
  if constexpr (SO == 0) {
    if (m2l_harmonic(t, 0)) {
      constexpr int W = m2l_waves<P, SO, double, kHarmonic>();
      const unsigned* idx = ensure_index_table<P, W, kHarmonic>(t, st);
      harm_fold<P, double>(t, t.M.get(), Op<P>::MS, t.Mh64, t.Lh64, st);
      k_m2l_gemm<P, SO, double, false, W, kHarmonic><<<grid, 64 * W, 0, st>>>(
          S.nchunks, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(), S.chunk_g1.get(),
          S.group_key.get(), S.group_src.get(), S.key_dlevel.get(), S.Dtab64.get(), idx, t.tile_order_ptr(), nullptr,
          Op<P>::MS, harm_size(P), harm_size(P), t.Mh64.get(), t.Lh64.get());
      harm_unfold<P, double>(t, t.Lh64, t.L.get(), Op<P>::LS, st);
      return;
    }
  }
  constexpr int W = m2l_waves<P, SO, double>();
  const unsigned* idx = ensure_index_table<P, W>(t, st);
  k_m2l_gemm<P, SO, double, false, W><<<grid, 64 * W, 0, st>>>(
      S.nchunks, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(), S.chunk_g1.get(), S.group_key.get(),
      S.group_src.get(), S.key_dlevel.get(), S.Dtab64.get(), idx, t.tile_order_ptr(), nullptr, Op<P>::MS,
      Op<P>::MS, Op<P>::LS, t.M.get(), t.L.get());
}

// L2L (example/calculate.cpp:219-234), x = child - parent, one thread per
// (child, slot chunk) of one level.
template <int P>
__global__ __launch_bounds__(64) void k_l2l(int beg, int end, const int* __restrict__ parent, const double* __restrict__ cx,
                      const double* __restrict__ cy, const double* __restrict__ cz, double* __restrict__ L) {
  constexpr int LS = Op<P>::LS, NCH = (LS + kChunk - 1) / kChunk;
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  const int p = parent[c];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < LS) ? LO + kChunk : LS;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.0;
    fmmgen_ops::L2L_chunk<P, LO / kChunk, double>(cx[c] - cx[p], cy[c] - cy[p], cz[c] - cz[p],
                                 static_cast<const double*>(L + (size_t)p * LS), acc);
    double* Lc = L + (size_t)c * LS;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) Lc[LO + j] += acc.v[j];
  });
}

template <int P>
__global__ __launch_bounds__(64) void k_l2p_final(int n, const int* __restrict__ leaf_of, const int* __restrict__ perm,
                            const double* __restrict__ cx, const double* __restrict__ cy,
                            const double* __restrict__ cz, const double* __restrict__ px,
                            const double* __restrict__ py, const double* __restrict__ pz,
                            const double* __restrict__ L, const double* __restrict__ Fnear, int use_l2p,
                            double* __restrict__ F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double acc[OUT];
#pragma unroll
  for (int k = 0; k < OUT; ++k) acc[k] = Fnear ? Fnear[(size_t)j * OUT + k] : 0.0;
  if (use_l2p) {
    const int c = leaf_of[j];
    fmmgen_ops::L2P_o<P, double>(px[j] - cx[c], py[j] - cy[c], pz[j] - cz[c], L + (size_t)c * Op<P>::LS, acc);
  }
  const int i = perm[j];
#pragma unroll
  for (int k = 0; k < OUT; ++k) F[(size_t)i * OUT + k] = acc[k];
}

// One operator over a batch of independent inputs (test entry point).
template <int P, int OPID>
__global__ void k_op_batch(int n, const double* __restrict__ xyz, const double* __restrict__ in, int in_stride,
                           double* __restrict__ out, int out_stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double* a = in + (size_t)i * in_stride;
  double* o = out + (size_t)i * out_stride;
  if constexpr (OPID == kOpP2M) fmmgen_ops::P2M_o<P, double>(x, y, z, a[0], o);
  if constexpr (OPID == kOpM2M) fmmgen_ops::M2M_o<P, double>(x, y, z, a, o);
  if constexpr (OPID == kOpM2L) fmmgen_ops::M2L_o<P, double>(x, y, z, a, o);
  if constexpr (OPID == kOpL2L) fmmgen_ops::L2L_o<P, double>(x, y, z, a, o);
  if constexpr (OPID == kOpL2P) fmmgen_ops::L2P_o<P, double>(x, y, z, a, o);
  if constexpr (OPID == kOpM2P) fmmgen_ops::M2P_o<P, double>(x, y, z, a, o);
  if constexpr (OPID == kOpP2MS) fmmgen_ops::P2M_S_o<P, double>(x, y, z, a, o);
}

}  // namespace FMM_VARIANT_NS
}  // namespace fmm

#include "variant_mixed.h"

namespace fmm {
namespace FMM_VARIANT_NS {

template <int P>
void m2l_mixed(Tree& t, int mode, hipStream_t st) {
  M2LSchedule& S = t.sched;
  if (S.ngroups == 0) return;
  S.Dtab32.alloc((size_t)S.nkeys * Op<P>::MS);
  k_dtab<P, float, true><<<nblk(S.nkeys, 64), 64, 0, st>>>(S.nkeys, S.key_dx.get(), S.key_r.get(), S.Dtab32.get());
  const bool harm = m2l_harmonic(t, mode);
  if constexpr (SO == 0)
    if (harm) harm_fold<P, float>(t, t.Mf.get(), Op<P>::MS, t.Mhf, t.Lhf, st);
  if (t.use_tile_order && t.m2l_classes) {  // learned order: one launch per tile-order class
    t.m2l_classes(t, mode, P, st);
  } else {
    m2l_class<P>(t, mode, nullptr, S.nchunks, P, st);
  }
  if constexpr (SO == 0)
    if (harm) harm_unfold<P, float>(t, t.Lhf, t.Lf.get(), Op<P>::LS, st);
}

// FP32 / BF16x3 M2L of the listed chunks (all when chunks == nullptr) at
// order Q, on expansions and D vectors stored at order ps (the layouts are
// prefixes: row strides of order ps).
template <int Q>
void m2l_class(Tree& t, int mode, const int* chunks, int nch, int ps, hipStream_t st) {
  M2LSchedule& S = t.sched;
  if (nch == 0) return;
  const int grid = ((nch + 7) / 8) * 8;
  const int ms = msize(ps, SO), ls = lsize(ps, SO);
  if (mode == 2) {  // BF16x3 split precision
    if constexpr (SO == 0) {
      if (m2l_harmonic(t, mode)) {
        constexpr int W32 = m2l_waves32<Q, SO, kHarmonic>();
        const unsigned* idx = ensure_index_table32<Q, W32, kHarmonic>(t, st);
        k_m2l_bf16x3<Q, SO, W32, kHarmonic><<<grid, 64 * W32, 0, st>>>(
            nch, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(), S.chunk_g1.get(), S.group_key.get(),
            S.group_src.get(), S.key_dlevel.get(), S.Dtab32.get(), idx, t.tile_order_ptr(), chunks, ms,
            harm_size(ps), harm_size(ps), t.Mhf.get(), t.Lhf.get());
        return;
      }
    }
    constexpr int W32 = m2l_waves32<Q, SO>();
    const unsigned* idx = ensure_index_table32<Q, W32>(t, st);
    k_m2l_bf16x3<Q, SO, W32><<<grid, 64 * W32, 0, st>>>(nch, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(),
                                             S.chunk_g1.get(), S.group_key.get(), S.group_src.get(),
                                             S.key_dlevel.get(), S.Dtab32.get(), idx, t.tile_order_ptr(), chunks,
                                             ms, ms, ls, t.Mf.get(), t.Lf.get());
    return;
  }
  if constexpr (SO == 0) {
    if (m2l_harmonic(t, mode)) {
      constexpr int W = m2l_waves<Q, SO, float, kHarmonic>();
      const unsigned* idx = ensure_index_table<Q, W, kHarmonic>(t, st);
      k_m2l_gemm<Q, SO, float, true, W, kHarmonic><<<grid, 64 * W, 0, st>>>(
          nch, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(), S.chunk_g1.get(), S.group_key.get(),
          S.group_src.get(), S.key_dlevel.get(), S.Dtab32.get(), idx, t.tile_order_ptr(), chunks, ms,
          harm_size(ps), harm_size(ps), t.Mhf.get(), t.Lhf.get());
      return;
    }
  }
  constexpr int W = m2l_waves<Q, SO, float>();
  const unsigned* idx = ensure_index_table<Q, W>(t, st);
  k_m2l_gemm<Q, SO, float, true, W><<<grid, 64 * W, 0, st>>>(
      nch, grid, S.tile_cells.get(), S.chunk_tile.get(), S.chunk_g0.get(), S.chunk_g1.get(), S.group_key.get(),
      S.group_src.get(), S.key_dlevel.get(), S.Dtab32.get(), idx, t.tile_order_ptr(), chunks, ms, ms, ls,
      t.Mf.get(), t.Lf.get());
}

template <int P>
void far_field_order(Tree& t, int mode, int mask, hipStream_t st) {
  if (mode != 0) {
    far_field_mixed<P>(t, mode, mask, st);
    return;
  }
  const int nc = t.ncells;
  if (mask & kPhaseP2M) {
    PhaseTimer pt(t, kTP2M, st);
    k_p2m<P><<<dim3(nblk(t.nleaves, 64), (Op<P>::MS + kChunk - 1) / kChunk), 64, 0, st>>>(t.nleaves, t.leaves.get(), t.pstart.get(), t.pcount.get(),
                                                 t.cx.get(), t.cy.get(), t.cz.get(), t.px.get(), t.py.get(),
                                                 t.pz.get(), t.pS.get(), t.M.get());
  }
  if (mask & kPhaseM2M) {
    PhaseTimer pt(t, kTM2M, st);
    for (int l = t.depth - 2; l >= 0; --l) {
      const int b = t.level_off[l], e = t.level_off[l + 1];
      k_m2m<P><<<dim3(nblk(e - b, 64), (Op<P>::MS + kChunk - 1) / kChunk), 64, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cx.get(),
                                               t.cy.get(), t.cz.get(), t.M.get());
    }
  }
  if (mask & kPhaseM2L) {
    PhaseTimer pt(t, kTM2L, st);
    m2l_plain<P>(t, st);
  }
  if (mask & kPhaseL2L) {
    PhaseTimer pt(t, kTL2L, st);
    for (int l = 1; l < t.depth; ++l) {
      const int b = t.level_off[l], e = t.level_off[l + 1];
      k_l2l<P><<<dim3(nblk(e - b, 64), (Op<P>::LS + kChunk - 1) / kChunk), 64, 0, st>>>(b, e, t.parent.get(), t.cx.get(), t.cy.get(), t.cz.get(),
                                               t.L.get());
    }
  }
}

template <int P>
void l2p_final_order(Tree& t, int mode, int mask, double* F, hipStream_t st) {
  if (mode != 0) {
    l2p_mixed<P>(t, mask, F, st);
    return;
  }
  const int n = (int)t.n;
  PhaseTimer pt(t, kTL2P, st);
  k_l2p_final<P><<<nblk(n, 64), 64, 0, st>>>(n, t.leaf_of.get(), t.perm.get(), t.cx.get(), t.cy.get(),
                                               t.cz.get(), t.px.get(), t.py.get(), t.pz.get(), t.L.get(),
                                               (mask & kPhaseP2P) ? t.Fp2p.get() : nullptr,
                                               (mask & kPhaseL2P) ? 1 : 0, F);
}

// Operators above the full-body order (their whole bodies are not
// compiled there).  P2M, M2M and L2L: the chunked bodies, one thread per
// (item, output chunk) as in the tree kernels; the charge P2M is the
// example's P2M_S at the opposite displacement (q (-x)^n / n!).
template <int P, int OPID>
__global__ __launch_bounds__(64) void k_op_batch_chunked(int n, const double* __restrict__ xyz,
                                                         const double* __restrict__ in, int in_stride,
                                                         double* __restrict__ out, int out_stride) {
  constexpr int NOUT = OPID == kOpL2L ? Op<P>::LS : Op<P>::MS, NCH = (NOUT + kChunk - 1) / kChunk;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double* a = in + (size_t)i * in_stride;
  double* o = out + (size_t)i * out_stride;
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < NOUT) ? LO + kChunk : NOUT;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) acc.v[k] = 0.0;
    if constexpr (OPID == kOpM2M) fmmgen_ops::M2M_chunk<P, LO / kChunk, double>(x, y, z, a, acc);
    if constexpr (OPID == kOpL2L) fmmgen_ops::L2L_chunk<P, LO / kChunk, double>(x, y, z, a, acc);
    if constexpr (OPID == kOpP2MS) fmmgen_ops::P2M_S_chunk<P, LO / kChunk, double>(x, y, z, a, acc);
    if constexpr (OPID == kOpP2M) {
      static_assert(SO == 0, "charge P2M through P2M_S needs source order 0");
      fmmgen_ops::P2M_S_chunk<P, LO / kChunk, double>(-x, -y, -z, a, acc);
    }
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) o[LO + k] += acc.v[k];
  });
}

// M2L and M2P above the full-body order, one block per item: the generated
// derivative block (orders up to P for M2L, P + 1 for M2P's field) is
// staged in LDS and the sums run over the multi-index arithmetic of the
// graded layout:
//   M2L  L_n += sum_m D_{n+m} M_m,  |n| + |m| <= P
//   M2P  phi += sum_m M_m D_m,  F_a -= sum_m M_m D_{m+e_a}
template <int P, int OPID>
__global__ __launch_bounds__(64) void k_op_batch_dsum(int n, const double* __restrict__ xyz,
                                                      const double* __restrict__ in, int in_stride,
                                                      double* __restrict__ out, int out_stride) {
  constexpr int MS = Op<P>::MS, LS = Op<P>::LS, OFF = nterms(SO - 1);
  constexpr int PD = OPID == kOpM2P && OUT != 1 ? P + 1 : P;  // derivative order
  constexpr int DS = msize(PD, SO);
  __shared__ double sD[DS];
  __shared__ double sM[MS];
  const int i = blockIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double* a = in + (size_t)i * in_stride;
  double* o = out + (size_t)i * out_stride;
  if (threadIdx.x == 0) {
    double D[DS];
    fmmgen_ops::Dgen_o<PD, double>(x, y, z, D);
    for (int j = 0; j < DS; ++j) sD[j] = D[j];
  }
  for (int j = threadIdx.x; j < MS; j += blockDim.x) sM[j] = a[j];
  __syncthreads();
  if constexpr (OPID == kOpM2L) {
    for (int r = threadIdx.x; r < LS; r += blockDim.x) {
      const MultiIdx na = slot_multi(r);
      const int dn = na.x + na.y + na.z;
      double acc = 0.0;
      for (int m = 0; m < MS; ++m) {
        const MultiIdx mb = slot_multi(m + OFF);
        if (dn + mb.x + mb.y + mb.z > P) break;  // M slots ascend in degree
        acc += sD[multi_slot(na.x + mb.x, na.y + mb.y, na.z + mb.z) - OFF] * sM[m];
      }
      o[r] += acc;
    }
  } else {
    // outputs: potential (when present), then the three field components
    constexpr bool POT = OUT != 3;
    const int j = threadIdx.x;
    if (j < OUT) {
      const int ax = POT ? j - 1 : j;  // -1: potential
      double acc = 0.0;
      for (int m = 0; m < MS; ++m) {
        const MultiIdx mb = slot_multi(m + OFF);
        const int k = multi_slot(mb.x + (ax == 0), mb.y + (ax == 1), mb.z + (ax == 2)) - OFF;
        acc += sM[m] * sD[k];
      }
      o[j] += ax < 0 ? acc : -acc;
    }
  }
}

// Barnes-Hut (Tree::compute_field_bh, example/tree.cpp:273-293, and
// evaluate_M2P_and_P2P, example/calculate.cpp:266-314), FP64.  Every
// particle walks the tree from the root: below an internal cell each child
// c is opened when r_c > theta |x_i - centre_c| and applied by M2P at
// x_i - centre_c otherwise; a leaf adds P2P from its particles except i.
// One wave takes 64 consecutive sorted particles (one or two leaves
// mostly) and walks the union of their paths on a stack in LDS whose
// entries carry the lanes that reached the cell; a lane applies M2P only
// at the cells where its own walk stops, so each particle sums exactly the
// reference's terms.  The distance of the criterion is rounded as the
// reference's host code rounds it, so every decision matches.
constexpr int kBHStack = 256;  // a pop pushes at most 8: 7 per level + 1

template <int P>
__device__ __forceinline__ void bh_m2p(double dx, double dy, double dz, const double* __restrict__ Mc,
                                       double* acc) {
  if constexpr (P <= FMMGEN_FULL_BODIES_MAXORDER) {
    fmmgen_ops::M2P_o<P, double>(dx, dy, dz, Mc, acc);
  } else {
    // the derivative block of order P + 1 (P without the field) and the
    // sums of M2P over the graded layout
    constexpr int MS = Op<P>::MS, OFF = nterms(SO - 1);
    constexpr int PD = OUT != 1 ? P + 1 : P, DS = msize(PD, SO);
    constexpr bool POT = OUT != 3;
    double D[DS];
    fmmgen_ops::Dgen_o<PD, double>(dx, dy, dz, D);
    double phi = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int m = 0; m < MS; ++m) {
      const MultiIdx b = slot_multi(m + OFF);
      const double w = Mc[m];
      if (POT) phi += w * D[m];
      if (OUT != 1) {
        fx += w * D[multi_slot(b.x + 1, b.y, b.z) - OFF];
        fy += w * D[multi_slot(b.x, b.y + 1, b.z) - OFF];
        fz += w * D[multi_slot(b.x, b.y, b.z + 1) - OFF];
      }
    }
    int k = 0;
    if (POT) acc[k++] += phi;
    if (OUT != 1) {
      acc[k] -= fx;
      acc[k + 1] -= fy;
      acc[k + 2] -= fz;
    }
  }
}

template <int P>
__global__ __launch_bounds__(64) void k_bh(int n, const int* __restrict__ first_child, const int* __restrict__ nchild,
                                           const double* __restrict__ cx, const double* __restrict__ cy,
                                           const double* __restrict__ cz, const double* __restrict__ cr,
                                           const int* __restrict__ pstart, const int* __restrict__ pcount,
                                           const double* __restrict__ px, const double* __restrict__ py,
                                           const double* __restrict__ pz, const double* __restrict__ pS,
                                           const int* __restrict__ perm, const double* __restrict__ M, double theta,
                                           double* __restrict__ F) {
  constexpr int MS = Op<P>::MS;
  __shared__ int s_cell[kBHStack];
  __shared__ unsigned long long s_lanes[kBHStack];
  const int lane = threadIdx.x;
  const int s = blockIdx.x * 64 + lane;
  const bool valid = s < n;
  const double xi = valid ? px[s] : 0.0, yi = valid ? py[s] : 0.0, zi = valid ? pz[s] : 0.0;
  double acc[OUT];
#pragma unroll
  for (int k = 0; k < OUT; ++k) acc[k] = 0.0;
  const unsigned long long all = __ballot(valid);
  if (lane == 0) {
    s_cell[0] = 0;
    s_lanes[0] = all;
  }
  int top = 1;  // uniform over the wave
  __syncthreads();
  while (top > 0) {
    --top;
    const int p = s_cell[top];
    const bool mine = (s_lanes[top] >> lane) & 1ull;
    __syncthreads();  // the entry is read before a push overwrites it
    const int nch = nchild[p];
    if (nch > 0) {
      const int fc = first_child[p];
      // children pushed last to first, so they are walked in octant order
      for (int k = nch - 1; k >= 0; --k) {
        const int c = fc + k;
        const double dx = xi - cx[c], dy = yi - cy[c], dz = zi - cz[c];
        const bool open = mine && cr[c] > theta * norm3_rn(dx, dy, dz);
        if (mine && !open) bh_m2p<P>(dx, dy, dz, M + (size_t)c * MS, acc);
        const unsigned long long lanes = __ballot(open);
        if (lanes != 0ull && top < kBHStack) {
          if (lane == 0) {
            s_cell[top] = c;
            s_lanes[top] = lanes;
          }
          ++top;
        }
      }
      __syncthreads();
    } else if (mine) {
      const int b = pstart[p], e = b + pcount[p];
      for (int j = b; j < e; ++j) {
        if (j == s) continue;
        fmmgen_ops::P2P<double>(xi - px[j], yi - py[j], zi - pz[j], pS + (size_t)j * SS, acc);
      }
    }
  }
  if (valid) {
    const size_t i = (size_t)perm[s];
#pragma unroll
    for (int k = 0; k < OUT; ++k) F[i * OUT + k] = acc[k];
  }
}

template <int P>
void bh_order(Tree& t, double* F, hipStream_t st) {
  const int n = (int)t.n;
  FMM_CHECK(7 * t.depth + 1 <= kBHStack, kUnsupported, "tree too deep for the Barnes-Hut walk");
  if (n == 0) return;
  PhaseTimer pt(t, kTBH, st);
  k_bh<P><<<nblk(n, 64), 64, 0, st>>>(n, t.first_child.get(), t.nchild.get(), t.cx.get(), t.cy.get(), t.cz.get(),
                                       t.cr.get(), t.pstart.get(), t.pcount.get(), t.px.get(), t.py.get(),
                                       t.pz.get(), t.pS.get(), t.perm.get(), t.M.get(), t.theta, F);
}

template <int P>
void op_batch_order(int op, int n, const double* xyz, const double* in, int is, double* out, int os,
                    hipStream_t st) {
  const int g = nblk(n, 64);
  if constexpr (P > FMMGEN_FULL_BODIES_MAXORDER) {
    // whole-operator bodies are compiled up to FMMGEN_FULL_BODIES_MAXORDER
    constexpr int NM = (Op<P>::MS + kChunk - 1) / kChunk, NL = (Op<P>::LS + kChunk - 1) / kChunk;
    switch (op) {
      case kOpP2M:
        if constexpr (SO == 0) {
          k_op_batch_chunked<P, kOpP2M><<<dim3(g, NM), 64, 0, st>>>(n, xyz, in, is, out, os);
          return;
        }
        throw Error{kUnsupported, "charge P2M above the full-body order needs source order 0"};
      case kOpP2MS: k_op_batch_chunked<P, kOpP2MS><<<dim3(g, NM), 64, 0, st>>>(n, xyz, in, is, out, os); return;
      case kOpM2M: k_op_batch_chunked<P, kOpM2M><<<dim3(g, NM), 64, 0, st>>>(n, xyz, in, is, out, os); return;
      case kOpL2L: k_op_batch_chunked<P, kOpL2L><<<dim3(g, NL), 64, 0, st>>>(n, xyz, in, is, out, os); return;
      case kOpM2L: k_op_batch_dsum<P, kOpM2L><<<n, 64, 0, st>>>(n, xyz, in, is, out, os); return;
      case kOpM2P:
        if constexpr (P + 1 <= FMMGEN_DGEN_MAXORDER) {
          k_op_batch_dsum<P, kOpM2P><<<n, 64, 0, st>>>(n, xyz, in, is, out, os);
          return;
        }
        throw Error{kUnsupported, "M2P above the compiled derivative order"};
      case kOpL2P: k_op_batch<P, kOpL2P><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); return;
      default: throw Error{kInvalidArgument, "operator has no order"};
    }
  }
  switch (op) {
    case kOpP2M: k_op_batch<P, kOpP2M><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpM2M: k_op_batch<P, kOpM2M><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpM2L: k_op_batch<P, kOpM2L><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpL2L: k_op_batch<P, kOpL2L><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpL2P: k_op_batch<P, kOpL2P><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpM2P: k_op_batch<P, kOpM2P><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    case kOpP2MS: k_op_batch<P, kOpP2MS><<<g, 64, 0, st>>>(n, xyz, in, is, out, os); break;
    default: throw Error{kInvalidArgument, "operator has no order"};
  }
}

}  // namespace FMM_VARIANT_NS
}  // namespace fmm

#define FMM_INSTANTIATE_ORDER(P)                                                                        \
  namespace fmm {                                                                                       \
  namespace FMM_VARIANT_NS {                                                                            \
  template void far_field_order<P>(Tree&, int, int, hipStream_t);                                       \
  template void l2p_final_order<P>(Tree&, int, int, double*, hipStream_t);                              \
  template void op_batch_order<P>(int, int, const double*, const double*, int, double*, int, hipStream_t); \
  template void bh_order<P>(Tree&, double*, hipStream_t);                                                 \
  template void m2l_class<P>(Tree&, int, const int*, int, int, hipStream_t);                         \
  }                                                                                                     \
  }
