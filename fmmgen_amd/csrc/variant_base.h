// Order-independent part of one operator variant: source gather, the P2P
// near field, the FP64 direct sum, order dispatch and the variant table.
#pragma once

#include "variant_decl.h"

namespace fmm {
namespace FMM_VARIANT_NS {

__global__ void k_gather_S(int n, const int* __restrict__ perm, const double* __restrict__ S,
                           double* __restrict__ pS) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int i = perm[j];
#pragma unroll
  for (int k = 0; k < SS; ++k) pS[(size_t)j * SS + k] = S[(size_t)i * SS + k];
}

// FP64 direct sum on sampled targets (example/calculate.cpp:252-264): one
// block per target, sources strided over the lanes, tree reduction in LDS.
__global__ void k_direct(int n, const double* __restrict__ pos, const double* __restrict__ S,
                         const long long* __restrict__ tg, int nt, double* __restrict__ F) {
  const int t = blockIdx.x;
  if (t >= nt) return;
  const long long i = tg[t];
  if (i < 0 || i >= n) {  // out-of-range target: NaN row, no access
    if (threadIdx.x == 0)
      for (int k = 0; k < OUT; ++k) F[(size_t)t * OUT + k] = __longlong_as_double(0x7ff8000000000000LL);
    return;
  }
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  double acc[OUT];
#pragma unroll
  for (int k = 0; k < OUT; ++k) acc[k] = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    if (j == i) continue;
    fmmgen_ops::P2P<double>(xi - pos[3 * j], yi - pos[3 * j + 1], zi - pos[3 * j + 2], S + (size_t)j * SS, acc);
  }
  __shared__ double red[256][OUT];
#pragma unroll
  for (int k = 0; k < OUT; ++k) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < OUT; ++k) red[threadIdx.x][k] += red[threadIdx.x + s][k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < OUT; ++k) F[(size_t)t * OUT + k] = red[0][k];
}

__global__ void k_op_p2p(int n, const double* __restrict__ xyz, const double* __restrict__ in, int in_stride,
                         double* __restrict__ out, int out_stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fmmgen_ops::P2P<double>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], in + (size_t)i * in_stride,
                          out + (size_t)i * out_stride);
}

// P2P near field (example/calculate.cpp:24-40, 208-216), one wavefront per
// work item: a segment of at most 8192 sources of one target leaf.
//
// The sources of a target leaf are its own particles (self pairs masked)
// and the merged particle ranges of its other P2P leaves; they are staged
// 64 at a time through LDS with coalesced loads.  A leaf whose sources span
// several items (a small leaf beside a dense cluster) adds its partial sums
// atomically into a zeroed output; other leaves store.  A leaf of nt <= 32
// particles is replicated over 64 / pow2(nt) lane slices that take
// alternate sources of each chunk, and the slices are summed at the end.
//   T = double: FP64 throughout (the reference arithmetic).
//   T = float : dx = float(x_t - x_s) from FP64 positions (no cancellation
//               loss for close pairs), the generated P2P body in FP32, and
//               compensated (Kahan) accumulation of each 64-source chunk's
//               partial sum into the running total.
// The result is written in FP64 per sorted slot.
template <typename T, bool SELF>
__device__ __forceinline__ void p2p_chunk(int filled, int sl, int nslice, int tslot, double xt, double yt,
                                          double zt, const double* sx, const double* sy, const double* sz,
                                          const T* sS, const int* sid, T (&part)[OUT]) {
  for (int j = sl; j < filled; j += nslice) {
    T dx, dy, dz;
    dx = (T)(xt - sx[j]);
    dy = (T)(yt - sy[j]);
    dz = (T)(zt - sz[j]);
    T sj[SS];
#pragma unroll
    for (int k = 0; k < SS; ++k) sj[k] = sS[j * SS + k];
    if (SELF && sid[j] == tslot) {  // self pair: zero source at unit distance
      dx = T(1);
      dy = T(0);
      dz = T(0);
#pragma unroll
      for (int k = 0; k < SS; ++k) sj[k] = T(0);
    }
    fmmgen_ops::P2P<T>(dx, dy, dz, sj, part);
  }
}

template <typename T>
__global__ __launch_bounds__(64) void k_p2p_leaf(int nitems, int nblocks, const int* __restrict__ it_leaf,
                                                 const int* __restrict__ it_r, const int* __restrict__ it_o,
                                                 const int* __restrict__ it_n, const int* __restrict__ pstart,
                                                 const int* __restrict__ pcount, const int* __restrict__ rng_off,
                                                 const int* __restrict__ rng_start, const int* __restrict__ rng_len,
                                                 const double* __restrict__ px, const double* __restrict__ py,
                                                 const double* __restrict__ pz, const double* __restrict__ pS,
                                                 int split, double* __restrict__ F) {
  constexpr bool kKahan = sizeof(T) == 4;
  __shared__ double sx[64], sy[64], sz[64];
  __shared__ T sS[64 * SS];
  __shared__ int sid[64];
  const int b = blockIdx.x;
  const int it = (b % 8) * ((nblocks + 7) / 8) + b / 8;  // XCD-contiguous item ranges
  if (it >= nitems) return;
  const int A = it_leaf[it];
  const int lane = threadIdx.x;
  const int t0 = pstart[A], nt = pcount[A];
  int tpad = 1;
  while (tpad < nt && tpad < 64) tpad <<= 1;
  const int nslice = 64 / tpad;
  const int rb = rng_off[A], re = rng_off[A + 1];
  // a leaf cut into several items accumulates atomically
  // (the items of a leaf are contiguous)
  const bool atomic = split && ((it > 0 && it_leaf[it - 1] == A) || (it + 1 < nitems && it_leaf[it + 1] == A));
  for (int tb = 0; tb < nt; tb += 64) {
    const int tl = tb + (lane % tpad), sl = lane / tpad;
    const bool active = tl < nt;
    const int tslot = t0 + tl;
    const double xt = active ? px[tslot] : 0.0, yt = active ? py[tslot] : 0.0, zt = active ? pz[tslot] : 0.0;
    T sum[OUT], comp[OUT];
#pragma unroll
    for (int k = 0; k < OUT; ++k) sum[k] = comp[k] = T(0);
    // source sequence: the own leaf (virtual range rb - 1, self pairs
    // masked), then the merged ranges rb .. re - 1; this item covers
    // it_n[it] sources from range it_r[it], offset it_o[it]
    int r = it_r[it], o = it_o[it], left = it_n[it];
    while (left > 0) {
      const bool has_self = r < rb;
      int filled = 0;
      __syncthreads();  // previous chunk consumed
      while (filled < 64 && left > 0 && r < re) {
        const int rs = (r < rb) ? t0 : rng_start[r];
        const int rl = (r < rb) ? nt : rng_len[r];
        const int take = min(min(64 - filled, rl - o), left);
        if (lane >= filled && lane < filled + take) {
          const int j = rs + o + (lane - filled);
          sx[lane] = px[j];
          sy[lane] = py[j];
          sz[lane] = pz[j];
          sid[lane] = (r < rb) ? j : -1;
#pragma unroll
          for (int k = 0; k < SS; ++k) sS[lane * SS + k] = (T)pS[(size_t)j * SS + k];
        }
        filled += take;
        o += take;
        left -= take;
        if (o == rl) {
          ++r;
          o = 0;
        }
      }
      if (filled == 0) break;  // defensive: sequence exhausted
      __syncthreads();  // chunk visible
      T part[OUT];
#pragma unroll
      for (int k = 0; k < OUT; ++k) part[k] = T(0);
      if (active) {
        if (has_self)
          p2p_chunk<T, true>(filled, sl, nslice, tslot, xt, yt, zt, sx, sy, sz, sS, sid, part);
        else
          p2p_chunk<T, false>(filled, sl, nslice, tslot, xt, yt, zt, sx, sy, sz, sS, sid, part);
      }
#pragma unroll
      for (int k = 0; k < OUT; ++k) {
        if constexpr (kKahan) {
          const T y = part[k] - comp[k];
          const T s2 = sum[k] + y;
          comp[k] = (s2 - sum[k]) - y;
          sum[k] = s2;
        } else {
          sum[k] += part[k];
        }
      }
    }
    // fold the slices of each target into slice 0
#pragma unroll
    for (int k = 0; k < OUT; ++k) {
      double v = kKahan ? (double)sum[k] - (double)comp[k] : (double)sum[k];
      for (int w = tpad; w < 64; w <<= 1) v += __shfl_xor(v, w, 64);
      if (active && sl == 0) {
        if (atomic) atomicAdd(&F[(size_t)tslot * OUT + k], v);
        else F[(size_t)tslot * OUT + k] = v;
      }
    }
  }
}

__global__ void k_fill_zero(long long n, double* __restrict__ a) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

void near_field(Tree& t, int mode, hipStream_t st) {
  const int ni = (int)t.n_p2p_items;
  PhaseTimer pt(t, kTP2P, st);
  if (t.p2p_split) {
    const long long m = (long long)t.n * OUT;
    k_fill_zero<<<(int)((m + 255) / 256), 256, 0, st>>>(m, t.Fp2p.get());
  }
  if (ni == 0) return;
  const int grid = ((ni + 7) / 8) * 8;
  auto launch = [&](auto kern) {
    kern<<<grid, 64, 0, st>>>(ni, grid, t.p2p_item_leaf.get(), t.p2p_item_r.get(), t.p2p_item_o.get(),
                              t.p2p_item_n.get(), t.pstart.get(), t.pcount.get(), t.rng_off.get(),
                              t.rng_start.get(), t.rng_len.get(), t.px.get(), t.py.get(), t.pz.get(), t.pS.get(),
                              t.p2p_split, t.Fp2p.get());
  };
  if (mode == 0) launch(k_p2p_leaf<double>);
  else launch(k_p2p_leaf<float>);
}

template <int P>
struct OrderTable {
  static void far(int order, Tree& t, int mode, int mask, hipStream_t st) {
    if (order == P) far_field_order<P>(t, mode, mask, st);
    else OrderTable<P + 1>::far(order, t, mode, mask, st);
  }
  static void l2p(int order, Tree& t, int mode, int mask, double* F, hipStream_t st) {
    if (order == P) l2p_final_order<P>(t, mode, mask, F, st);
    else OrderTable<P + 1>::l2p(order, t, mode, mask, F, st);
  }
  static void op(int order, int op, int n, const double* xyz, const double* in, int is, double* out, int os,
                 hipStream_t st) {
    if (order == P) op_batch_order<P>(op, n, xyz, in, is, out, os, st);
    else OrderTable<P + 1>::op(order, op, n, xyz, in, is, out, os, st);
  }
};
template <>
struct OrderTable<PMAX> {
  static void far(int, Tree&, int, int, hipStream_t) { throw Error{kUnsupported, "order not compiled"}; }
  static void l2p(int, Tree&, int, int, double*, hipStream_t) { throw Error{kUnsupported, "order not compiled"}; }
  static void op(int, int, int, const double*, const double*, int, double*, int, hipStream_t) {
    throw Error{kUnsupported, "order not compiled"};
  }
};

// Tree::compute_field_fmm (example/tree.cpp:243-271): zero F, M and L,
// then P2M, M2M, {M2L, P2P}, L2L, L2P and the final per-target sum.
void evaluate(Tree& t, int mode, int mask, double* F, hipStream_t st) {
  const int n = (int)t.n;
  const int nc = t.ncells;
  const size_t ms = (size_t)msize(t.order, SO), ls = (size_t)lsize(t.order, SO);
  t.M.alloc((size_t)nc * ms);
  t.L.alloc((size_t)nc * ls);
  t.Fp2p.alloc((size_t)n * OUT);
  t.pS.alloc((size_t)n * SS);
  {
    PhaseTimer pt(t, kTPrep, st);
    k_gather_S<<<nblk(n, 256), 256, 0, st>>>(n, t.perm.get(), t.user_S, t.pS.get());
    FMM_HIP(hipMemsetAsync(t.M.get(), 0, sizeof(double) * nc * ms, st));
    FMM_HIP(hipMemsetAsync(t.L.get(), 0, sizeof(double) * nc * ls, st));
  }
  OrderTable<PMIN>::far(t.order, t, mode, mask, st);
  if (mask & kPhaseP2P) near_field(t, mode, st);
  OrderTable<PMIN>::l2p(t.order, t, mode, mask, F, st);
}

void op_batch(int op, int order, int n, const double* xyz, const double* in, int in_stride, double* out,
              int out_stride, hipStream_t st) {
  // operand sizes of each operator; strides below them would read or
  // write past an item
  const int ms = msize(order, SO), ls = lsize(order, SO);
  int need_in = 0, need_out = 0;
  switch (op) {
    case kOpP2M: need_in = 1; need_out = ms; break;
    case kOpM2M: need_in = ms; need_out = ms; break;
    case kOpM2L: need_in = ms; need_out = ls; break;
    case kOpL2L: need_in = ls; need_out = ls; break;
    case kOpL2P: need_in = ls; need_out = OUT; break;
    case kOpM2P: need_in = ms; need_out = OUT; break;
    case kOpP2P: need_in = SS; need_out = OUT; break;
    case kOpP2MS: need_in = SS; need_out = ms; break;
    default: throw Error{kInvalidArgument, "unknown operator id"};
  }
  FMM_CHECK(in_stride >= need_in && out_stride >= need_out, kInvalidArgument,
            "operand stride smaller than the operator's array size (in " + std::to_string(need_in) + ", out " +
                std::to_string(need_out) + ")");
  if (op == kOpP2P) {
    k_op_p2p<<<nblk(n, 64), 64, 0, st>>>(n, xyz, in, in_stride, out, out_stride);
    return;
  }
  OrderTable<PMIN>::op(order, op, n, xyz, in, in_stride, out, out_stride, st);
}

void direct(int n, const double* pos, const double* S, const long long* targets, int nt, double* F,
            hipStream_t st) {
  if (nt > 0) k_direct<<<nt, 256, 0, st>>>(n, pos, S, targets, nt, F);
}

}  // namespace FMM_VARIANT_NS

// The table holds host function pointers; it exists in the host pass only.
#if !defined(__HIP_DEVICE_COMPILE__)
const Variant FMM_VARIANT_NAME = {FMM_VARIANT_NS::SO,   FMM_VARIANT_NS::OUT,       FMM_VARIANT_NS::PMIN,
                                  FMM_VARIANT_NS::PMAX, &FMM_VARIANT_NS::evaluate, &FMM_VARIANT_NS::op_batch,
                                  &FMM_VARIANT_NS::direct};
#endif

}  // namespace fmm
