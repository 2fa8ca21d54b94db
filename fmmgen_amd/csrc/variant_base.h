// Order-independent part of one operator variant: source gather, the P2P
// near field, the FP64 direct sum, order dispatch and the variant table.
#pragma once

#include <algorithm>

#include "learned.h"
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

// Sampled judge of the learned order: FP64 direct sums (self excluded) at
// sorted slots.  A block takes kJudgeTile targets (held in registers, so
// each source read serves kJudgeTile pairs) against one slice of
// kJudgeSlice sources, and adds its partial sums atomically into the
// zero-filled Fd[J][OUT] (for charges with the field: the field components;
// the potential component stays zero).
constexpr int kJudgeSlice = 16384;
constexpr int kJudgeTile = 8;

// One charge pair of the judge's field, q r^-3 (dx, dy, dz) (the field terms
// of the example's P2P, example/calculate.cpp:24-40; the judge's check and
// the field scale read only the field, so the potential is not summed).
// r^-3 = x^(-3/2), x = r^2, from the hardware estimate y of x^(-1/2):
// x^(-3/2) = y^3 (1 - e)^(-3/2) with e = 1 - x y^2, and the series
// 1 + 3e/2 + 15e^2/8 cubes the estimate's relative error.  A coincident
// pair (the target itself, x = 0) adds nothing: the factor's high word is
// cleared, which leaves a zero or subnormal factor times a zero
// displacement (one select instead of two for the whole FP64 value).
template <int NOUT>
__device__ __forceinline__ void judge_pair(double dx, double dy, double dz, double q, double (&acc)[NOUT]) {
  const double x = fma(dx, dx, fma(dy, dy, dz * dz));
  const double y = __builtin_amdgcn_rsq(x);
  const double y2 = y * y;
  const double e = fma(-x, y2, 1.0);
  const double qy3 = q * (y2 * y);
  const double t = qy3 * e;
  double s = fma(t * e, 1.875, fma(t, 1.5, qy3));
  s = __hiloint2double(x > 0.0 ? __double2hiint(s) : 0, __double2loint(s));
  constexpr int F0 = NOUT == 4 ? 1 : 0;
  acc[F0] = fma(s, dx, acc[F0]);
  acc[F0 + 1] = fma(s, dy, acc[F0 + 1]);
  acc[F0 + 2] = fma(s, dz, acc[F0 + 2]);
}
__global__ void __launch_bounds__(256) k_judge_direct(int n, const double* __restrict__ px,
                                                      const double* __restrict__ py, const double* __restrict__ pz,
                                                      const double* __restrict__ pS, const int* __restrict__ slots,
                                                      int J, double* __restrict__ Fd) {
  const int t0 = blockIdx.x * kJudgeTile;
  int ti[kJudgeTile];
  double tx[kJudgeTile], ty[kJudgeTile], tz[kJudgeTile], acc[kJudgeTile][OUT];
#pragma unroll
  for (int u = 0; u < kJudgeTile; ++u) {
    const int tt = min(t0 + u, J - 1);  // a short last tile repeats its last target (not stored)
    ti[u] = slots[tt];
    tx[u] = px[ti[u]];
    ty[u] = py[ti[u]];
    tz[u] = pz[ti[u]];
#pragma unroll
    for (int k = 0; k < OUT; ++k) acc[u][k] = 0.0;
  }
  const int b = blockIdx.y * kJudgeSlice, e = min(n, b + kJudgeSlice);
  for (int j = b + (int)threadIdx.x; j < e; j += blockDim.x) {
    const double xj = px[j], yj = py[j], zj = pz[j];
    double sj[SS];
#pragma unroll
    for (int k = 0; k < SS; ++k) sj[k] = pS[(size_t)j * SS + k];
#pragma unroll
    for (int u = 0; u < kJudgeTile; ++u) {
      double dx = tx[u] - xj, dy = ty[u] - yj, dz = tz[u] - zj;
      if constexpr (SO == 0 && (OUT == 4 || OUT == 3)) {
        judge_pair<OUT>(dx, dy, dz, sj[0], acc[u]);  // r = 0 only for the target itself
      } else {
        double s[SS];
        const bool self = j == ti[u];
#pragma unroll
        for (int k = 0; k < SS; ++k) s[k] = self ? 0.0 : sj[k];
        if (self) dx = 1.0;  // finite distance, zero strength
        fmmgen_ops::P2P<double>(dx, dy, dz, s, acc[u]);
      }
    }
  }
  // wave sums, block sums in LDS, one atomic per block and value
  __shared__ double red[4][kJudgeTile * OUT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < kJudgeTile; ++u)
#pragma unroll
    for (int k = 0; k < OUT; ++k) {
      double v = acc[u][k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][u * OUT + k] = v;
    }
  __syncthreads();
  const int i = threadIdx.x;
  if (i < kJudgeTile * OUT && t0 + i / OUT < J)
    atomicAdd(&Fd[(size_t)t0 * OUT + i], red[0][i] + red[1][i] + red[2][i] + red[3][i]);
}

// Judge check (one block): a target fails when |F - Fd| over the field
// components exceeds tol times the rms |Fd| of the sample; the leaf holding
// it is flagged.  acc = (sum |F - Fd|^2, sum |Fd|^2).
__global__ void k_judge_check(int J, const int* __restrict__ slots, const int* __restrict__ perm,
                              const int* __restrict__ leaf_of, const double* __restrict__ F,
                              const double* __restrict__ Fd, double tol, double* __restrict__ acc,
                              int* __restrict__ fail) {
  constexpr int FO = OUT == 4 ? 1 : 0;
  __shared__ double r1[256], r2[256];
  double e2 = 0.0, f2 = 0.0;
  for (int t = threadIdx.x; t < J; t += blockDim.x) {
    const size_t i = (size_t)perm[slots[t]];
    for (int k = FO; k < OUT; ++k) {
      const double d = F[i * OUT + k] - Fd[(size_t)t * OUT + k];
      e2 += d * d;
      f2 += Fd[(size_t)t * OUT + k] * Fd[(size_t)t * OUT + k];
    }
  }
  r1[threadIdx.x] = e2;
  r2[threadIdx.x] = f2;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double lim = tol * sqrt(r2[0] / (double)(J > 0 ? J : 1));
  if (threadIdx.x == 0) {
    acc[0] = r1[0];
    acc[1] = r2[0];
  }
  for (int t = threadIdx.x; t < J; t += blockDim.x) {
    const size_t i = (size_t)perm[slots[t]];
    double d2 = 0.0;
    for (int k = FO; k < OUT; ++k) {
      const double d = F[i * OUT + k] - Fd[(size_t)t * OUT + k];
      d2 += d * d;
    }
    if (sqrt(d2) > lim) {
      atomicAdd(fail, 1);
      fail[1 + leaf_of[slots[t]]] = 1;
    }
  }
}

__global__ void k_set_scalar(double* __restrict__ p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_fill_int(long long n, int* __restrict__ a) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// P2P near field (example/calculate.cpp:24-40, 208-216), one wavefront per
// work item: a segment of at most 8192 sources of one target leaf.
//
// The sources of a target leaf are its own particles (self pairs masked)
// and the merged particle ranges of its other P2P leaves; they are staged
// 128 at a time through LDS with coalesced loads.  A leaf whose sources span
// several items (a small leaf beside a dense cluster) adds its partial sums
// atomically into a zeroed output; other leaves store.  Targets go 64 to a
// round; a round of r targets runs as two power-of-two pieces (r = 40 as
// 32 + 8, r = 20 as 16 + 4), each piece's targets replicated over 64 / p
// lane slices that take alternate sources of every staged chunk, so fewer
// lanes idle than with one pow2(r) piece (40 of 64 lanes busy for r = 40).
//   T = double: FP64 throughout (the reference arithmetic).
//   T = float : dx = float(x_t - x_s) from FP64 positions (no cancellation
//               loss for close pairs), the generated P2P body in FP32, and
//               compensated (Kahan) accumulation of each chunk's partial
//               sum into the running total.
// The result is written in FP64 per sorted slot.
template <typename T, bool SELF>
__device__ __forceinline__ void p2p_pair(int j, int tslot, double xt, double yt, double zt, const double* sx,
                                         const double* sy, const double* sz, const T* sS, const int* sid,
                                         T (&part)[OUT]) {
  T dx = (T)(xt - sx[j]), dy = (T)(yt - sy[j]), dz = (T)(zt - sz[j]);
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

// Sources sl, sl + NS, sl + 2 NS, ... of a chunk of `filled`: a uniform trip
// count (filled / NS) unrolled with immediate LDS offsets, then one masked
// remainder pair.
template <typename T, bool SELF, int NS>
__device__ __forceinline__ void p2p_chunk(int filled, int sl, int tslot, double xt, double yt, double zt,
                                          const double* sx, const double* sy, const double* sz, const T* sS,
                                          const int* sid, T (&part)[OUT]) {
  const int full = filled / NS;
  const double *x = sx + sl, *y = sy + sl, *z = sz + sl;
  const T* S = sS + sl * SS;
  const int* id = sid + sl;
#pragma unroll 4
  for (int k = 0; k < full; ++k)
    p2p_pair<T, SELF>(k * NS, tslot, xt, yt, zt, x, y, z, S, id, part);
  if (sl < filled - full * NS) p2p_pair<T, SELF>(full * NS, tslot, xt, yt, zt, x, y, z, S, id, part);
}

template <typename T, bool SELF>
__device__ __forceinline__ void p2p_chunk_any(int nslice, int filled, int sl, int tslot, double xt, double yt,
                                              double zt, const double* sx, const double* sy, const double* sz,
                                              const T* sS, const int* sid, T (&part)[OUT]) {
  switch (nslice) {
#define FMM_P2P_NS(ns) \
  case ns:             \
    p2p_chunk<T, SELF, ns>(filled, sl, tslot, xt, yt, zt, sx, sy, sz, sS, sid, part); \
    break;
    FMM_P2P_NS(1) FMM_P2P_NS(2) FMM_P2P_NS(4) FMM_P2P_NS(8) FMM_P2P_NS(16) FMM_P2P_NS(32) FMM_P2P_NS(64)
#undef FMM_P2P_NS
  }
}

// Chunk partial sum into a running total: compensated (Kahan) in FP32,
// plain in FP64.
template <typename T>
__device__ __forceinline__ void p2p_accumulate(const T (&part)[OUT], T (&sum)[OUT], T (&comp)[OUT]) {
#pragma unroll
  for (int k = 0; k < OUT; ++k) {
    if constexpr (sizeof(T) == 4) {
      const T y = part[k] - comp[k];
      const T s2 = sum[k] + y;
      comp[k] = (s2 - sum[k]) - y;
      sum[k] = s2;
    } else {
      sum[k] += part[k];
    }
  }
}

// Sources staged in LDS per chunk (coalesced loads of up to 64 per pass);
// every lane sums its pairs of a chunk before one compensated fold.  128
// halves the chunk folds and LDS waits of 64 (FP32 near field 14.2 -> 13.1
// ms at 2^22 blobs); 256 gains nothing more and doubles the LDS.
constexpr int kP2PChunk = 128;

// One work item of the near field (the body of k_p2p_leaf).
template <typename T>
__device__ __forceinline__ void p2p_item(int it, int nitems, const int* __restrict__ it_leaf,
                                         const int* __restrict__ it_r, const int* __restrict__ it_o,
                                         const int* __restrict__ it_n, const int* __restrict__ pstart,
                                         const int* __restrict__ pcount, const int* __restrict__ rng_off,
                                         const int* __restrict__ rng_start, const int* __restrict__ rng_len,
                                         const double* __restrict__ px, const double* __restrict__ py,
                                         const double* __restrict__ pz, const double* __restrict__ pS, int split,
                                         double* __restrict__ F, double* sx, double* sy, double* sz, T* sS,
                                         int* sid) {
  constexpr bool kKahan = sizeof(T) == 4;
  const int lane = threadIdx.x;
  const int A = it_leaf[it];
  const int t0 = pstart[A], nt = pcount[A];
  const int rb = rng_off[A], re = rng_off[A + 1];
  // a leaf cut into several items accumulates atomically
  // (the items of a leaf are contiguous)
  const bool atomic = split && ((it > 0 && it_leaf[it - 1] == A) || (it + 1 < nitems && it_leaf[it + 1] == A));
  for (int tb = 0; tb < nt; tb += 64) {
    // this round's r targets as two power-of-two pieces: pa = the largest
    // power of two <= r (64 at most) and pb = r - pa rounded up to a power
    // of two (0 when r is one); piece targets are replicated over 64 / p
    // lane slices taking alternate sources of each chunk
    const int r = min(64, nt - tb);
    const int pa = 1 << (31 - __clz(r));
    const int pb = r > pa ? 1 << (32 - __clz(r - pa - 1)) : 0;
    const int sla = lane / pa, tia = lane % pa;
    const int slb = pb ? lane / pb : 0, tib = pb ? pa + lane % pb : r;
    const bool actb = tib < r;
    const int ta = t0 + tb + tia, tbs = t0 + tb + (actb ? tib : tia);
    const double xa = px[ta], ya = py[ta], za = pz[ta];
    const double xb = px[tbs], yb = py[tbs], zb = pz[tbs];
    T suma[OUT], compa[OUT], sumb[OUT], compb[OUT];
#pragma unroll
    for (int k = 0; k < OUT; ++k) suma[k] = compa[k] = sumb[k] = compb[k] = T(0);
    // source sequence: the own leaf (virtual range rb - 1, self pairs
    // masked), then the merged ranges rb .. re - 1; this item covers
    // it_n[it] sources from range it_r[it], offset it_o[it]
    int rg = it_r[it], o = it_o[it], left = it_n[it];
    while (left > 0) {
      const bool has_self = rg < rb;
      int filled = 0;
      __syncthreads();  // previous chunk consumed
      while (filled < kP2PChunk && left > 0 && rg < re) {
        const int rs = (rg < rb) ? t0 : rng_start[rg];
        const int rl = (rg < rb) ? nt : rng_len[rg];
        const int take = min(min(min(kP2PChunk - filled, 64), rl - o), left);
        if (lane < take) {  // lane stages slot filled + lane
          const int j = rs + o + lane, slot = filled + lane;
          sx[slot] = px[j];
          sy[slot] = py[j];
          sz[slot] = pz[j];
          sid[slot] = (rg < rb) ? j : -1;
#pragma unroll
          for (int k = 0; k < SS; ++k) sS[slot * SS + k] = (T)pS[(size_t)j * SS + k];
        }
        filled += take;
        o += take;
        left -= take;
        if (o == rl) {
          ++rg;
          o = 0;
        }
      }
      if (filled == 0) break;  // defensive: sequence exhausted
      __syncthreads();  // chunk visible
      T part[OUT];
#pragma unroll
      for (int k = 0; k < OUT; ++k) part[k] = T(0);
      if (has_self)
        p2p_chunk_any<T, true>(64 / pa, filled, sla, ta, xa, ya, za, sx, sy, sz, sS, sid, part);
      else
        p2p_chunk_any<T, false>(64 / pa, filled, sla, ta, xa, ya, za, sx, sy, sz, sS, sid, part);
      p2p_accumulate<T>(part, suma, compa);
      if (pb) {
#pragma unroll
        for (int k = 0; k < OUT; ++k) part[k] = T(0);
        if (actb) {
          if (has_self)
            p2p_chunk_any<T, true>(64 / pb, filled, slb, tbs, xb, yb, zb, sx, sy, sz, sS, sid, part);
          else
            p2p_chunk_any<T, false>(64 / pb, filled, slb, tbs, xb, yb, zb, sx, sy, sz, sS, sid, part);
        }
        p2p_accumulate<T>(part, sumb, compb);
      }
    }
    // fold each piece's slices; slice 0 of lane i holds target i of the piece
#pragma unroll
    for (int k = 0; k < OUT; ++k) {
      double va = kKahan ? (double)suma[k] - (double)compa[k] : (double)suma[k];
      for (int w = pa; w < 64; w <<= 1) va += __shfl_xor(va, w, 64);
      if (sla == 0) {
        if (atomic) atomicAdd(&F[(size_t)ta * OUT + k], va);
        else F[(size_t)ta * OUT + k] = va;
      }
      if (pb) {
        double vb = kKahan ? (double)sumb[k] - (double)compb[k] : (double)sumb[k];
        for (int w = pb; w < 64; w <<= 1) vb += __shfl_xor(vb, w, 64);
        if (slb == 0 && actb) {
          if (atomic) atomicAdd(&F[(size_t)tbs * OUT + k], vb);
          else F[(size_t)tbs * OUT + k] = vb;
        }
      }
    }
  }
}

// Persistent form: a grid of a few waves per SIMD takes work items
// from eight counters, one per XCD range of items (blocks are dealt to the
// XCDs round robin, and neighbouring leaves share their sources in that
// XCD's L2); a block whose range is exhausted takes items of the next
// XCDs'.  next[8] is zero before the launch.
template <typename T>
__global__ __launch_bounds__(64) void k_p2p_leaf(int nitems, int* __restrict__ next, const int* __restrict__ it_leaf,
                                                 const int* __restrict__ it_r, const int* __restrict__ it_o,
                                                 const int* __restrict__ it_n, const int* __restrict__ pstart,
                                                 const int* __restrict__ pcount, const int* __restrict__ rng_off,
                                                 const int* __restrict__ rng_start, const int* __restrict__ rng_len,
                                                 const double* __restrict__ px, const double* __restrict__ py,
                                                 const double* __restrict__ pz, const double* __restrict__ pS,
                                                 int split, double* __restrict__ F) {
  constexpr bool kKahan = sizeof(T) == 4;
  __shared__ double sx[kP2PChunk], sy[kP2PChunk], sz[kP2PChunk];
  __shared__ T sS[kP2PChunk * SS];
  __shared__ int sid[kP2PChunk];
  const int lane = threadIdx.x;
  const int per = (nitems + 7) / 8;
  for (int k = 0; k < 8; ++k) {
    const int x = (blockIdx.x + k) & 7;
    const int lo = x * per, hi = min(nitems, lo + per);
    for (;;) {
      int got = 0;
      if (lane == 0) got = atomicAdd(&next[x], 1);
      const int it = lo + __shfl(got, 0, 64);
      if (it >= hi) break;
      p2p_item<T>(it, nitems, it_leaf, it_r, it_o, it_n, pstart, pcount, rng_off, rng_start, rng_len, px, py, pz,
                  pS, split, F, sx, sy, sz, sS, sid);
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
  t.p2p_next.alloc(8);
  k_fill_int<<<1, 64, 0, st>>>(8, t.p2p_next.get());
  const int grid = std::min(ni, p2p_grid(t));
  auto launch = [&](auto kern) {
    kern<<<grid, 64, 0, st>>>(ni, t.p2p_next.get(), t.p2p_item_leaf.get(), t.p2p_item_r.get(), t.p2p_item_o.get(),
                              t.p2p_item_n.get(), t.pstart.get(), t.pcount.get(), t.rng_off.get(),
                              t.rng_start.get(), t.rng_len.get(), t.px.get(), t.py.get(), t.pz.get(), t.pS.get(),
                              t.p2p_split, t.Fp2p.get());
  };
  if (mode == 0) launch(k_p2p_leaf<double>);
  else launch(k_p2p_leaf<float>);
}

template <int P>
struct OrderTable {
  static void m2l(int order, Tree& t, int mode, const int* chunks, int nch, int ps, hipStream_t st) {
    if (order == P) m2l_class<P>(t, mode, chunks, nch, ps, st);
    else OrderTable<P + 1>::m2l(order, t, mode, chunks, nch, ps, st);
  }
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
  static void bh(int order, Tree& t, double* F, hipStream_t st) {
    if (order == P) bh_order<P>(t, F, st);
    else OrderTable<P + 1>::bh(order, t, F, st);
  }
};
template <>
struct OrderTable<PMAX> {
  static void m2l(int, Tree&, int, const int*, int, int, hipStream_t) {
    throw Error{kUnsupported, "order not compiled"};
  }
  static void far(int, Tree&, int, int, hipStream_t) { throw Error{kUnsupported, "order not compiled"}; }
  static void l2p(int, Tree&, int, int, double*, hipStream_t) { throw Error{kUnsupported, "order not compiled"}; }
  static void op(int, int, int, const double*, const double*, int, double*, int, hipStream_t) {
    throw Error{kUnsupported, "order not compiled"};
  }
  static void bh(int, Tree&, double*, hipStream_t) { throw Error{kUnsupported, "order not compiled"}; }
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
  t.ml_order = mode == 0 ? t.order : -1;  // FP64 M / L hold this evaluation's expansions
}

// Tree::compute_field_bh (example/tree.cpp:273-293): F overwritten; the
// multipoles by P2M and M2M as in the FMM, then every particle's walk
// (FP64).
void evaluate_bh(Tree& t, double* F, hipStream_t st) {
  const int n = (int)t.n;
  const int nc = t.ncells;
  const size_t ms = (size_t)msize(t.order, SO);
  t.M.alloc((size_t)nc * ms);
  t.pS.alloc((size_t)std::max(n, 1) * SS);
  {
    PhaseTimer pt(t, kTPrep, st);
    k_gather_S<<<nblk(n, 256), 256, 0, st>>>(n, t.perm.get(), t.user_S, t.pS.get());
  }
  OrderTable<PMIN>::far(t.order, t, 0, kPhaseP2M | kPhaseM2M, st);
  OrderTable<PMIN>::bh(t.order, t, F, st);
  t.ml_order = -1;  // L does not belong to this evaluation
}

// The caller's current source strengths in sorted order (t.pS).
void gather_sources(Tree& t, hipStream_t st) {
  const int n = (int)t.n;
  t.pS.alloc((size_t)std::max(n, 1) * SS);
  k_gather_S<<<nblk(n, 256), 256, 0, st>>>(n, t.perm.get(), t.user_S, t.pS.get());
}

// M2L of the learned evaluation: work items grouped by their supertile's
// order (the ranges read back with the far field's order), each group
// launched with the kernel of its own order on the expansions of the far
// field's order P (row strides of order P).
// The classes alternate between the evaluation stream and a second stream
// (they add into disjoint or atomically updated local rows), so that one
// class's last blocks overlap the next class's first.
void m2l_by_class(Tree& t, int mode, int P, hipStream_t st) {
  const int* start = t.class_start_h;
  if (!t.side) {
    FMM_HIP(hipStreamCreateWithFlags(&t.side, hipStreamNonBlocking));
    FMM_HIP(hipEventCreateWithFlags(&t.side_fork, hipEventDisableTiming));
    FMM_HIP(hipEventCreateWithFlags(&t.side_join, hipEventDisableTiming));
  }
  FMM_HIP(hipEventRecord(t.side_fork, st));
  FMM_HIP(hipStreamWaitEvent(t.side, t.side_fork, 0));
  int k = 0;
  for (int q = PMIN; q <= P && q < M2LSchedule::kMaxOrder; ++q) {
    const int nch = start[q + 1] - start[q];
    if (nch > 0)
      OrderTable<PMIN>::m2l(q, t, mode, t.sched.class_chunks.get() + start[q], nch, P, (k++ & 1) ? t.side : st);
  }
  FMM_HIP(hipEventRecord(t.side_join, t.side));
  FMM_HIP(hipStreamWaitEvent(st, t.side_join, 0));
}

constexpr int kJudgeTargets = 1024;
// A sample error above kJudgeDistrust times the tolerance: the network is
// not trusted on this input.
constexpr double kJudgeDistrust = 4.0;

// Judge sample and field scale: FP64 direct sums at kJudgeTargets
// stratified sorted slots (t.judge_Fd) and F_s, their rms |F| without the
// largest 0.1% (at least the largest: a near-coincident pair in the sample
// would dominate the rms), into t.lw[kLwScale] on the device.  Needs the
// gathered sources t.pS.
void field_scale(Tree& t, hipStream_t st) {
  constexpr int FO = OUT == 4 ? 1 : 0;
  const int n = (int)t.n;
  const int J = std::min(kJudgeTargets, n);
  if (t.judge_n != J) {  // once per tree
    std::vector<int> h(J);
    for (int i = 0; i < J; ++i) h[i] = (int)(((long long)(2 * i + 1) * n) / (2LL * J));
    t.judge_slot.alloc(std::max(J, 1));
    FMM_HIP(hipMemcpyAsync(t.judge_slot.get(), h.data(), sizeof(int) * J, hipMemcpyHostToDevice, st));
    FMM_HIP(hipStreamSynchronize(st));
    t.judge_n = J;
  }
  t.judge_Fd.alloc((size_t)std::max(J, 1) * OUT);
  learned_scalars(t);
  if (J > 0) {
    k_fill_zero<<<nblk((long long)J * OUT, 256), 256, 0, st>>>((long long)J * OUT, t.judge_Fd.get());
    const dim3 g((J + kJudgeTile - 1) / kJudgeTile, (n + kJudgeSlice - 1) / kJudgeSlice);
    k_judge_direct<<<g, 256, 0, st>>>(n, t.px.get(), t.py.get(), t.pz.get(), t.pS.get(), t.judge_slot.get(), J,
                                      t.judge_Fd.get());
  }
  field_scale_from_samples(t, J, OUT, FO, st);
}

// The learned cell orders for the tree's current sources (t.cell_q),
// without an evaluation: the order hints of the M2L tiling.
void predict_cell_orders(Tree& t, hipStream_t st) {
  const int nl = t.nleaves;
  gather_sources(t, st);
  field_scale(t, st);
  leaf_features(t, SS, st);
  t.leaf_y.alloc((size_t)std::max(nl, 1) * kMlpO);
  mlp_infer_f16(nl, t.leaf_feat.get(), t.mlp_half.get(), t.leaf_y.get(), nullptr, 0.f, 0.f, 0, 0, st);
  select_orders(t, t.leaf_y.get(), t.margin, std::max(PMIN, kMlpOrder0), t.order, false, st);
}

// Learned-order evaluation: the judge sample's FP64 direct sums and field
// scale, leaf features, FP16 network inference -> leaf orders (the
// cheapest whose predicted norm-wise error meets the tolerance) -> cell
// and supertile orders, mixed-precision far field at the largest order
// with every supertile truncated at its own order, FP32 P2P, then the
// judge's check at its kJudgeTargets targets.  The order selection runs on
// the device; the host waits once for the far field's order and once for
// the judge's verdict.  When the sample's relative field error e exceeds
// the tolerance, the leaves holding a target whose error exceeds
// tol * rms |F| rerun at the tree's order and the other leaves are
// re-allocated for 0.5 tol^2 / e (the predictions ran optimistic by about
// e / tol); all leaves run at the tree's order when e exceeds
// kJudgeDistrust * tol, or when the partial rerun still misses.
// mode 3: FP32 M2L, mode 4: BF16x3 M2L.
void evaluate_learned(Tree& t, int mode, double* F, hipStream_t st) {
  FMM_CHECK(t.has_model, kInvalidArgument, "no order model set on the tree");
  t.ml_order = -1;
  const int n = (int)t.n, nl = t.nleaves;
  const int fmode = mode == 4 ? 2 : 1;
  const int qmin = std::max(PMIN, kMlpOrder0);
  t.Fp2p.alloc((size_t)std::max(n, 1) * OUT);
  t.pS.alloc((size_t)std::max(n, 1) * SS);
  {
    PhaseTimer pt(t, kTPrep, st);
    k_gather_S<<<nblk(n, 256), 256, 0, st>>>(n, t.perm.get(), t.user_S, t.pS.get());
  }
  {
    PhaseTimer pt(t, kTJudge, st);
    field_scale(t, st);
  }
  const int J = t.judge_n;
  int peff;
  {
    PhaseTimer pt(t, kTOrder, st);
    leaf_features(t, SS, st);
    t.leaf_y.alloc((size_t)std::max(nl, 1) * kMlpO);
    mlp_infer_f16(nl, t.leaf_feat.get(), t.mlp_half.get(), t.leaf_y.get(), nullptr, 0.f, 0.f, 0, 0, st);
    select_orders(t, t.leaf_y.get(), t.margin, qmin, t.order, true, st);
    chunk_classes_async(t, st);
    peff = read_orders(t, st, t.class_start_h);
  }
  peff = std::max(std::min(peff, t.order), PMIN);
  struct Restore {  // the supertile orders apply to this evaluation only
    Tree& t;
    ~Restore() {
      t.use_tile_order = false;
      t.m2l_classes = nullptr;
    }
  } restore{t};
  t.use_tile_order = true;
  t.m2l_classes = &m2l_by_class;
  bool near_done = false;
  auto far = [&](int p) {
    OrderTable<PMIN>::far(p, t, fmode, kPhaseAll & ~kPhaseP2P, st);
    if (!near_done) {  // near field (FP32, compensated) once; reruns reuse it
      near_field(t, 1, st);
      near_done = true;
    }
    OrderTable<PMIN>::l2p(p, t, fmode, kPhaseAll, F, st);
  };
  far(peff);
  t.judge_acc.alloc(4);
  t.judge_fail.alloc((size_t)t.ncells + 1);
  int fails = 0, reruns = 0;
  double jerr = 0.0;
  auto check = [&]() {
    k_fill_int<<<nblk(t.ncells + 1, 256), 256, 0, st>>>(t.ncells + 1, t.judge_fail.get());
    k_judge_check<<<1, 256, 0, st>>>(J, t.judge_slot.get(), t.perm.get(), t.leaf_of.get(), F, t.judge_Fd.get(),
                                     t.tol, t.judge_acc.get(), t.judge_fail.get());
    double h[2];
    int nf = 0;
    FMM_HIP(hipMemcpyAsync(h, t.judge_acc.get(), sizeof(h), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(&nf, t.judge_fail.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    jerr = h[1] > 0 ? sqrt(h[0] / h[1]) : 0.0;
    return nf;
  };
  // orders after a change of the leaf orders: cells, supertiles, classes
  auto reorder = [&]() {
    orders_from_leaves(t, t.leaf_q.get(), t.cell_q.get(), t.tile_order.get(), st);
    chunk_classes_async(t, st);
    return std::max(std::min(read_orders(t, st, t.class_start_h), t.order), PMIN);
  };
  bool rerun = false;
  if (J > 0) {
    PhaseTimer pt(t, kTJudge, st);
    fails = check();
    rerun = jerr > t.tol;  // the sample's norm-wise error decides
  }
  if (rerun) {
    PhaseTimer pt(t, kTJudge, st);
    // a near miss: the failing targets' leaves rerun at the full order and
    // the other leaves are re-allocated for a tolerance lowered by the
    // factor the sample missed by, and half again; a wide miss: the network
    // is not trusted on this input and every leaf reruns at the full order
    const bool all = jerr > kJudgeDistrust * t.tol;
    if (all) {
      raise_all_leaves(t, t.order, st);
    } else {
      k_set_scalar<<<1, 64, 0, st>>>(t.lw.get() + kLwTol2, 0.5 * t.tol * (t.tol / jerr));
      k_set_scalar<<<1, 64, 0, st>>>(t.lw.get() + kLwRealloc, 1.0);
      allocate_orders(t, t.leaf_y.get(), kLwTol2, t.margin, qmin, t.order, t.leaf_q.get(), kLwPred2, false, st);
      raise_failing_leaves(t, t.order, st);
    }
    peff = reorder();
    far(peff);
    reruns = 1;
    check();
    if (!all && jerr > t.tol) {
      // the partial rerun still misses: the fixed-order method at the
      // tree's full order, every leaf
      raise_all_leaves(t, t.order, st);
      peff = reorder();
      far(peff);
      reruns = 2;
      check();
    }
  }
  t.lstats[0] = peff;
  t.lstats[1] = fails;
  t.lstats[2] = reruns;
  t.lstats[3] = jerr;
  t.lstats[4] = J;
  t.lstats_ready = false;  // the order statistics are computed when asked for
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
const Variant FMM_VARIANT_NAME = {FMM_VARIANT_NS::SO,
                                  FMM_VARIANT_NS::OUT,
                                  FMM_VARIANT_NS::PMIN,
                                  FMM_VARIANT_NS::PMAX,
                                  &FMM_VARIANT_NS::evaluate,
                                  &FMM_VARIANT_NS::evaluate_learned,
                                  &FMM_VARIANT_NS::field_scale,
                                  &FMM_VARIANT_NS::predict_cell_orders,
                                  &FMM_VARIANT_NS::gather_sources,
                                  &FMM_VARIANT_NS::op_batch,
                                  &FMM_VARIANT_NS::direct,
                                  &FMM_VARIANT_NS::evaluate_bh};
#endif

}  // namespace fmm
