// M2L as offset-grouped GEMMs on the matrix cores.
//
// For the pairs of one (tile, key) group the translation is the matrix
// T[n][m] = D_{n+m}(x), x = target centre - source centre, non-zero only for
// |n| + |m| <= p (fmmgen/expansions.py:144-172: L_n += sum_m D_{n+m} M_m).
// One block of W wavefronts owns a tile of 16 target cells and computes
//     L[n][t] += sum_m T[n][m] * M_src(t)[m]        (t = 0..15)
// sharing the staged operands in LDS; wave w issues the row blocks I = w
// mod W.
// with 16x16x4 MFMAs: A = 16 rows of T by 4 multipole slots, B = the same 4
// slots of the 16 gathered source multipoles, the 16x16 accumulator block
// holds 16 rows of the tile's local expansions.  Row blocks I and slot
// slices K whose lowest degrees already exceed p are never issued (the
// (I, K) plan is built at compile time from the degree layout), and the
// accumulators stay in registers across a chunk of up to kGroupsPerChunk
// groups of the tile; the chunks of a tile add their results atomically.
//
// A operand: lane l holds T[16 I + l%16][4 K + l/16] = D[idx(n + m)], read
// from the group's D vector in LDS through a per-(step, lane) index table;
// B operand: lane l holds M_src(l%16)[4 K + l/16], read from the group's 16
// source multipoles staged in LDS by coalesced row loads; accumulator register r of
// lane l holds target column l%16 and row 16 I + acc_row(l/16, r), where
// acc_row = 4 (l/16) + r for the f32 MFMA and (l/16) + 4 r for the f64 MFMA
// (measured on gfx950 with single-element operands).
//
// Mixed precision works on normalised expansions (cell half-width r):
//   Mhat_m = M_m / r^|m|,  Lhat_n = L_n r^(|n|+1),  Dhat_k = D_k r_A^(|k|+1)
// so Lhat_A,n = sum_m Dhat_{n+m} Mhat_B,m (r_B / r_A)^|m|; r_B / r_A is a
// power of two (2^(level_A - level_B)) applied exactly with ldexp.
#pragma once

#include <type_traits>
#include <utility>

#include "fmm_common.h"

namespace fmm {

template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// total degree of slot j of a source-order-0 array
__host__ __device__ constexpr int slot_degree(int j) {
  int d = 0;
  while (nterms(d) <= j) ++d;
  return d;
}

// Slot layouts of the GEMM's rows (L) and columns (M).
//   kGraded: every multi-index of degree <= p in the expansions' own graded
//     order (below).
//   kHarmonic (charges only): the multi-indices with n_z <= 1, (p + 1)^2
//     slots; degree d holds slots d^2 .. (d + 1)^2 - 1 in the graded order
//     restricted to n_z <= 1 (by a = n_y + n_z, then n_z), which keeps the
//     D slots one MFMA step reads within the 8-bit offsets of its table.
// Every D_k = d^k (1/R) is harmonic, D_{k+2z} = -D_{k+2x} - D_{k+2y}, so a
// multipole term M_m with m_z >= 2 acts on the local expansion exactly as
// -M_m placed at m - 2z + 2x and at m - 2z + 2y (the same degree: every
// truncation |n| + |m| <= p is kept).  Repeated, M_(a,b,c+2k) adds
// (-1)^k C(k,i) M to slot (a + 2i, b + 2(k - i), c), c = 0 or 1 (the fold);
// the translation of the folded multipole gives every L_n exactly, and the
// L_n with n_z >= 2 follow from the same identity,
//   L_(a,b,c+2k) = (-1)^k sum_i C(k,i) L_(a+2i, b+2(k-i), c)   (the unfold).
// The GEMM then has (p + 1)^2 rows and columns instead of C(p + 3, 3).
enum : int { kGraded = 0, kHarmonic = 1 };

__host__ __device__ constexpr int harm_degree(int j) {
  int d = 0;
  while ((d + 1) * (d + 1) <= j) ++d;
  return d;
}
__host__ __device__ constexpr int harm_size(int p) { return (p + 1) * (p + 1); }

// MFMA steps of one translation: row blocks of 16 L slots by slices of KW
// M slots, the steps whose lowest degrees sum to more than P dropped.
template <int P, int SO, int KW, int LAY = kGraded>
struct StepPlan {
  static_assert(LAY == kGraded || SO == 0, "the harmonic layout is for charges");
  static constexpr int LS = LAY == kHarmonic ? harm_size(P) : lsize(P, SO);
  static constexpr int MS = LAY == kHarmonic ? harm_size(P) : msize(P, SO);
  static constexpr int NI = (LS + 15) / 16, NK = (MS + KW - 1) / KW;
  static constexpr int OFF = nterms(SO - 1);  // M slot -> full index
  static constexpr int row_degree(int n) { return LAY == kHarmonic ? harm_degree(n) : slot_degree(n); }
  static constexpr int col_degree(int m) { return LAY == kHarmonic ? harm_degree(m) : slot_degree(m + OFF); }
  static constexpr bool valid(int I, int K) {
    return 16 * I < LS && KW * K < MS && row_degree(16 * I) + col_degree(KW * K) <= P;
  }
  static constexpr int count_steps() {
    int s = 0;
    for (int K = 0; K < NK; ++K)
      for (int I = 0; I < NI; ++I)
        if (valid(I, K)) s = s + 1;
    return s;
  }
  static constexpr int NST0 = count_steps();
  struct Tab {
    int nI[NK];
    int base[NK + 1];
    short sI[NST0 + 1], sK[NST0 + 1];  // row block and slice of step s (K-major order)
  };
  // valid row blocks of slice K are a prefix (rows are ordered by degree)
  static constexpr Tab make() {
    Tab t{};
    int s = 0;
    for (int K = 0; K < NK; ++K) {
      int c = 0;
      for (int I = 0; I < NI; ++I)
        if (valid(I, K)) c = I + 1;
      t.nI[K] = c;
      t.base[K] = s;
      for (int I = 0; I < c; ++I) {
        t.sI[s + I] = (short)I;
        t.sK[s + I] = (short)K;
      }
      s += c;
    }
    t.base[NK] = s;
    return t;
  }
  static constexpr Tab tab = make();
  static constexpr int NSTEP = tab.base[NK];
};
template <int P, int SO, int LAY = kGraded>
using M2LPlan = StepPlan<P, SO, 4, LAY>;  // FP32 / FP64 16x16x4 MFMAs
template <int P, int SO, int LAY = kGraded>
using M2LPlan32 = StepPlan<P, SO, 32, LAY>;  // BF16 16x16x32 MFMAs

// Multi-index of slot j of the graded layout and its inverse:
// idx(n) = Nterms(|n| - 1) + T(|n| - n_x) + (|n| - n_x - n_y), T(a) = a(a+1)/2.
struct MultiIdx {
  int x, y, z;
};
__host__ __device__ constexpr MultiIdx slot_multi(int j) {
  const int d = slot_degree(j);
  const int r = j - nterms(d - 1);
  int a = 0;
  while ((a + 1) * (a + 2) / 2 <= r) ++a;
  const int b = r - a * (a + 1) / 2;
  return {d - a, a - b, b};
}
__host__ __device__ constexpr int multi_slot(int x, int y, int z) {
  const int d = x + y + z, a = d - x;
  return nterms(d - 1) + a * (a + 1) / 2 + (d - x - y);
}
// Multi-index of every graded slot up to order 15, packed x | y << 8 |
// z << 16, for slot indices known only at run time (one load instead of
// slot_multi's two searches).
struct SlotMultiTab {
  unsigned v[nterms(15)];
  constexpr SlotMultiTab() : v{} {
    for (int j = 0; j < nterms(15); ++j) {
      const MultiIdx m = slot_multi(j);
      v[j] = (unsigned)m.x | ((unsigned)m.y << 8) | ((unsigned)m.z << 16);
    }
  }
};
__device__ __forceinline__ MultiIdx slot_multi_rt(int j) {
  constexpr SlotMultiTab tab{};
  const unsigned u = tab.v[j];
  return MultiIdx{(int)(u & 255u), (int)((u >> 8) & 255u), (int)(u >> 16)};
}

// The same for the harmonic layout (n_z <= 1): within degree d, slot
// d^2 + r holds a = n_y + n_z = (r + 1) / 2 and n_z = (r + 1) % 2 (r > 0).
__host__ __device__ constexpr MultiIdx harm_multi(int j) {
  const int d = harm_degree(j), r = j - d * d;
  const int a = (r + 1) / 2, z = r == 0 ? 0 : (r + 1) % 2;
  return MultiIdx{d - a, a - z, z};
}
__host__ __device__ constexpr int harm_slot(int x, int y, int z) {
  const int d = x + y + z, a = y + z;
  return d * d + (a == 0 ? 0 : 2 * a - 1 + z);
}
constexpr bool harm_layout_ok(int p) {
  for (int j = 0; j < harm_size(p); ++j) {
    const MultiIdx m = harm_multi(j);
    if (m.x < 0 || m.y < 0 || m.z < 0 || m.z > 1 || m.x + m.y + m.z != harm_degree(j) ||
        harm_slot(m.x, m.y, m.z) != j)
      return false;
  }
  return true;
}
static_assert(harm_layout_ok(15), "harmonic slots: a bijection onto the multi-indices with n_z <= 1");

template <int LAY>
__host__ __device__ constexpr MultiIdx lay_multi(int j, int off) {
  return LAY == kHarmonic ? harm_multi(j) : slot_multi(j + off);
}

// D-vector slot (full index - OFF) lane l reads at step (I, K): the entry
// T[16 I + l%16][4 K + l/16] = D[n + m], or -1 outside the expansion.
// (D is always in the graded layout.)
template <int P, int SO, int LAY = kGraded>
__host__ __device__ constexpr int step_lane_slot(int I, int K, int l) {
  using Plan = M2LPlan<P, SO, LAY>;
  constexpr int LS = Plan::LS, MS = Plan::MS, OFF = Plan::OFF;
  const int n = 16 * I + (l & 15), m = 4 * K + (l >> 4);
  if (n >= LS || m >= MS) return -1;
  const MultiIdx a = lay_multi<LAY>(n, 0), b = lay_multi<LAY>(m, OFF);
  if (a.x + a.y + a.z + b.x + b.y + b.z > P) return -1;
  return multi_slot(a.x + b.x, a.y + b.y, a.z + b.z) - OFF;
}

// Lowest D slot a step reads: the index table holds 8-bit offsets from it
// (255 marks an empty entry), the base is a compile-time immediate.
template <int P, int SO, int LAY = kGraded>
__host__ __device__ constexpr int step_base(int I, int K) {
  int b = 1 << 30;
  for (int l = 0; l < 64; ++l) {
    const int v = step_lane_slot<P, SO, LAY>(I, K, l);
    if (v >= 0 && v < b) b = v;
  }
  return b == (1 << 30) ? 0 : b;
}

template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};

__device__ __forceinline__ Vec4<double>::type mfma16x16x4(double a, double b, Vec4<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Vec4<float>::type mfma16x16x4(float a, float b, Vec4<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row (within the 16-row block) of accumulator register r of lane group q = l/16.
template <typename T>
__device__ __forceinline__ constexpr int acc_row(int q, int r) {
  return sizeof(T) == 8 ? q + 4 * r : 4 * q + r;
}

// Blocks are dealt to the 8 XCDs round robin; hand each XCD one contiguous
// range of tiles so neighbouring tiles (which share sources) share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = (nb + 7) / 8;
  const int x = b % 8, k = b / 8;
  return x * per + k;
}

// Steps of wave wv of a W-wave block, balanced so that every wave issues
// about NSTEP / W MFMAs.  Low-degree row blocks are valid for nearly every
// slice and high ones for few, so whole rows dealt round robin would leave
// wave 0 with about half of the work.  Rows I >= R0 are dealt round robin
// (I = wv mod W); the steps of the rows I < R0 (R0 the least for which no
// wave's round-robin share exceeds NSTEP / W) are taken in slice-major order
// and cut into consecutive runs that fill every wave up to its share.  A
// wave holds one accumulator per row it touches (rows[0 .. nrow-1]; a low
// row cut between waves is added into L by each).  Within the wave steps
// are ordered by slice K, so consecutive steps of one slice share their B
// operand read.
template <class Plan, int W, int WV>
struct WavePlan {
  struct Tab {
    short s[Plan::NSTEP + 1];
    int n, nrow;
    short rows[Plan::NI + 1];
    short slot[Plan::NI + 1];  // accumulator of row I (-1: none)
  };
  static constexpr int rowcount(int I) {
    int c = 0;
    for (int K = 0; K < Plan::NK; ++K)
      if (I < Plan::tab.nI[K]) ++c;
    return c;
  }
  static constexpr int high_share(int R0, int w) {
    int c = 0;
    for (int I = R0; I < Plan::NI; ++I)
      if (I % W == w) c += rowcount(I);
    return c;
  }
  static constexpr Tab make() {
    Tab t{};
    const int N = Plan::NSTEP;
    int R0 = 0;
    for (; R0 < Plan::NI; ++R0) {
      int mx = 0;
      for (int w = 0; w < W; ++w) mx = high_share(R0, w) > mx ? high_share(R0, w) : mx;
      if ((long long)mx * W <= N) break;
    }
    int lowN = 0;
    for (int K = 0; K < Plan::NK; ++K) lowN += Plan::tab.nI[K] < R0 ? Plan::tab.nI[K] : R0;
    // low-step run of wave w: [cut(w), cut(w + 1))
    auto cut = [&](int w) {
      if (w >= W) return lowN;
      long long h = 0;
      for (int v = 0; v < w; ++v) h += high_share(R0, v);
      long long c = (long long)N * w / W - h;
      return (int)(c < 0 ? 0 : (c > lowN ? lowN : c));
    };
    const int lb = cut(WV), le = cut(WV + 1) > lb ? cut(WV + 1) : lb;
    bool used[Plan::NSTEP + 1] = {};
    int idx = 0;
    for (int K = 0; K < Plan::NK; ++K)
      for (int I = 0; I < Plan::tab.nI[K] && I < R0; ++I, ++idx)
        if (idx >= lb && idx < le) used[Plan::tab.base[K] + I] = true;
    for (int K = 0; K < Plan::NK; ++K)
      for (int I = R0; I < Plan::tab.nI[K]; ++I)
        if (I % W == WV) used[Plan::tab.base[K] + I] = true;
    int n = 0;
    for (int st = 0; st < N; ++st)
      if (used[st]) t.s[n++] = (short)st;  // K-major table: slice, then row order
    t.n = n;
    for (int I = 0; I <= Plan::NI; ++I) t.slot[I] = -1;
    int nr = 0;
    for (int I = 0; I < Plan::NI; ++I) {
      bool has = false;
      for (int i = 0; i < n; ++i) has = has || Plan::tab.sI[t.s[i]] == I;
      if (has) {
        t.slot[I] = (short)nr;
        t.rows[nr++] = (short)I;
      }
    }
    t.nrow = nr;
    return t;
  }
  static constexpr Tab tab = make();
};

// Round-robin plan: wave wv takes the row blocks I = wv mod W, its steps
// in slice-major order (consecutive steps of one slice share the B read).
// Used below order 6, where a block's few steps leave little to balance.
template <class Plan, int W, int WV>
struct WavePlanRR {
  using Tab = typename WavePlan<Plan, W, WV>::Tab;
  static constexpr Tab make() {
    Tab t{};
    int n = 0;
    for (int st = 0; st < Plan::NSTEP; ++st)
      if (Plan::tab.sI[st] % W == WV) t.s[n++] = (short)st;
    t.n = n;
    int nr = 0;
    for (int I = 0; I <= Plan::NI; ++I) t.slot[I] = -1;
    for (int I = WV; I < Plan::NI; I += W) {
      t.slot[I] = (short)nr;
      t.rows[nr++] = (short)I;
    }
    t.nrow = n > 0 ? nr : 0;
    return t;
  }
  static constexpr Tab tab = make();
};

// The plan a kernel of order P uses (measured: the balanced plan is faster
// from order 7 up, equal at 6).
template <int P, int SO, int W, int WV, int LAY = kGraded>
using WavePlanFor = std::conditional_t<(P >= 6), WavePlan<M2LPlan<P, SO, LAY>, W, WV>,
                                       WavePlanRR<M2LPlan<P, SO, LAY>, W, WV>>;
template <int P, int SO, int W, int WV, int LAY = kGraded>
using WavePlan32For = WavePlan<M2LPlan32<P, SO, LAY>, W, WV>;

// Steps before wave wv in the BF16 plan's index table (4 words per lane
// and step: eight 16-bit D slots).
template <int P, int SO, int W, int WV, int LAY = kGraded>
constexpr int wave_steps32() {
  if constexpr (WV == 0) {
    return 0;
  } else {
    return wave_steps32<P, SO, W, WV - 1, LAY>() + WavePlan32For<P, SO, W, WV - 1, LAY>::tab.n;
  }
}

// The index table holds, per wave, its steps' 8-bit D offsets packed four
// to a 32-bit word (one LDS read serves four steps): word d of lane l of
// wave wv sits at wave_words<..., wv>() + 64 d + l and holds steps
// 4d .. 4d + 3 of the wave's order in bytes 0 .. 3.
template <int P, int SO, int W, int WV, int LAY = kGraded>
constexpr int wave_words() {
  if constexpr (WV == 0) {
    return 0;
  } else {
    return wave_words<P, SO, W, WV - 1, LAY>() + 64 * ((WavePlanFor<P, SO, W, WV - 1, LAY>::tab.n + 3) / 4);
  }
}

// Waves per block: the LDS tables of a chunk are shared by W waves.
template <int P, int SO, typename T, int LAY = kGraded>
constexpr int m2l_waves() {
  using Plan = M2LPlan<P, SO, LAY>;
  constexpr int lds = Plan::NSTEP * 64 + (int)sizeof(T) * (kTile * (Plan::MS | 1) + msize(P, SO) + 256);
  if (LAY == kHarmonic && P <= 8) return 2;  // few steps per group: fewer waves per barrier
  return lds > 80 * 1024 ? 8 : 4;
}

template <int W, typename F>
__device__ __forceinline__ void dispatch_wave(int w, F&& f) {
  static_for<W>([&](auto WV) {
    if (w == WV) f(WV);
  });
}

// tile_order (optional, per tile): the order q <= P the pairs of the tile's
// groups are truncated at (D entries of degree > q are zeroed).  chunk_list
// (optional): the chunks to run (else 0 .. nchunks-1).  dstride / mstride /
// lstride: row strides of the D table / of M / of L (the sizes of a higher
// order when the expansions are stored for one: every layout is a prefix).
// LAY: the layout of M's and L's rows (the folded multipoles and the
// unfolded local expansions for kHarmonic); D is graded either way.
template <int P, int SO, typename T, bool SCALED, int W, int LAY = kGraded>
__global__ __launch_bounds__(64 * W) void k_m2l_gemm(int nchunks, int nblocks, const int* __restrict__ tile_cells,
                                                     const int* __restrict__ chunk_tile,
                                                     const int* __restrict__ chunk_g0,
                                                     const int* __restrict__ chunk_g1,
                                                     const int* __restrict__ group_key,
                                                     const int* __restrict__ group_src,
                                                     const int* __restrict__ key_dlevel, const T* __restrict__ Dtab,
                                                     const unsigned* __restrict__ idxtab,
                                                     const signed char* __restrict__ tile_order,
                                                     const int* __restrict__ chunk_list, int dstride, int mstride,
                                                     int lstride, const T* __restrict__ M, T* __restrict__ L) {
  using Plan = M2LPlan<P, SO, LAY>;
  constexpr int MS = Plan::MS, LS = Plan::LS;
  constexpr int MD = msize(P, SO);       // D vector slots
  constexpr int MSP = (MS | 1);          // odd row stride of the staged multipoles: conflict-free B reads
  constexpr int NJ = (MS + 63) / 64;     // 64-slot segments of one column
  constexpr int CPW = kTile / W;         // columns staged by each wave
  constexpr int DJ = ((MD + 63) / 64 + W - 1) / W;  // D segments staged by each wave
  using V4 = typename Vec4<T>::type;
  constexpr int NWORDS = wave_words<P, SO, W, W, LAY>();
  __shared__ unsigned sidx[NWORDS];  // offsets from each step's base slot, four steps a word
  __shared__ T sD[MD + 256];                 // slots MD .. MD + 255 stay zero
  __shared__ T sB[kTile * MSP + 4];          // tail slots stay zero
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ci = xcd_tile(blockIdx.x, nblocks);
  if (ci >= nchunks) return;  // uniform over the block
  const int chunk = chunk_list ? chunk_list[ci] : ci;
  const int tile = chunk_tile[chunk];
  for (int i = tid; i < NWORDS; i += 64 * W) sidx[i] = idxtab[i];
  for (int i = tid; i < 256; i += 64 * W) sD[MD + i] = T(0);
  for (int i = tid; i < kTile * MSP + 4; i += 64 * W) sB[i] = T(0);
  const int col = lane & 15, kk = lane >> 4;
  const int g0 = chunk_g0[chunk], g1 = chunk_g1[chunk];
  const int q = tile_order ? (int)tile_order[tile] : P;
  const int qc = q < 0 ? 0 : (q > P ? P : q);

  dispatch_wave<W>(w, [&](auto WV) {
    constexpr int wv = WV;
    using WP = WavePlanFor<P, SO, W, wv, LAY>;
    constexpr int NIW = WP::tab.nrow;  // accumulators: one per row block touched
    constexpr int NSW = WP::tab.n;
    V4 acc[NIW > 0 ? NIW : 1];
#pragma unroll
    for (int i = 0; i < (NIW > 0 ? NIW : 1); ++i) acc[i] = V4{0, 0, 0, 0};
    T rD[DJ], rB[CPW][NJ];
    // degrees of the D and M slots this lane stages (fixed per lane)
    int degD[DJ], degB[NJ];
#pragma unroll
    for (int i = 0; i < DJ; ++i) {
      const int j = 64 * (wv + W * i) + lane;
      degD[i] = j < MD ? slot_degree(j + Plan::OFF) : 0;
    }
#pragma unroll
    for (int i = 0; i < NJ; ++i) {
      const int j = 64 * i + lane;
      degB[i] = j < MS ? Plan::col_degree(j) : 0;
    }
    // loads are issued unconditionally (empty columns read cell 0, slots
    // past the row read its last slot) and masked afterwards: no branch
    // around a load
    auto prefetch = [&](int g) {
      const int key = group_key[g];
#pragma unroll
      for (int i = 0; i < DJ; ++i) {
        const int j = 64 * (wv + W * i) + lane;
        const T v = Dtab[(size_t)key * dstride + (j < MD ? j : MD - 1)];
        rD[i] = j < MD ? v : T(0);
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const int src = group_src[(size_t)g * kTile + wv + W * c];
        const T* row = M + (size_t)(src >= 0 ? src : 0) * mstride;
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
          const int j = 64 * i + lane;
          const T v = row[j < MS ? j : MS - 1];
          rB[c][i] = (src >= 0 && j < MS) ? v : T(0);
        }
      }
    };
    if (g0 < g1) prefetch(g0);
    for (int g = g0; g < g1; ++g) {
      const int dl = SCALED ? key_dlevel[group_key[g]] : 0;
      __syncthreads();  // previous group finished reading sD / sB
#pragma unroll
      for (int i = 0; i < DJ; ++i) {
        const int j = 64 * (wv + W * i) + lane;
        if (j < MD) sD[j] = degD[i] <= qc ? rD[i] : T(0);
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
          const int j = 64 * i + lane;
          if (j < MS) {
            T v = rB[c][i];
            if constexpr (SCALED) v = ldexpf(v, __mul24(dl, degB[i]));  // exact: a power of two
            sB[(wv + W * c) * MSP + j] = v;
          }
        }
      __syncthreads();  // staged operands visible to every wave
      if (g + 1 < g1) prefetch(g + 1);  // in flight during the MFMAs below
      // batches of kBatch steps: all operand reads of a batch, then its
      // MFMAs (a tile of a lower order than P has its D entries of higher
      // degree zeroed; the learned evaluation launches each tile order with
      // the kernel of that order, so no step is wasted there)
      constexpr int kBatch = 16;  // a multiple of 4: a batch starts a packed word
      constexpr int WW = wave_words<P, SO, W, wv, LAY>();
      unsigned word = 0;
      static_for<(NSW + kBatch - 1) / kBatch>([&](auto bb) {
        {
          T av[kBatch], bv[kBatch];
          static_for<kBatch>([&](auto u) {
            constexpr int i = bb * kBatch + u;
            if constexpr (i < NSW) {
              constexpr int s = WP::tab.s[i];
              constexpr int K = Plan::tab.sK[s], I = Plan::tab.sI[s];
              constexpr int base = step_base<P, SO, LAY>(I, K);
              if constexpr (i % 4 == 0) word = sidx[WW + 64 * (i / 4) + lane];
              const int off = (int)((word >> (8 * (i % 4))) & 255u);
              av[u] = sD[base + off];
              // an empty entry (offset 255) reads a zero slot when the
              // step's base is within 255 of MD, else it is masked here
              if constexpr (MD - base > 255) av[u] = off == 255 ? T(0) : av[u];
              constexpr int KP = (u > 0) ? Plan::tab.sK[WP::tab.s[i > 0 ? i - 1 : 0]] : -1;
              if constexpr (K == KP) bv[u] = bv[u > 0 ? u - 1 : 0];
              else bv[u] = sB[col * MSP + 4 * K + kk];
            }
          });
          static_for<kBatch>([&](auto u) {
            constexpr int i = bb * kBatch + u;
            if constexpr (i < NSW) {
              constexpr int s = WP::tab.s[i];
              constexpr int I = Plan::tab.sI[s];
              constexpr int A = WP::tab.slot[I];
              acc[A] = mfma16x16x4(av[u], bv[u], acc[A]);
            }
          });
        }
      });
    }
    const int cell = tile_cells[tile * kTile + col];
    if (cell < 0) return;
    T* Lc = L + (size_t)cell * lstride;
    static_for<NIW>([&](auto ii) {
      constexpr int I = WP::tab.rows[ii];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = 16 * I + acc_row<T>(kk, r);
        if (n < LS) atomicAdd(&Lc[n], acc[ii][r]);  // chunks of one tile add into the same L
      }
    });
  });
}

// Fold and unfold as term tables (see the harmonic layout above): for each
// harmonic slot j, the graded multipole slots that fold onto it and their
// weights (-1)^k C(k,i); for each graded local slot n, the harmonic rows it
// is rebuilt from.  Built at compile time per order.
template <int P>
struct HarmTables {
  static constexpr int HS = harm_size(P), GS = nterms(P);
  static constexpr int binom(int n, int k) {
    int c = 1;
    for (int i = 0; i < k; ++i) c = c * (n - i) / (i + 1);
    return c;
  }
  static constexpr int count_fold() {
    int t = 0;
    for (int j = 0; j < HS; ++j) {
      const MultiIdx h = harm_multi(j);
      for (int k = 0; 2 * k <= h.x + h.y; ++k)
        for (int a = 0; a <= k; ++a)
          if (h.x - 2 * a >= 0 && h.y - 2 * (k - a) >= 0) ++t;
    }
    return t;
  }
  static constexpr int count_unfold() {
    int t = 0;
    for (int n = 0; n < GS; ++n) t += slot_multi(n).z / 2 + 1;
    return t;
  }
  static constexpr int NF = count_fold(), NU = count_unfold();
  short fold_off[HS + 1], fold_src[NF];
  signed char fold_w[NF];
  short unf_off[GS + 1], unf_src[NU];
  signed char unf_w[NU];
  constexpr HarmTables() : fold_off{}, fold_src{}, fold_w{}, unf_off{}, unf_src{}, unf_w{} {
    int t = 0;
    for (int j = 0; j < HS; ++j) {
      fold_off[j] = (short)t;
      const MultiIdx h = harm_multi(j);
      for (int k = 0; 2 * k <= h.x + h.y; ++k)
        for (int a = 0; a <= k; ++a) {
          const int sx = h.x - 2 * a, sy = h.y - 2 * (k - a);
          if (sx >= 0 && sy >= 0) {
            fold_src[t] = (short)multi_slot(sx, sy, h.z + 2 * k);
            fold_w[t] = (signed char)((k & 1) ? -binom(k, a) : binom(k, a));
            ++t;
          }
        }
    }
    fold_off[HS] = (short)t;
    t = 0;
    for (int n = 0; n < GS; ++n) {
      unf_off[n] = (short)t;
      const MultiIdx m = slot_multi(n);
      const int k = m.z / 2, c = m.z % 2;
      for (int a = 0; a <= k; ++a) {
        unf_src[t] = (short)harm_slot(m.x + 2 * a, m.y + 2 * (k - a), c);
        unf_w[t] = (signed char)((k & 1) ? -binom(k, a) : binom(k, a));
        ++t;
      }
    }
    unf_off[GS] = (short)t;
  }
};

// Fold and unfold kernels: a block of 4 waves stages the term table in LDS
// once, then each wave takes kHarmCellsPerWave cells in turn, copying the
// cell's source row into LDS with coalesced loads and summing its terms (FP64)
// from LDS.
constexpr int kHarmCellsPerWave = 4;

// Fold: Mh[cell][j] for the (P + 1)^2 harmonic slots j from the graded
// multipole M[cell] (stride ms).
template <int P, typename T>
__global__ __launch_bounds__(256) void k_harm_fold(int nc, const T* __restrict__ M, int ms, T* __restrict__ Mh,
                                                   int hs) {
  using H = HarmTables<P>;
  constexpr H tab{};
  __shared__ short s_off[H::HS + 1], s_src[H::NF];
  __shared__ signed char s_w[H::NF];
  __shared__ T s_row[4][H::GS];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  for (int i = tid; i <= H::HS; i += 256) s_off[i] = tab.fold_off[i];
  for (int i = tid; i < H::NF; i += 256) {
    s_src[i] = tab.fold_src[i];
    s_w[i] = tab.fold_w[i];
  }
  const int c0 = (blockIdx.x * 4 + w) * kHarmCellsPerWave;
  for (int c = 0; c < kHarmCellsPerWave; ++c) {
    const int cell = c0 + c;
    __syncthreads();  // tables staged; the previous cell's row read
    if (cell < nc)
      for (int j = lane; j < H::GS; j += 64) s_row[w][j] = M[(size_t)cell * ms + j];
    __syncthreads();
    if (cell < nc)
      for (int j = lane; j < H::HS; j += 64) {
        double v = 0.0;
        for (int t = s_off[j]; t < s_off[j + 1]; ++t) v += (double)s_w[t] * (double)s_row[w][s_src[t]];
        Mh[(size_t)cell * hs + j] = (T)v;
      }
  }
}

// Unfold: L[cell][n] = the graded local expansion rebuilt from the harmonic
// rows Lh[cell] (stride hs), stored into every slot of every cell (the M2L
// is the first writer of L in an evaluation).
template <int P, typename T>
__global__ __launch_bounds__(256) void k_harm_unfold(int nc, const T* __restrict__ Lh, int hs, T* __restrict__ L,
                                                     int ls) {
  using H = HarmTables<P>;
  constexpr H tab{};
  __shared__ short s_off[H::GS + 1], s_src[H::NU];
  __shared__ signed char s_w[H::NU];
  __shared__ T s_row[4][H::HS];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  for (int i = tid; i <= H::GS; i += 256) s_off[i] = tab.unf_off[i];
  for (int i = tid; i < H::NU; i += 256) {
    s_src[i] = tab.unf_src[i];
    s_w[i] = tab.unf_w[i];
  }
  const int c0 = (blockIdx.x * 4 + w) * kHarmCellsPerWave;
  for (int c = 0; c < kHarmCellsPerWave; ++c) {
    const int cell = c0 + c;
    __syncthreads();
    if (cell < nc)
      for (int j = lane; j < H::HS; j += 64) s_row[w][j] = Lh[(size_t)cell * hs + j];
    __syncthreads();
    if (cell < nc)
      for (int n = lane; n < H::GS; n += 64) {
        double v = 0.0;
        for (int t = s_off[n]; t < s_off[n + 1]; ++t) v += (double)s_w[t] * (double)s_row[w][s_src[t]];
        L[(size_t)cell * ls + n] = (T)v;  // the first writer of L: every slot of every cell
      }
  }
}

// ---------------------------------------------------------------- BF16x3
// Split-precision M2L: each FP32 operand x is carried as hi = bf16(x) plus
// lo = bf16(x - hi), and the product as hi*hi + hi*lo + lo*hi on the BF16
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  Operand
// layout (measured on gfx950): A[l%16][8 (l/16) + i], B[8 (l/16) + j][l%16],
// i, j = 0..7; the accumulator is laid out as for the f32 16x16 MFMA.
//
// Organised as k_m2l_gemm: W waves per block share a tile chunk's staged
// operands (the group's D vector split once into packed hi | lo, the 16
// source columns in FP32) and deal the steps of the 16x16x32 plan by the
// balanced wave plan.  Each lane's eight D slots of each of its steps are
// held in registers (four words of 16-bit slots, loaded once per block),
// so a step costs eight D reads from LDS and no index reads.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// One FP32 value as packed (hi | lo << 16) bf16 halves.
__device__ __forceinline__ unsigned pack_split(float x) {
  const __bf16 h = (__bf16)x;
  const __bf16 l = (__bf16)(x - (float)h);
  return (unsigned)__builtin_bit_cast(unsigned short, h) | ((unsigned)__builtin_bit_cast(unsigned short, l) << 16);
}

// Eight packed values -> hi and lo operand vectors (two byte permutes per pair).
__device__ __forceinline__ void unpack8(const unsigned (&w)[8], bf16x8& hi, bf16x8& lo) {
  u32x4 h, l;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    h[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], 0x05040100u);
    l[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], 0x07060302u);
  }
  hi = __builtin_bit_cast(bf16x8, h);
  lo = __builtin_bit_cast(bf16x8, l);
}

__device__ __forceinline__ void split_bf16(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)v[i];
    hi[i] = h;
    lo[i] = (__bf16)(v[i] - (float)h);
  }
}

// Waves per block of the BF16 kernel: a wave's index words (4 per step)
// stay within about 40 registers.
template <int P, int SO, int LAY = kGraded>
constexpr int m2l_waves32() {
  if (LAY == kHarmonic && P <= 8) return 2;  // few steps per group: fewer waves per barrier
  return M2LPlan32<P, SO, LAY>::NSTEP > 40 ? 8 : 4;
}

template <int P, int SO, int W, int LAY = kGraded>
__global__ __launch_bounds__(64 * W) void k_m2l_bf16x3(int nchunks, int nblocks, const int* __restrict__ tile_cells,
                                                       const int* __restrict__ chunk_tile,
                                                       const int* __restrict__ chunk_g0,
                                                       const int* __restrict__ chunk_g1,
                                                       const int* __restrict__ group_key,
                                                       const int* __restrict__ group_src,
                                                       const int* __restrict__ key_dlevel,
                                                       const float* __restrict__ Dtab,
                                                       const unsigned* __restrict__ idxtab,
                                                       const signed char* __restrict__ tile_order,
                                                       const int* __restrict__ chunk_list, int dstride,
                                                       int mstride, int lstride, const float* __restrict__ M,
                                                       float* __restrict__ L) {
  using Plan = M2LPlan32<P, SO, LAY>;
  constexpr int MS = Plan::MS, LS = Plan::LS, NK = Plan::NK;
  constexpr int MD = msize(P, SO);                 // D vector slots
  constexpr int RS = ((32 * NK + 3) / 4) * 4 + 4;  // row stride: 16-byte aligned reads, zero tail
  constexpr int NJ = (MS + 63) / 64;
  constexpr int CPW = kTile / W;
  constexpr int DJ = ((MD + 63) / 64 + W - 1) / W;
  using V4 = typename Vec4<float>::type;
  __shared__ unsigned sD[MD + 1];  // D split once per group: packed (hi | lo << 16); slot MD stays zero
  __shared__ __attribute__((aligned(16))) float sB[kTile * RS];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ci = xcd_tile(blockIdx.x, nblocks);
  if (ci >= nchunks) return;
  const int chunk = chunk_list ? chunk_list[ci] : ci;
  const int tile = chunk_tile[chunk];
  if (tid == 0) sD[MD] = 0u;
  for (int i = tid; i < kTile * RS; i += 64 * W) sB[i] = 0.f;
  const int col = lane & 15, kk = lane >> 4;
  const int g0 = chunk_g0[chunk], g1 = chunk_g1[chunk];
  const int q = tile_order ? (int)tile_order[tile] : P;
  const int qc = q < 0 ? 0 : (q > P ? P : q);

  dispatch_wave<W>(w, [&](auto WV) {
    constexpr int wv = WV;
    using WP = WavePlan32For<P, SO, W, wv, LAY>;
    constexpr int NIW = WP::tab.nrow;
    constexpr int NSW = WP::tab.n;
    constexpr int WS = wave_steps32<P, SO, W, wv, LAY>();
    u32x4 ix[NSW > 0 ? NSW : 1];
    static_for<NSW>([&](auto i) {
      ix[i] = *reinterpret_cast<const u32x4*>(idxtab + ((size_t)(WS + i) * 64 + lane) * 4);
    });
    V4 acc[NIW > 0 ? NIW : 1];
#pragma unroll
    for (int i = 0; i < (NIW > 0 ? NIW : 1); ++i) acc[i] = V4{0, 0, 0, 0};
    float rD[DJ], rB[CPW][NJ];
    int degD[DJ], degB[NJ];  // degrees of the D and M slots this lane stages
#pragma unroll
    for (int i = 0; i < DJ; ++i) {
      const int j = 64 * (wv + W * i) + lane;
      degD[i] = j < MD ? slot_degree(j + Plan::OFF) : 0;
    }
#pragma unroll
    for (int i = 0; i < NJ; ++i) {
      const int j = 64 * i + lane;
      degB[i] = j < MS ? Plan::col_degree(j) : 0;
    }
    // loads are issued unconditionally and masked afterwards
    auto prefetch = [&](int g) {
      const int key = group_key[g];
#pragma unroll
      for (int i = 0; i < DJ; ++i) {
        const int j = 64 * (wv + W * i) + lane;
        const float v = Dtab[(size_t)key * dstride + (j < MD ? j : MD - 1)];
        rD[i] = j < MD ? v : 0.f;
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const int src = group_src[(size_t)g * kTile + wv + W * c];
        const float* row = M + (size_t)(src >= 0 ? src : 0) * mstride;
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
          const int j = 64 * i + lane;
          const float v = row[j < MS ? j : MS - 1];
          rB[c][i] = (src >= 0 && j < MS) ? v : 0.f;
        }
      }
    };
    if (g0 < g1) prefetch(g0);
    for (int g = g0; g < g1; ++g) {
      const int dl = key_dlevel[group_key[g]];
      __syncthreads();  // previous group finished reading sD / sB
#pragma unroll
      for (int i = 0; i < DJ; ++i) {
        const int j = 64 * (wv + W * i) + lane;
        if (j < MD) sD[j] = degD[i] <= qc ? pack_split(rD[i]) : 0u;
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
          const int j = 64 * i + lane;
          if (j < MS) sB[(wv + W * c) * RS + j] = ldexpf(rB[c][i], __mul24(dl, degB[i]));  // exact
        }
      __syncthreads();  // staged operands visible to every wave
      if (g + 1 < g1) prefetch(g + 1);  // in flight during the MFMAs below
      bf16x8 bH, bL;
      static_for<NSW>([&](auto i) {
        constexpr int s = WP::tab.s[i];
        constexpr int K = Plan::tab.sK[s], I = Plan::tab.sI[s];
        constexpr int KP = i > 0 ? Plan::tab.sK[WP::tab.s[i > 0 ? i - 1 : 0]] : -1;
        if constexpr (K != KP) {  // consecutive steps of one slice share the B operand
          const float* src = sB + col * RS + 32 * K + 8 * kk;
          const float4 v0 = *reinterpret_cast<const float4*>(src);
          const float4 v1 = *reinterpret_cast<const float4*>(src + 4);
          const float bv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
          split_bf16(bv, bH, bL);
        }
        unsigned aw[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          aw[2 * e] = sD[ix[i][e] & 0xffffu];
          aw[2 * e + 1] = sD[ix[i][e] >> 16];
        }
        bf16x8 aH, aL;
        unpack8(aw, aH, aL);
        constexpr int A = WP::tab.slot[I];
        acc[A] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH, bH, acc[A], 0, 0, 0);
        acc[A] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH, bL, acc[A], 0, 0, 0);
        acc[A] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aL, bH, acc[A], 0, 0, 0);
      });
    }
    const int cell = tile_cells[tile * kTile + col];
    if (cell < 0) return;
    float* Lc = L + (size_t)cell * lstride;
    static_for<NIW>([&](auto ii) {
      constexpr int I = WP::tab.rows[ii];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = 16 * I + acc_row<float>(kk, r);
        if (n < LS) atomicAdd(&Lc[n], acc[ii][r]);  // chunks of one tile add into the same L
      }
    });
  });
}

// Host side: the BF16 plan's table for a W-wave block, per wave step (wave
// order) and lane four words holding the eight 16-bit D slots
// A[l%16][8 (l/16) + i] reads (msize(P, SO), a zero slot, for an empty
// entry); rows and columns in the layout LAY, D graded.
template <int P, int SO, int W, int WV, int LAY>
void fill_wave_index32(std::vector<unsigned>& t) {
  if constexpr (WV < W) {
    using WP = WavePlan32For<P, SO, W, WV, LAY>;
    using Plan = M2LPlan32<P, SO, LAY>;
    const size_t s0 = (size_t)wave_steps32<P, SO, W, WV, LAY>();
    for (int i = 0; i < WP::tab.n; ++i) {
      const int st = WP::tab.s[i];
      const int I = Plan::tab.sI[st], K = Plan::tab.sK[st];
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          const int n = 16 * I + (l & 15), m = 32 * K + 8 * (l >> 4) + e;
          unsigned u = (unsigned)msize(P, SO);
          if (n < Plan::LS && m < Plan::MS) {
            const MultiIdx a = lay_multi<LAY>(n, 0), b = lay_multi<LAY>(m, Plan::OFF);
            if (a.x + a.y + a.z + b.x + b.y + b.z <= P)
              u = (unsigned)(multi_slot(a.x + b.x, a.y + b.y, a.z + b.z) - Plan::OFF);
          }
          unsigned& word = t[((s0 + i) * 64 + l) * 4 + e / 2];
          word = (e % 2 == 0) ? ((word & 0xffff0000u) | u) : ((word & 0xffffu) | (u << 16));
        }
    }
    fill_wave_index32<P, SO, W, WV + 1, LAY>(t);
  }
}

template <int P, int SO, int W, int LAY = kGraded>
std::vector<unsigned> m2l_index_table32() {
  std::vector<unsigned> t((size_t)M2LPlan32<P, SO, LAY>::NSTEP * 64 * 4, 0u);
  fill_wave_index32<P, SO, W, 0, LAY>(t);
  return t;
}

// Host side: the index table of a W-wave block (layout at wave_words):
// 8-bit offsets of each lane's D slot from its step's base slot, 255 for
// an empty entry.
template <int P, int SO, int W, int WV, int LAY>
void fill_wave_index(std::vector<unsigned>& t) {
  if constexpr (WV < W) {
    using WP = WavePlanFor<P, SO, W, WV, LAY>;
    using Plan = M2LPlan<P, SO, LAY>;
    const size_t w0 = (size_t)wave_words<P, SO, W, WV, LAY>();
    for (int i = 0; i < WP::tab.n; ++i) {
      const int st = WP::tab.s[i];
      const int I = Plan::tab.sI[st], K = Plan::tab.sK[st];
      const int base = step_base<P, SO, LAY>(I, K);
      for (int l = 0; l < 64; ++l) {
        const int v = step_lane_slot<P, SO, LAY>(I, K, l);
        if (v >= 0 && v - base > 254) throw Error{kUnsupported, "M2L step spans more than 255 D slots"};
        const unsigned b = v < 0 ? 255u : (unsigned)(v - base);
        unsigned& word = t[w0 + 64 * (size_t)(i / 4) + l];
        word = (word & ~(255u << (8 * (i % 4)))) | (b << (8 * (i % 4)));
      }
    }
    fill_wave_index<P, SO, W, WV + 1, LAY>(t);
  }
}

template <int P, int SO, int W, int LAY = kGraded>
std::vector<unsigned> m2l_index_table() {
  std::vector<unsigned> t((size_t)wave_words<P, SO, W, W, LAY>(), 0xffffffffu);
  fill_wave_index<P, SO, W, 0, LAY>(t);
  return t;
}

}  // namespace fmm
