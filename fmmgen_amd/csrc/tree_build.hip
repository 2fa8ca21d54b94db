// Octree build and dual-traversal interaction lists on the GPU.
//
// Reproduces example/tree.cpp:130-233 (build_tree) and
// example/calculate.cpp:81-133 (interact_dehnen_lazy) bit for bit on the
// geometry, with a canonical cell numbering:
//   * root centre = sequential FP64 mean (one lane per axis, in index
//     order, as example/tree.cpp:145-156 sums), half-width = 1.001 x the
//     largest per-axis deviation (order-free max reduction);
//   * a cell splits iff it holds >= ncrit particles; only non-empty octants
//     become children; child centre = parent centre +- r/2 by the same FP64
//     operations as example/tree.cpp:95-98, octant bits x->1, y->2, z->4
//     from strict '>' compares (example/tree.cpp:193);
//   * the subdivision runs one level at a time: an octant histogram per
//     split cell, a scan that numbers the children (level-major, octant-path
//     order inside a level), and a per-particle 64-bit path key; one stable
//     radix sort of the keys puts every cell's particles in one contiguous
//     range with increasing caller index inside a leaf (the order the
//     reference leaf lists have);
//   * the dual traversal runs breadth first over a frontier of (target,
//     source) cell pairs with scan compaction.  Each pair takes the same
//     decision as the recursion (opening criterion R*theta > rmax_A +
//     rmax_B with correctly rounded FP64 and no contraction), so the M2L and
//     P2P lists are the reference's lists as multisets.
#pragma clang fp contract(off)

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>

#include "fmm_common.h"

namespace fmm {

namespace {

constexpr int kMaxSplitLevels = 21;  // 3 key bits per level in 63 bits

__device__ __forceinline__ double rmax_of(double r) { return sqrt_rn((3.0 * r) * r); }

// Sequential mean: the block stages chunks of positions through LDS
// (coalesced loads), lanes 0..2 add them in index order.
// Mean of the positions with the reference's sequential FP64 sum
// (lanes 0..2 of wave 0 sum x, y, z in particle order, one rounding per
// addition, so the root centre is bit-exact).  Waves 1..3 stage the next
// chunk into the other half of a double buffer while wave 0 sums the
// current one, reading runs of 16 values before their 16 ordered additions.
__global__ __launch_bounds__(256) void k_mean(const double* __restrict__ pos, long long n, double* __restrict__ out) {
  constexpr int kChunk = 1024, kRun = 16;
  __shared__ double buf[2][3 * kChunk];
  double s = 0.0;
  const int tid = threadIdx.x;
  auto stage = [&](long long base, int half) {  // waves 1..3
    const int m = (int)(n - base < kChunk ? n - base : kChunk);
    for (int k = tid - 64; k < 3 * m; k += 192) buf[half][k] = pos[3 * base + k];
  };
  if (tid >= 64 && n > 0) stage(0, 0);
  __syncthreads();
  int half = 0;
  for (long long base = 0; base < n; base += kChunk, half ^= 1) {
    if (tid >= 64 && base + kChunk < n) stage(base + kChunk, half ^ 1);
    if (tid < 3) {
      const int m = (int)(n - base < kChunk ? n - base : kChunk);
      const double* b = buf[half] + tid;
      int i = 0;
      for (; i + kRun <= m; i += kRun) {
        double v[kRun];
#pragma unroll
        for (int u = 0; u < kRun; ++u) v[u] = b[3 * (i + u)];
#pragma unroll
        for (int u = 0; u < kRun; ++u) s = s + v[u];
      }
      for (; i < m; ++i) s = s + b[3 * i];
    }
    __syncthreads();  // this half consumed, the next one staged
  }
  if (tid < 3) out[tid] = s / (double)n;
}

__global__ void k_maxdev(const double* __restrict__ pos, long long n, const double* __restrict__ mean,
                         unsigned long long* __restrict__ mx) {
  const double m0 = mean[0], m1 = mean[1], m2 = mean[2];
  double a0 = 0, a1 = 0, a2 = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = fabs(pos[3 * i] - m0), y = fabs(pos[3 * i + 1] - m1), z = fabs(pos[3 * i + 2] - m2);
    a0 = x > a0 ? x : a0;
    a1 = y > a1 ? y : a1;
    a2 = z > a2 ? z : a2;
  }
  // non-negative doubles order like their bit patterns
  atomicMax(&mx[0], (unsigned long long)__double_as_longlong(a0));
  atomicMax(&mx[1], (unsigned long long)__double_as_longlong(a1));
  atomicMax(&mx[2], (unsigned long long)__double_as_longlong(a2));
}

__global__ void k_root(const double* __restrict__ mean, const unsigned long long* __restrict__ mx,
                       double* cx, double* cy, double* cz, double* cr, int* count, long long n) {
  const double xm = __longlong_as_double((long long)mx[0]);
  const double ym = __longlong_as_double((long long)mx[1]);
  const double zm = __longlong_as_double((long long)mx[2]);
  const double r = (xm > ym ? (xm > zm ? xm : zm) : (ym > zm ? ym : zm)) * 1.001;
  cx[0] = mean[0];
  cy[0] = mean[1];
  cz[0] = mean[2];
  cr[0] = r;
  count[0] = (int)n;
}

// Per particle still inside a split cell: octant histogram of the level.
// Particles whose cell at this level is a leaf are finalised.
__global__ void k_level_octants(const double* __restrict__ pos, int n, int* __restrict__ pcell,
                                unsigned char* __restrict__ poct, int* __restrict__ leaf_of,
                                const double* __restrict__ lcx, const double* __restrict__ lcy,
                                const double* __restrict__ lcz, const int* __restrict__ lcount,
                                int* __restrict__ ccount, int ncrit, int level_base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = pcell[i];
  if (c < 0) return;
  if (lcount[c] < ncrit) {
    leaf_of[i] = level_base + c;
    pcell[i] = -1;
    return;
  }
  const int oct = (pos[3 * i] > lcx[c]) + ((pos[3 * i + 1] > lcy[c]) << 1) + ((pos[3 * i + 2] > lcz[c]) << 2);
  poct[i] = (unsigned char)oct;
  atomicAdd(&ccount[8 * c + oct], 1);
}

__global__ void k_flags(const int* __restrict__ ccount, int* __restrict__ flag, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) flag[j] = ccount[j] > 0 ? 1 : 0;
}

// Children of the level, numbered by the scan over (parent, octant) slots.
__global__ void k_make_children(int nparent, const int* __restrict__ ccount, const int* __restrict__ cidx,
                                const double* __restrict__ lcx, const double* __restrict__ lcy,
                                const double* __restrict__ lcz, const double* __restrict__ lr,
                                double* __restrict__ ncx, double* __restrict__ ncy, double* __restrict__ ncz,
                                double* __restrict__ nr, int* __restrict__ ncount, int* __restrict__ nparent_local,
                                int* __restrict__ noctant, int* __restrict__ first_child_local,
                                int* __restrict__ nchild, int* __restrict__ octmask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nparent) return;
  int first = -1, cnt = 0, mask = 0;
  const double r = lr[c] / 2.0;
  for (int o = 0; o < 8; ++o) {
    const int j = 8 * c + o;
    const int k = ccount[j];
    if (k == 0) continue;
    const int id = cidx[j];
    if (first < 0) first = id;
    ++cnt;
    mask |= 1 << o;
    ncx[id] = lcx[c] + r * ((o & 1) * 2 - 1);
    ncy[id] = lcy[c] + r * ((o & 2) - 1);
    ncz[id] = lcz[c] + r * ((o & 4) / 2 - 1);
    nr[id] = r;
    ncount[id] = k;
    nparent_local[id] = c;
    noctant[id] = o;
  }
  first_child_local[c] = first;
  nchild[c] = cnt;
  octmask[c] = mask;
}

__global__ void k_descend(int n, int* __restrict__ pcell, const unsigned char* __restrict__ poct,
                          const int* __restrict__ cidx, unsigned long long* __restrict__ key, int level) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = pcell[i];
  if (c < 0) return;
  const int o = poct[i];
  pcell[i] = cidx[8 * c + o];
  key[i] |= (unsigned long long)o << (3 * (kMaxSplitLevels - 1 - level));
}

__global__ void k_iota(int* v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_gather_particles(int n, const int* __restrict__ perm, const double* __restrict__ pos,
                                   const int* __restrict__ leaf_of_orig, double* __restrict__ px,
                                   double* __restrict__ py, double* __restrict__ pz, int* __restrict__ leaf_of) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int i = perm[j];
  px[j] = pos[3 * i];
  py[j] = pos[3 * i + 1];
  pz[j] = pos[3 * i + 2];
  leaf_of[j] = leaf_of_orig[i];
}

__global__ void k_leaf_starts(int n, const int* __restrict__ leaf_of, int* __restrict__ pstart) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (j == 0 || leaf_of[j] != leaf_of[j - 1]) pstart[leaf_of[j]] = j;
}

__global__ void k_internal_starts(int beg, int end, const int* __restrict__ nchild,
                                  const int* __restrict__ first_child, int* __restrict__ pstart) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  if (nchild[c] > 0) pstart[c] = pstart[first_child[c]];
}

__global__ void k_rmax(int n, const double* __restrict__ r, double* __restrict__ rmax) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) rmax[c] = rmax_of(r[c]);
}

__global__ void k_offset_add(int* v, int n, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && v[i] >= 0) v[i] += add;
}

__global__ void k_leaf_flags(int n, const int* __restrict__ nchild, int* __restrict__ flag) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) flag[c] = nchild[c] == 0 ? 1 : 0;
}

__global__ void k_leaf_compact(int n, const int* __restrict__ flag, const int* __restrict__ pos,
                               int* __restrict__ leaves) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && flag[c]) leaves[pos[c]] = c;
}

// ---------------------------------------------------------------- lists

struct Counts3 {
  int m2l, p2p, nxt;
};
struct Plus3 {
  __host__ __device__ Counts3 operator()(const Counts3& a, const Counts3& b) const {
    return Counts3{a.m2l + b.m2l, a.p2p + b.p2p, a.nxt + b.nxt};
  }
};

// 0: M2L, 1: P2P, 2: split target A, 3: split source B
__device__ __forceinline__ int decide(int A, int B, const double* __restrict__ cx, const double* __restrict__ cy,
                                      const double* __restrict__ cz, const double* __restrict__ rmax,
                                      const int* __restrict__ nchild, const int* __restrict__ pcount,
                                      double theta, int ncrit) {
  const double dx = cx[A] - cx[B];
  const double dy = cy[A] - cy[B];
  const double dz = cz[A] - cz[B];
  const double R = sqrt_rn(dx * dx + dy * dy + dz * dz);
  if (R * theta > (rmax[A] + rmax[B])) return 0;
  const bool la = nchild[A] == 0, lb = nchild[B] == 0;
  if (la && lb) return pcount[B] >= ncrit ? 0 : 1;
  if (lb || (rmax[A] >= rmax[B] && !la)) return 2;
  return 3;
}

__global__ void k_pair_count(int nf, const int2* __restrict__ fr, Counts3* __restrict__ cnt,
                             unsigned char* __restrict__ dec, const double* __restrict__ cx,
                             const double* __restrict__ cy, const double* __restrict__ cz,
                             const double* __restrict__ rmax, const int* __restrict__ nchild,
                             const int* __restrict__ pcount, double theta, int ncrit) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int2 ab = fr[f];
  const int d = decide(ab.x, ab.y, cx, cy, cz, rmax, nchild, pcount, theta, ncrit);
  dec[f] = (unsigned char)d;
  Counts3 c{0, 0, 0};
  if (d == 0) c.m2l = 1;
  else if (d == 1) c.p2p = 1;
  else if (d == 2) c.nxt = nchild[ab.x];
  else c.nxt = nchild[ab.y];
  cnt[f] = c;
}

// M2L entries are written as (target, source) like the CSR key; the
// reference stores (source, target) (example/calculate.cpp:95).
__global__ void k_pair_emit(int nf, const int2* __restrict__ fr, const Counts3* __restrict__ off,
                            const unsigned char* __restrict__ dec, const int* __restrict__ first_child,
                            const int* __restrict__ nchild, int2* __restrict__ m2l, long long m2l_base,
                            int2* __restrict__ p2p, long long p2p_base, int2* __restrict__ next) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int2 ab = fr[f];
  const int d = dec[f];
  const Counts3 o = off[f];
  if (d == 0) {
    m2l[m2l_base + o.m2l] = make_int2(ab.x, ab.y);
  } else if (d == 1) {
    p2p[p2p_base + o.p2p] = make_int2(ab.x, ab.y);
  } else if (d == 2) {
    const int fc = first_child[ab.x], k = nchild[ab.x];
    for (int i = 0; i < k; ++i) next[o.nxt + i] = make_int2(fc + i, ab.y);
  } else {
    const int fc = first_child[ab.y], k = nchild[ab.y];
    for (int i = 0; i < k; ++i) next[o.nxt + i] = make_int2(ab.x, fc + i);
  }
}

__global__ void k_pair_keys(long long m, const int2* __restrict__ pr, unsigned long long* __restrict__ key) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) key[i] = ((unsigned long long)(unsigned)pr[i].x << 32) | (unsigned)pr[i].y;
}

__global__ void k_csr_fill(long long m, const unsigned long long* __restrict__ key, int* __restrict__ src,
                           int* __restrict__ row_first, int* __restrict__ row_end) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const unsigned long long k = key[i];
  const int t = (int)(k >> 32);
  src[i] = (int)(k & 0xffffffffu);
  if (i == 0 || (int)(key[i - 1] >> 32) != t) row_first[t] = (int)i;
  if (i == m - 1 || (int)(key[i + 1] >> 32) != t) row_end[t] = (int)(i + 1);
}

// Entries per row (0 for rows without any: both bounds stay zero).
__global__ void k_csr_count(int ncells, const int* __restrict__ row_first, const int* __restrict__ row_end,
                            int* __restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) cnt[c] = row_end[c] - row_first[c];
}

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

inline int blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// P2P particle pairs: sum over list entries of nA * nB, minus self pairs.
__global__ void k_p2p_work(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                           const int* __restrict__ pcount, unsigned long long* __restrict__ total) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long w = 0;
  if (A < ncells) {
    const long long na = pcount[A];
    for (int i = off[A]; i < off[A + 1]; ++i) {
      const int B = src[i];
      w += (unsigned long long)(na * pcount[B] - (B == A ? na : 0));
    }
  }
  if (w) atomicAdd(total, w);
}

template <typename T>
void exclusive_scan_int(Scratch& s, const T* in, T* out, long long n, hipStream_t st) {
  size_t bytes = 0;
  FMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  void* tmp = s.get(bytes);
  FMM_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
}

// CSR offsets off[0..nc] from each row's first and one-past-last entry
// (both zero for a row without entries): the exclusive scan of the row
// lengths, so an empty row's offset is the next row's start; off[nc] = m.
void close_csr(Tree& t, const int* first, const int* end, int* off, int nc, long long m, hipStream_t st) {
  DevBuf<int> cnt;
  cnt.alloc(nc);
  k_csr_count<<<blocks(nc, 256), 256, 0, st>>>(nc, first, end, cnt.get());
  exclusive_scan_int(t.scratch, cnt.get(), off, nc, st);
  k_set_int<<<1, 64, 0, st>>>(off + nc, (int)m);
  FMM_HIP(hipStreamSynchronize(st));  // cnt is freed on return
}

void build_csr(Tree& t, DevBuf<int2>& pairs, long long m, DevBuf<int>& off, DevBuf<int>& src,
               hipStream_t st) {
  const int nc = t.ncells;
  off.alloc(nc + 1);
  src.alloc(std::max<long long>(m, 1));
  DevBuf<int> first, end;
  first.alloc(nc);
  end.alloc(nc);
  FMM_HIP(hipMemsetAsync(first.get(), 0, sizeof(int) * nc, st));
  FMM_HIP(hipMemsetAsync(end.get(), 0, sizeof(int) * nc, st));
  if (m > 0) {
    DevBuf<unsigned long long> k0, k1;
    k0.alloc(m);
    k1.alloc(m);
    k_pair_keys<<<blocks(m, 256), 256, 0, st>>>(m, pairs.get(), k0.get());
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_keys(nullptr, bytes, k0.get(), k1.get(), (size_t)m, 0, 64, st));
    void* tmp = t.scratch.get(bytes);
    FMM_HIP(rocprim::radix_sort_keys(tmp, bytes, k0.get(), k1.get(), (size_t)m, 0, 64, st));
    k_csr_fill<<<blocks(m, 256), 256, 0, st>>>(m, k1.get(), src.get(), first.get(), end.get());
    FMM_HIP(hipStreamSynchronize(st));
  }
  close_csr(t, first.get(), end.get(), off.get(), nc, m, st);
}

// Source ranges: one entry per (target, non-self source) pair keyed by
// (target, source particle start).
__global__ void k_rng_keys(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                           const int* __restrict__ pstart, unsigned long long* __restrict__ key,
                           int* __restrict__ cnt_out, int* __restrict__ pos) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells) return;
  int w = pos[A];
  for (int i = off[A]; i < off[A + 1]; ++i) {
    const int B = src[i];
    if (B == A) continue;
    key[w] = ((unsigned long long)(unsigned)A << 32) | (unsigned)pstart[B];
    cnt_out[w] = B;
    ++w;
  }
}

__global__ void k_rng_count(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                            int* __restrict__ n) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells) return;
  int c = 0;
  for (int i = off[A]; i < off[A + 1]; ++i) c += src[i] != A;
  n[A] = c;
}

// run starts: new target, or a gap to the previous source's end
__global__ void k_rng_flags(long long m, const unsigned long long* __restrict__ key, const int* __restrict__ srcB,
                            const int* __restrict__ pcount, int* __restrict__ flag) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  int f = 1;
  if (i > 0 && (key[i] >> 32) == (key[i - 1] >> 32)) {
    const unsigned prev_end = (unsigned)(key[i - 1] & 0xffffffffu) + (unsigned)pcount[srcB[i - 1]];
    f = ((unsigned)(key[i] & 0xffffffffu) != prev_end);
  }
  flag[i] = f;
}

__global__ void k_rng_fill(long long m, const unsigned long long* __restrict__ key, const int* __restrict__ srcB,
                           const int* __restrict__ pcount, const int* __restrict__ flag, const int* __restrict__ pos,
                           int* __restrict__ rstart, int* __restrict__ rlen, int* __restrict__ row_first,
                           int* __restrict__ row_end) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int r = pos[i] + flag[i] - 1;
  const int A = (int)(key[i] >> 32);
  if (flag[i]) {
    rstart[r] = (int)(key[i] & 0xffffffffu);
    if (i == 0 || (int)(key[i - 1] >> 32) != A) row_first[A] = r;
  }
  if (i == m - 1 || (int)(key[i + 1] >> 32) != A) row_end[A] = r + 1;
  atomicAdd(&rlen[r], pcount[srcB[i]]);
}

// Per leaf: number of P2P segments; per range: its length as 64-bit.
__global__ void k_seg_count(int ncells, const int* __restrict__ nchild, const int* __restrict__ pcount,
                            const int* __restrict__ rng_off, const long long* __restrict__ cum,
                            long long* __restrict__ nseg, int seg) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells) return;
  long long k = 0;
  if (nchild[A] == 0 && pcount[A] > 0) {
    const long long tot = pcount[A] + cum[rng_off[A + 1]] - cum[rng_off[A]];
    k = (tot + seg - 1) / seg;
  }
  nseg[A] = k;
}

__global__ void k_widen(long long m, const int* __restrict__ in, long long* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) out[i] = in[i];
}

// Items of one leaf: segment s starts at position s * seg of the source
// sequence (own particles first, then the ranges in order).
__global__ void k_seg_fill(int ncells, const int* __restrict__ nchild, const int* __restrict__ pcount,
                           const int* __restrict__ rng_off, const long long* __restrict__ cum,
                           const long long* __restrict__ nseg, const long long* __restrict__ seg_off, int seg,
                           int* __restrict__ it_leaf, int* __restrict__ it_r, int* __restrict__ it_o,
                           int* __restrict__ it_n) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells || nseg[A] == 0) return;
  const int nt = pcount[A], rb = rng_off[A], re = rng_off[A + 1];
  const long long base = cum[rb], tot = nt + cum[re] - base;
  for (long long s = 0; s < nseg[A]; ++s) {
    const long long P = s * seg, w = seg_off[A] + s;
    int r, o;
    if (P < nt) {
      r = rb - 1;
      o = (int)P;
    } else {
      // last range r in [rb, re) with cum[r] - base <= P - nt
      const long long q = P - nt + base;
      int lo = rb, hi = re - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      r = lo;
      o = (int)(q - cum[r]);
    }
    it_leaf[w] = A;
    it_r[w] = r;
    it_o[w] = o;
    it_n[w] = (int)((tot - P < seg) ? tot - P : seg);
  }
}

}  // namespace

// P2P work items (segments of each leaf's source sequence).
void build_p2p_items(Tree& t, hipStream_t st) {
  const int nc = t.ncells;
  const long long nr = t.n_rng;
  DevBuf<long long> len, cum, nseg, soff;
  len.alloc(nr + 1);
  cum.alloc(nr + 1);
  FMM_HIP(hipMemsetAsync(len.get(), 0, sizeof(long long) * (nr + 1), st));
  if (nr > 0) k_widen<<<blocks(nr, 256), 256, 0, st>>>(nr, t.rng_len.get(), len.get());
  exclusive_scan_int(t.scratch, len.get(), cum.get(), nr + 1, st);
  nseg.alloc(nc + 1);
  soff.alloc(nc + 1);
  FMM_HIP(hipMemsetAsync(nseg.get(), 0, sizeof(long long) * (nc + 1), st));
  const int seg = Tree::kP2PSegment;
  k_seg_count<<<blocks(nc, 256), 256, 0, st>>>(nc, t.nchild.get(), t.pcount.get(), t.rng_off.get(), cum.get(),
                                               nseg.get(), seg);
  exclusive_scan_int(t.scratch, nseg.get(), soff.get(), nc + 1, st);
  long long total = 0;
  FMM_HIP(hipMemcpyAsync(&total, soff.get() + nc, sizeof(long long), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  FMM_CHECK(total < (1LL << 31), kOutOfMemory, "P2P work items exceed 2^31");
  t.n_p2p_items = total;
  const long long na = std::max<long long>(total, 1);
  t.p2p_item_leaf.alloc(na);
  t.p2p_item_r.alloc(na);
  t.p2p_item_o.alloc(na);
  t.p2p_item_n.alloc(na);
  k_seg_fill<<<blocks(nc, 256), 256, 0, st>>>(nc, t.nchild.get(), t.pcount.get(), t.rng_off.get(), cum.get(),
                                              nseg.get(), soff.get(), seg, t.p2p_item_leaf.get(), t.p2p_item_r.get(),
                                              t.p2p_item_o.get(), t.p2p_item_n.get());
  FMM_HIP(hipStreamSynchronize(st));
  t.p2p_split = total > t.nleaves;
}

// Merged source particle ranges of every P2P target (self excluded).
void build_p2p_ranges(Tree& t, hipStream_t st) {
  const int nc = t.ncells;
  DevBuf<int> cnt, pos;
  cnt.alloc(nc + 1);
  pos.alloc(nc + 1);
  k_rng_count<<<blocks(nc, 256), 256, 0, st>>>(nc, t.p2p_off.get(), t.p2p_src.get(), cnt.get());
  exclusive_scan_int(t.scratch, cnt.get(), pos.get(), nc, st);
  int a = 0, b = 0;
  FMM_HIP(hipMemcpyAsync(&a, cnt.get() + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipMemcpyAsync(&b, pos.get() + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  const long long m = (long long)a + b;
  t.rng_off.alloc(nc + 1);
  if (m == 0) {
    t.n_rng = 0;
    t.rng_start.alloc(1);
    t.rng_len.alloc(1);
    FMM_HIP(hipMemsetAsync(t.rng_off.get(), 0, sizeof(int) * (nc + 1), st));
    FMM_HIP(hipStreamSynchronize(st));
    return;
  }
  DevBuf<unsigned long long> k0, k1;
  DevBuf<int> b0, b1, fl, ps;
  k0.alloc(m); k1.alloc(m); b0.alloc(m); b1.alloc(m); fl.alloc(m); ps.alloc(m);
  k_rng_keys<<<blocks(nc, 256), 256, 0, st>>>(nc, t.p2p_off.get(), t.p2p_src.get(), t.pstart.get(), k0.get(),
                                              b0.get(), pos.get());
  size_t bytes = 0;
  FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k0.get(), k1.get(), b0.get(), b1.get(), (size_t)m, 0, 64, st));
  void* tmp = t.scratch.get(bytes);
  FMM_HIP(rocprim::radix_sort_pairs(tmp, bytes, k0.get(), k1.get(), b0.get(), b1.get(), (size_t)m, 0, 64, st));
  k_rng_flags<<<blocks(m, 256), 256, 0, st>>>(m, k1.get(), b1.get(), t.pcount.get(), fl.get());
  exclusive_scan_int(t.scratch, fl.get(), ps.get(), m, st);
  FMM_HIP(hipMemcpyAsync(&a, fl.get() + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipMemcpyAsync(&b, ps.get() + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  t.n_rng = (long long)a + b;
  t.rng_start.alloc(t.n_rng);
  t.rng_len.alloc(t.n_rng);
  FMM_HIP(hipMemsetAsync(t.rng_len.get(), 0, sizeof(int) * t.n_rng, st));
  DevBuf<int> first, end;
  first.alloc(nc);
  end.alloc(nc);
  FMM_HIP(hipMemsetAsync(first.get(), 0, sizeof(int) * nc, st));
  FMM_HIP(hipMemsetAsync(end.get(), 0, sizeof(int) * nc, st));
  k_rng_fill<<<blocks(m, 256), 256, 0, st>>>(m, k1.get(), b1.get(), t.pcount.get(), fl.get(), ps.get(),
                                             t.rng_start.get(), t.rng_len.get(), first.get(), end.get());
  close_csr(t, first.get(), end.get(), t.rng_off.get(), nc, t.n_rng, st);
}

// Builds the tree of `t` from caller positions (device, [n*3]).
void build_tree_device(Tree& t, const double* d_pos, hipStream_t st) {
  const int n = (int)t.n;
  const int ncrit = (int)t.ncrit;
  DevBuf<double> mean;
  DevBuf<unsigned long long> mx;
  mean.alloc(3);
  mx.alloc(3);
  FMM_HIP(hipMemsetAsync(mx.get(), 0, 3 * sizeof(unsigned long long), st));
  k_mean<<<1, 256, 0, st>>>(d_pos, n, mean.get());
  k_maxdev<<<std::min(blocks(n, 256), 2048), 256, 0, st>>>(d_pos, n, mean.get(), mx.get());

  // per-level arrays
  struct Level {
    int n = 0;
    DevBuf<double> cx, cy, cz, r;
    DevBuf<int> count, parent, first_child, nchild, octmask, oct;
  };
  std::vector<Level*> lv;
  auto new_level = [&](int m) {
    Level* L = new Level();
    L->n = m;
    L->cx.alloc(m); L->cy.alloc(m); L->cz.alloc(m); L->r.alloc(m);
    L->count.alloc(m); L->parent.alloc(m); L->first_child.alloc(m); L->nchild.alloc(m); L->octmask.alloc(m);
    L->oct.alloc(m);
    FMM_HIP(hipMemsetAsync(L->oct.get(), 0, sizeof(int) * m, st));
    FMM_HIP(hipMemsetAsync(L->nchild.get(), 0, sizeof(int) * m, st));
    FMM_HIP(hipMemsetAsync(L->octmask.get(), 0, sizeof(int) * m, st));
    FMM_HIP(hipMemsetAsync(L->first_child.get(), 0xff, sizeof(int) * m, st));
    lv.push_back(L);
    return L;
  };
  struct Cleanup {
    std::vector<Level*>& v;
    ~Cleanup() { for (auto* p : v) delete p; }
  } cleanup{lv};

  Level* root = new_level(1);
  FMM_HIP(hipMemsetAsync(root->parent.get(), 0, sizeof(int), st));
  k_root<<<1, 1, 0, st>>>(mean.get(), mx.get(), root->cx.get(), root->cy.get(), root->cz.get(), root->r.get(),
                          root->count.get(), n);

  DevBuf<int> pcell, leaf_orig, ccount, flag, cidx;
  DevBuf<unsigned char> poct;
  DevBuf<unsigned long long> key;
  pcell.alloc(n);
  leaf_orig.alloc(n);
  poct.alloc(n);
  key.alloc(n);
  FMM_HIP(hipMemsetAsync(pcell.get(), 0, sizeof(int) * n, st));
  FMM_HIP(hipMemsetAsync(key.get(), 0, sizeof(unsigned long long) * n, st));

  std::vector<int> level_off{0};
  int level = 0;
  while (true) {
    Level* cur = lv.back();
    const int m = cur->n;
    level_off.push_back(level_off.back() + m);
    ccount.alloc(8 * (size_t)m);
    FMM_HIP(hipMemsetAsync(ccount.get(), 0, sizeof(int) * 8 * (size_t)m, st));
    k_level_octants<<<blocks(n, 256), 256, 0, st>>>(d_pos, n, pcell.get(), poct.get(), leaf_orig.get(),
                                                    cur->cx.get(), cur->cy.get(), cur->cz.get(),
                                                    cur->count.get(), ccount.get(), ncrit, level_off[level]);
    flag.alloc(8 * (size_t)m);
    cidx.alloc(8 * (size_t)m + 1);
    k_flags<<<blocks(8LL * m, 256), 256, 0, st>>>(ccount.get(), flag.get(), 8 * m);
    exclusive_scan_int(t.scratch, flag.get(), cidx.get(), 8LL * m, st);
    int last_flag = 0, last_idx = 0;
    FMM_HIP(hipMemcpyAsync(&last_flag, flag.get() + 8 * (size_t)m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(&last_idx, cidx.get() + 8 * (size_t)m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    const int nnext = last_idx + last_flag;
    if (nnext == 0) break;
    FMM_CHECK(level < kMaxSplitLevels, kTreeTooDeep,
              "tree deeper than 21 split levels (coincident particles?)");
    Level* nx = new_level(nnext);
    k_make_children<<<blocks(m, 256), 256, 0, st>>>(
        m, ccount.get(), cidx.get(), cur->cx.get(), cur->cy.get(), cur->cz.get(), cur->r.get(), nx->cx.get(),
        nx->cy.get(), nx->cz.get(), nx->r.get(), nx->count.get(), nx->parent.get(), nx->oct.get(),
        cur->first_child.get(), cur->nchild.get(), cur->octmask.get());
    k_descend<<<blocks(n, 256), 256, 0, st>>>(n, pcell.get(), poct.get(), cidx.get(), key.get(), level);
    ++level;
  }
  t.depth = (int)lv.size();
  t.level_off = level_off;
  t.ncells = level_off.back();
  const int nc = t.ncells;

  // concatenate levels into canonical global arrays
  t.cx.alloc(nc); t.cy.alloc(nc); t.cz.alloc(nc); t.cr.alloc(nc); t.crmax.alloc(nc);
  t.parent.alloc(nc); t.first_child.alloc(nc); t.nchild.alloc(nc); t.octmask.alloc(nc);
  t.level.alloc(nc); t.pstart.alloc(nc); t.pcount.alloc(nc); t.octant.alloc(nc);
  std::vector<int> hlevel(nc);
  for (int l = 0; l < t.depth; ++l) {
    Level* L = lv[l];
    const int o = level_off[l], m = L->n;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      FMM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    };
    cp(t.cx.get() + o, L->cx.get(), sizeof(double) * m);
    cp(t.cy.get() + o, L->cy.get(), sizeof(double) * m);
    cp(t.cz.get() + o, L->cz.get(), sizeof(double) * m);
    cp(t.cr.get() + o, L->r.get(), sizeof(double) * m);
    cp(t.pcount.get() + o, L->count.get(), sizeof(int) * m);
    cp(t.parent.get() + o, L->parent.get(), sizeof(int) * m);
    cp(t.first_child.get() + o, L->first_child.get(), sizeof(int) * m);
    cp(t.nchild.get() + o, L->nchild.get(), sizeof(int) * m);
    cp(t.octmask.get() + o, L->octmask.get(), sizeof(int) * m);
    cp(t.octant.get() + o, L->oct.get(), sizeof(int) * m);
    if (l > 0) k_offset_add<<<blocks(m, 256), 256, 0, st>>>(t.parent.get() + o, m, level_off[l - 1]);
    if (l + 1 < t.depth) k_offset_add<<<blocks(m, 256), 256, 0, st>>>(t.first_child.get() + o, m, level_off[l + 1]);
    for (int c = 0; c < m; ++c) hlevel[o + c] = l;
  }
  FMM_HIP(hipMemcpyAsync(t.level.get(), hlevel.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
  k_rmax<<<blocks(nc, 256), 256, 0, st>>>(nc, t.cr.get(), t.crmax.get());

  // stable sort of particles by path key
  {
    DevBuf<unsigned long long> k1;
    DevBuf<int> v0;
    k1.alloc(n);
    v0.alloc(n);
    t.perm.alloc(n);
    k_iota<<<blocks(n, 256), 256, 0, st>>>(v0.get(), n);
    const unsigned begin_bit = 3u * (unsigned)(kMaxSplitLevels - std::max(level, 1));
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.get(), k1.get(), v0.get(), t.perm.get(), (size_t)n,
                                      begin_bit, 63u, st));
    void* tmp = t.scratch.get(bytes);
    FMM_HIP(rocprim::radix_sort_pairs(tmp, bytes, key.get(), k1.get(), v0.get(), t.perm.get(), (size_t)n,
                                      begin_bit, 63u, st));
    FMM_HIP(hipStreamSynchronize(st));
  }
  t.px.alloc(n); t.py.alloc(n); t.pz.alloc(n); t.leaf_of.alloc(n);
  k_gather_particles<<<blocks(n, 256), 256, 0, st>>>(n, t.perm.get(), d_pos, leaf_orig.get(), t.px.get(),
                                                     t.py.get(), t.pz.get(), t.leaf_of.get());
  k_leaf_starts<<<blocks(n, 256), 256, 0, st>>>(n, t.leaf_of.get(), t.pstart.get());
  for (int l = t.depth - 2; l >= 0; --l) {
    const int b = level_off[l], e = level_off[l + 1];
    k_internal_starts<<<blocks(e - b, 256), 256, 0, st>>>(b, e, t.nchild.get(), t.first_child.get(),
                                                          t.pstart.get());
  }
  // leaves in canonical order
  {
    DevBuf<int> lf, lp;
    lf.alloc(nc + 1);
    lp.alloc(nc + 1);
    k_leaf_flags<<<blocks(nc, 256), 256, 0, st>>>(nc, t.nchild.get(), lf.get());
    exclusive_scan_int(t.scratch, lf.get(), lp.get(), nc, st);
    int a = 0, b = 0;
    FMM_HIP(hipMemcpyAsync(&a, lf.get() + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(&b, lp.get() + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    t.nleaves = a + b;
    t.leaves.alloc(t.nleaves);
    k_leaf_compact<<<blocks(nc, 256), 256, 0, st>>>(nc, lf.get(), lp.get(), t.leaves.get());
    FMM_HIP(hipStreamSynchronize(st));
  }
  FMM_HIP(hipStreamSynchronize(st));
}

// Breadth-first dual traversal from (root, root) and CSR lists by target.
void build_lists_device(Tree& t, hipStream_t st) {
  DevBuf<int2> cur, nxt, m2l, p2p;
  DevBuf<Counts3> cnt, off;
  DevBuf<unsigned char> dec;
  long long n_m2l = 0, n_p2p = 0;
  long long cap_m2l = 1 << 20, cap_p2p = 1 << 20;
  m2l.alloc(cap_m2l);
  p2p.alloc(cap_p2p);
  cur.alloc(1);
  int2 rootpair = make_int2(0, 0);
  FMM_HIP(hipMemcpyAsync(cur.get(), &rootpair, sizeof(int2), hipMemcpyHostToDevice, st));
  int nf = 1;
  auto grow = [&](DevBuf<int2>& b, long long& cap, long long used, long long need) {
    if (need <= cap) return;
    long long nc = std::max(need, 2 * cap);
    DevBuf<int2> nb;
    nb.alloc(nc);
    if (used > 0) FMM_HIP(hipMemcpyAsync(nb.get(), b.get(), sizeof(int2) * used, hipMemcpyDeviceToDevice, st));
    FMM_HIP(hipStreamSynchronize(st));
    std::swap(b.p, nb.p);
    std::swap(b.n, nb.n);
    cap = nc;
  };
  while (nf > 0) {
    cnt.alloc(nf + 1);
    off.alloc(nf + 1);
    dec.alloc(nf);
    k_pair_count<<<blocks(nf, 256), 256, 0, st>>>(nf, cur.get(), cnt.get(), dec.get(), t.cx.get(), t.cy.get(),
                                                  t.cz.get(), t.crmax.get(), t.nchild.get(), t.pcount.get(),
                                                  t.theta, (int)t.ncrit);
    {
      size_t bytes = 0;
      FMM_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt.get(), off.get(), Counts3{0, 0, 0}, (size_t)nf, Plus3(), st));
      void* tmp = t.scratch.get(bytes);
      FMM_HIP(rocprim::exclusive_scan(tmp, bytes, cnt.get(), off.get(), Counts3{0, 0, 0}, (size_t)nf, Plus3(), st));
    }
    Counts3 lastc, lasto;
    FMM_HIP(hipMemcpyAsync(&lastc, cnt.get() + nf - 1, sizeof(Counts3), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(&lasto, off.get() + nf - 1, sizeof(Counts3), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    const long long add_m2l = (long long)lastc.m2l + lasto.m2l;
    const long long add_p2p = (long long)lastc.p2p + lasto.p2p;
    const long long add_nxt = (long long)lastc.nxt + lasto.nxt;
    grow(m2l, cap_m2l, n_m2l, n_m2l + add_m2l);
    grow(p2p, cap_p2p, n_p2p, n_p2p + add_p2p);
    nxt.alloc(std::max<long long>(add_nxt, 1));
    k_pair_emit<<<blocks(nf, 256), 256, 0, st>>>(nf, cur.get(), off.get(), dec.get(), t.first_child.get(),
                                                 t.nchild.get(), m2l.get(), n_m2l, p2p.get(), n_p2p, nxt.get());
    FMM_HIP(hipStreamSynchronize(st));
    n_m2l += add_m2l;
    n_p2p += add_p2p;
    FMM_CHECK(add_nxt < (1LL << 31), kOutOfMemory, "traversal frontier exceeds 2^31 pairs");
    nf = (int)add_nxt;
    std::swap(cur.p, nxt.p);
    std::swap(cur.n, nxt.n);
  }
  FMM_HIP(hipStreamSynchronize(st));
  t.n_m2l = n_m2l;
  t.n_p2p = n_p2p;
  build_csr(t, m2l, n_m2l, t.m2l_off, t.m2l_src, st);
  build_csr(t, p2p, n_p2p, t.p2p_off, t.p2p_src, st);
  build_p2p_ranges(t, st);
  build_p2p_items(t, st);
  {
    DevBuf<unsigned long long> tot;
    tot.alloc(1);
    FMM_HIP(hipMemsetAsync(tot.get(), 0, sizeof(unsigned long long), st));
    k_p2p_work<<<blocks(t.ncells, 256), 256, 0, st>>>(t.ncells, t.p2p_off.get(), t.p2p_src.get(), t.pcount.get(),
                                                      tot.get());
    unsigned long long h = 0;
    FMM_HIP(hipMemcpyAsync(&h, tot.get(), sizeof(h), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    t.p2p_particle_pairs = (long long)h;
  }
  FMM_HIP(hipStreamSynchronize(st));
}

}  // namespace fmm
