// M2L schedule: target tiles and (tile, offset) groups.
//
// The M2L list (example/calculate.cpp:190-206, one M2L(A - B, M_B, L_A) per
// pair) is regrouped so that the translation runs as small GEMMs:
//   * target cells that have M2L sources are cut into tiles of kTile cells
//     of one level, ordered by (octant, parent octant) class and then
//     canonically, so the cells of a tile share most relative offsets;
//   * a pair's key is (target level, source level, centre offset in units
//     of the finer half-width).  Centres sit on a dyadic lattice, so all
//     pairs of one key share one translation matrix T[n][m] = D_{n+m}
//     (up to the last-bit rounding of the FP64 centres);
//   * a group is (tile, key): at most one source per tile column (the
//     source of a target at a given key is unique), so the group is one
//     16-column operand of the tile's GEMM; the tile's group list is cut
//     into chunks (one wavefront each) whose results add atomically.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>

#include "evaluate.h"
#include "fmm_common.h"

namespace fmm {

namespace {

inline int blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

__global__ void k_has_m2l(int b, int e, const int* __restrict__ off, int* __restrict__ flag) {
  const int c = b + blockIdx.x * blockDim.x + threadIdx.x;
  if (c < e) flag[c - b] = off[c + 1] > off[c] ? 1 : 0;
}

// Tile order of a level's targets: (octant, parent octant) class first,
// canonical order inside a class.  Away from the domain boundary, cells of
// one class see the same set of relative M2L offsets, so a tile of them
// fills its 16 columns for most offsets.  Cells without M2L work sort last.
// With a per-cell order hint (the learned order the cell is predicted to
// run at), cells of one class are further ordered by hint, so that a tile
// (which runs at the largest order of its cells) holds cells of one order.
__global__ void k_tile_keys(int b, int e, const int* __restrict__ off, const int* __restrict__ octant,
                            const int* __restrict__ parent, const int* __restrict__ hint,
                            unsigned long long* __restrict__ key, int* __restrict__ cell) {
  const int c = b + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= e) return;
  const bool has = off[c + 1] > off[c];
  const unsigned cls = (unsigned)(octant[c] * 8 + octant[parent[c]]);
  const unsigned h = hint ? (unsigned)min(max(hint[c], 0), 15) : 0u;
  key[c - b] = has ? (((unsigned long long)cls << 36) | ((unsigned long long)h << 32) | (unsigned)c) : ~0ULL;
  cell[c - b] = c;
}

__global__ void k_assign_tiles(int cnt, const int* __restrict__ cell_sorted, int tile_base,
                               int* __restrict__ tile_of, int* __restrict__ col_of, int* __restrict__ tile_cells) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int c = cell_sorted[i];
  const int t = tile_base + i / kTile, col = i % kTile;
  tile_of[c] = t;
  col_of[c] = col;
  tile_cells[t * kTile + col] = c;
}

// Per pair: key (target level, source level, quantised offset) and the
// pair's tile.  Offsets are integer multiples of the finer half-width.
__global__ void k_pair_key(long long m, const int* __restrict__ off, const int* __restrict__ src, int ncells,
                           const int* __restrict__ level, const double* __restrict__ cx,
                           const double* __restrict__ cy, const double* __restrict__ cz,
                           const double* __restrict__ cr, unsigned long long* __restrict__ key,
                           int* __restrict__ tgt, int* __restrict__ bad) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  // target of entry i by binary search in the CSR offsets
  int lo = 0, hi = ncells;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  const int A = lo, B = src[i];
  tgt[i] = A;
  const double rf = fmin(cr[A], cr[B]);
  const long long qx = llrint((cx[A] - cx[B]) / rf);
  const long long qy = llrint((cy[A] - cy[B]) / rf);
  const long long qz = llrint((cz[A] - cz[B]) / rf);
  const long long lim = 1LL << 17;
  if (qx <= -lim || qx >= lim || qy <= -lim || qy >= lim || qz <= -lim || qz >= lim) atomicOr(bad, 1);
  key[i] = ((unsigned long long)(level[A] & 31) << 59) | ((unsigned long long)(level[B] & 31) << 54) |
           ((unsigned long long)((qx + lim) & 0x3ffff) << 36) | ((unsigned long long)((qy + lim) & 0x3ffff) << 18) |
           (unsigned long long)((qz + lim) & 0x3ffff);
}

__global__ void k_iota_ll(long long m, int* __restrict__ v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) v[i] = (int)i;
}

__global__ void k_run_flags64(long long m, const unsigned long long* __restrict__ k, int* __restrict__ flag) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) flag[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// Key ids in pair order and one representative pair per key.
__global__ void k_key_ids(long long m, const int* __restrict__ flag, const int* __restrict__ pos,
                          const int* __restrict__ order, const int* __restrict__ tgt, const int* __restrict__ src,
                          const int* __restrict__ level, const double* __restrict__ cx,
                          const double* __restrict__ cy, const double* __restrict__ cz,
                          const double* __restrict__ cr, int* __restrict__ keyid, double* __restrict__ key_dx,
                          double* __restrict__ key_r, int* __restrict__ key_dlevel) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int id = pos[j] + flag[j] - 1;  // inclusive rank
  const int i = order[j];
  keyid[i] = id;
  if (flag[j]) {
    const int A = tgt[i], B = src[i];
    key_dx[3 * id + 0] = cx[A] - cx[B];
    key_dx[3 * id + 1] = cy[A] - cy[B];
    key_dx[3 * id + 2] = cz[A] - cz[B];
    key_r[id] = cr[A];
    key_dlevel[id] = level[A] - level[B];
  }
}

__global__ void k_group_keys(long long m, const int* __restrict__ tgt, const int* __restrict__ keyid,
                             const int* __restrict__ tile_of, unsigned long long* __restrict__ comp) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < m) comp[i] = ((unsigned long long)(unsigned)tile_of[tgt[i]] << 32) | (unsigned)keyid[i];
}

__global__ void k_fill_groups(long long m, const unsigned long long* __restrict__ comp,
                              const int* __restrict__ order, const int* __restrict__ flag,
                              const int* __restrict__ pos, const int* __restrict__ tgt,
                              const int* __restrict__ src, const int* __restrict__ col_of,
                              int* __restrict__ group_key, int* __restrict__ group_src,
                              int* __restrict__ tile_first) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int g = pos[j] + flag[j] - 1;
  const int i = order[j];
  const unsigned long long c = comp[j];
  if (flag[j]) {
    group_key[g] = (int)(c & 0xffffffffu);
    const int t = (int)(c >> 32);
    if (j == 0 || (int)(comp[j - 1] >> 32) != t) tile_first[t] = g;
  }
  group_src[(long long)g * kTile + col_of[tgt[i]]] = src[i];
}

__global__ void k_close_offsets(int n, int* __restrict__ off, int total) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int next = total;
  off[n] = total;
  for (int c = n - 1; c >= 0; --c) {
    if (off[c] < 0) off[c] = next;
    else next = off[c];
  }
}

template <typename T>
void scan_ex(Scratch& s, const T* in, T* out, long long n, hipStream_t st) {
  size_t bytes = 0;
  FMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  void* tmp = s.get(bytes);
  FMM_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
}

int read_int(const int* d, hipStream_t st) {
  int h = 0;
  FMM_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  return h;
}

}  // namespace

void build_m2l_schedule(Tree& t, hipStream_t st) {
  M2LSchedule& S = t.sched;
  S.tile_pairs_ready = false;  // statistics and maps of the previous tiles
  S.cell_tile_ready = false;
  const int nc = t.ncells;
  const long long m = t.n_m2l;
  // tiles, one level at a time
  DevBuf<int> tile_of, col_of, flag, rank;
  tile_of.alloc(nc);
  col_of.alloc(nc);
  std::vector<int> tiles_per_level(t.depth), tile_base(t.depth + 1, 0);
  {
    DevBuf<int> fl, rk;
    for (int l = 0; l < t.depth; ++l) {
      const int b = t.level_off[l], e = t.level_off[l + 1], k = e - b;
      fl.alloc(k + 1);
      rk.alloc(k + 1);
      k_has_m2l<<<blocks(k, 256), 256, 0, st>>>(b, e, t.m2l_off.get(), fl.get());
      scan_ex(t.scratch, fl.get(), rk.get(), k, st);
      const int cnt = read_int(rk.get() + k - 1, st) + read_int(fl.get() + k - 1, st);
      tiles_per_level[l] = (cnt + kTile - 1) / kTile;
      tile_base[l + 1] = tile_base[l] + tiles_per_level[l];
    }
    S.ntiles = tile_base[t.depth];
    S.tile_cells.alloc((size_t)std::max(S.ntiles, 1) * kTile);
    FMM_HIP(hipMemsetAsync(S.tile_cells.get(), 0xff, sizeof(int) * std::max(S.ntiles, 1) * kTile, st));
    DevBuf<unsigned long long> k0, k1;
    DevBuf<int> c0, c1;
    for (int l = 0; l < t.depth; ++l) {
      const int b = t.level_off[l], e = t.level_off[l + 1], k = e - b;
      const int cnt = tiles_per_level[l] == 0 ? 0 : -1;
      if (cnt == 0) continue;
      k0.alloc(k); k1.alloc(k); c0.alloc(k); c1.alloc(k);
      fl.alloc(k + 1);
      rk.alloc(k + 1);
      k_has_m2l<<<blocks(k, 256), 256, 0, st>>>(b, e, t.m2l_off.get(), fl.get());
      scan_ex(t.scratch, fl.get(), rk.get(), k, st);
      const int ncnt = read_int(rk.get() + k - 1, st) + read_int(fl.get() + k - 1, st);
      k_tile_keys<<<blocks(k, 256), 256, 0, st>>>(b, e, t.m2l_off.get(), t.octant.get(), t.parent.get(),
                                                  t.tile_hint_ptr(), k0.get(), c0.get());
      size_t bytes = 0;
      FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k0.get(), k1.get(), c0.get(), c1.get(), (size_t)k, 0, 64, st));
      void* tmp = t.scratch.get(bytes);
      FMM_HIP(rocprim::radix_sort_pairs(tmp, bytes, k0.get(), k1.get(), c0.get(), c1.get(), (size_t)k, 0, 64, st));
      k_assign_tiles<<<blocks(ncnt, 256), 256, 0, st>>>(ncnt, c1.get(), tile_base[l], tile_of.get(), col_of.get(),
                                                        S.tile_cells.get());
      FMM_HIP(hipStreamSynchronize(st));
    }
  }
  S.tile_group_off.alloc(S.ntiles + 1);
  if (m == 0) {
    S.ngroups = 0;
    S.nkeys = 0;
    FMM_HIP(hipMemsetAsync(S.tile_group_off.get(), 0, sizeof(int) * (S.ntiles + 1), st));
    FMM_HIP(hipStreamSynchronize(st));
    return;
  }
  FMM_CHECK(m < (1LL << 31), kOutOfMemory, "more than 2^31 M2L pairs");
  DevBuf<unsigned long long> key, key_sorted;
  DevBuf<int> tgt, order0, order, fl, pos, keyid, bad;
  key.alloc(m); key_sorted.alloc(m); tgt.alloc(m); order0.alloc(m); order.alloc(m);
  fl.alloc(m); pos.alloc(m); keyid.alloc(m); bad.alloc(1);
  FMM_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), st));
  k_pair_key<<<blocks(m, 256), 256, 0, st>>>(m, t.m2l_off.get(), t.m2l_src.get(), nc, t.level.get(), t.cx.get(),
                                             t.cy.get(), t.cz.get(), t.cr.get(), key.get(), tgt.get(), bad.get());
  FMM_CHECK(read_int(bad.get(), st) == 0, kUnsupported, "M2L offset beyond 2^17 half-widths (tree too deep)");
  // distinct keys
  k_iota_ll<<<blocks(m, 256), 256, 0, st>>>(m, order0.get());
  {
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.get(), key_sorted.get(), order0.get(), order.get(),
                                      (size_t)m, 0, 64, st));
    void* tmp = t.scratch.get(bytes);
    FMM_HIP(rocprim::radix_sort_pairs(tmp, bytes, key.get(), key_sorted.get(), order0.get(), order.get(),
                                      (size_t)m, 0, 64, st));
  }
  k_run_flags64<<<blocks(m, 256), 256, 0, st>>>(m, key_sorted.get(), fl.get());
  scan_ex(t.scratch, fl.get(), pos.get(), m, st);
  S.nkeys = read_int(pos.get() + m - 1, st) + read_int(fl.get() + m - 1, st);
  S.key_dx.alloc(3 * (size_t)S.nkeys);
  S.key_r.alloc(S.nkeys);
  S.key_dlevel.alloc(S.nkeys);
  k_key_ids<<<blocks(m, 256), 256, 0, st>>>(m, fl.get(), pos.get(), order.get(), tgt.get(), t.m2l_src.get(),
                                            t.level.get(), t.cx.get(), t.cy.get(), t.cz.get(), t.cr.get(),
                                            keyid.get(), S.key_dx.get(), S.key_r.get(), S.key_dlevel.get());
  // groups = runs of (tile, key)
  k_group_keys<<<blocks(m, 256), 256, 0, st>>>(m, tgt.get(), keyid.get(), tile_of.get(), key.get());
  k_iota_ll<<<blocks(m, 256), 256, 0, st>>>(m, order0.get());
  {
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.get(), key_sorted.get(), order0.get(), order.get(),
                                      (size_t)m, 0, 64, st));
    void* tmp = t.scratch.get(bytes);
    FMM_HIP(rocprim::radix_sort_pairs(tmp, bytes, key.get(), key_sorted.get(), order0.get(), order.get(),
                                      (size_t)m, 0, 64, st));
  }
  k_run_flags64<<<blocks(m, 256), 256, 0, st>>>(m, key_sorted.get(), fl.get());
  scan_ex(t.scratch, fl.get(), pos.get(), m, st);
  S.ngroups = read_int(pos.get() + m - 1, st) + read_int(fl.get() + m - 1, st);
  S.group_key.alloc(S.ngroups);
  S.group_src.alloc((size_t)S.ngroups * kTile);
  FMM_HIP(hipMemsetAsync(S.group_src.get(), 0xff, sizeof(int) * (size_t)S.ngroups * kTile, st));
  FMM_HIP(hipMemsetAsync(S.tile_group_off.get(), 0xff, sizeof(int) * (S.ntiles + 1), st));
  k_fill_groups<<<blocks(m, 256), 256, 0, st>>>(m, key_sorted.get(), order.get(), fl.get(), pos.get(), tgt.get(),
                                                t.m2l_src.get(), col_of.get(), S.group_key.get(), S.group_src.get(),
                                                S.tile_group_off.get());
  k_close_offsets<<<1, 1, 0, st>>>(S.ntiles, S.tile_group_off.get(), S.ngroups);
  FMM_HIP(hipStreamSynchronize(st));
  // work items: runs of groups of one tile
  std::vector<int> goff(S.ntiles + 1);
  FMM_HIP(hipMemcpy(goff.data(), S.tile_group_off.get(), sizeof(int) * (S.ntiles + 1), hipMemcpyDeviceToHost));
  std::vector<int> ct, c0, c1;
  // a tile's groups cut into equal runs of at most kGroupsPerChunk
  for (int tl = 0; tl < S.ntiles; ++tl) {
    const int G = goff[tl + 1] - goff[tl];
    const int nch = (G + kGroupsPerChunk - 1) / kGroupsPerChunk;
    for (int k = 0; k < nch; ++k) {
      ct.push_back(tl);
      c0.push_back(goff[tl] + (int)((long long)G * k / nch));
      c1.push_back(goff[tl] + (int)((long long)G * (k + 1) / nch));
    }
  }
  S.nchunks = (int)ct.size();
  S.chunk_tile.alloc(S.nchunks);
  S.chunk_g0.alloc(S.nchunks);
  S.chunk_g1.alloc(S.nchunks);
  FMM_HIP(hipMemcpy(S.chunk_tile.get(), ct.data(), sizeof(int) * S.nchunks, hipMemcpyHostToDevice));
  FMM_HIP(hipMemcpy(S.chunk_g0.get(), c0.data(), sizeof(int) * S.nchunks, hipMemcpyHostToDevice));
  FMM_HIP(hipMemcpy(S.chunk_g1.get(), c1.data(), sizeof(int) * S.nchunks, hipMemcpyHostToDevice));
}

}  // namespace fmm
