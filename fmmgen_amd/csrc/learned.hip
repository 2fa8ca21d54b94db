// Features, order propagation and bookkeeping of the learned order
// (semantics in learned.h).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "learned.h"

namespace fmm {
namespace {

inline int nb(long long n, int bs) { return (int)((n + bs - 1) / bs); }
constexpr int kAcc = 7;  // per-cell list sums

__global__ void k_zero_d(long long n, double* __restrict__ a) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

__global__ void k_fill_i(long long n, int* __restrict__ a, int v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// sum |S| over the particles of each leaf
__global__ void k_leaf_qabs(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pstart,
                            const int* __restrict__ pcount, const double* __restrict__ pS, int ss,
                            double* __restrict__ qabs) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  double a = 0.0;
  for (int j = pstart[c]; j < pstart[c] + pcount[c]; ++j) {
    double s2 = 0.0;
    for (int k = 0; k < ss; ++k) s2 += pS[(size_t)j * ss + k] * pS[(size_t)j * ss + k];
    a += sqrt(s2);
  }
  qabs[c] = a;
}

// parents of one level: sum of the children (children are contiguous)
__global__ void k_parent_sum(int beg, int end, const int* __restrict__ first_child, const int* __restrict__ nchild,
                             double* __restrict__ v) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end || nchild[c] == 0) return;
  double a = 0.0;
  for (int k = first_child[c]; k < first_child[c] + nchild[c]; ++k) a += v[k];
  v[c] = a;
}

__global__ void k_parent_max(int beg, int end, const int* __restrict__ first_child, const int* __restrict__ nchild,
                             int* __restrict__ q) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end || nchild[c] == 0) return;
  int a = 0;
  for (int k = first_child[c]; k < first_child[c] + nchild[c]; ++k) a = max(a, q[k]);
  q[c] = a;
}

// list sums of every target cell: acc[c] = (S0, S3, S6, S9, S12, max rho, count)
__global__ void k_cell_lists(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                             const double* __restrict__ cx, const double* __restrict__ cy,
                             const double* __restrict__ cz, const double* __restrict__ crmax,
                             const double* __restrict__ qabs, double* __restrict__ acc) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells) return;
  double s0 = 0, s3 = 0, s6 = 0, s9 = 0, s12 = 0, rmx = 0;
  const double xa = cx[A], ya = cy[A], za = cz[A], ra = crmax[A];
  for (int i = off[A]; i < off[A + 1]; ++i) {
    const int B = src[i];
    const double dx = xa - cx[B], dy = ya - cy[B], dz = za - cz[B];
    const double R2 = dx * dx + dy * dy + dz * dz;
    const double rho = (ra + crmax[B]) * rsqrt(R2);
    const double w = qabs[B] / R2;
    const double r3 = rho * rho * rho, r6 = r3 * r3;
    s0 += w;
    s3 += w * r3;
    s6 += w * r6;
    s9 += w * r6 * r3;
    s12 += w * r6 * r6;
    rmx = fmax(rmx, rho);
  }
  double* a = acc + (size_t)A * kAcc;
  a[0] = s0;
  a[1] = s3;
  a[2] = s6;
  a[3] = s9;
  a[4] = s12;
  a[5] = rmx;
  a[6] = (double)(off[A + 1] - off[A]);
}

// cells of one level add their parent's accumulated sums (root first)
__global__ void k_acc_down(int beg, int end, const int* __restrict__ parent, double* __restrict__ acc) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  const double* p = acc + (size_t)parent[c] * kAcc;
  double* a = acc + (size_t)c * kAcc;
  for (int k = 0; k < 5; ++k) a[k] += p[k];
  a[5] = fmax(a[5], p[5]);
  a[6] += p[6];
}

// sum of squares of the field components of the sorted near field
__global__ void k_sumsq(long long n, int outsz, int fo, const double* __restrict__ F, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    for (int k = fo; k < outsz; ++k) s += F[i * outsz + k] * F[i * outsz + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

__global__ void k_leaf_feat(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pcount,
                            const double* __restrict__ acc,
                            const double* __restrict__ fscale, long long n, float* __restrict__ feat) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  const double fn = sqrt(fscale[0] / (double)(n > 0 ? n : 1));
  const double* a = acc + (size_t)c * kAcc;
  float* f = feat + (size_t)li * kMlpF;
  auto lg = [&](double v) { return (float)fmin(fmax(log10(fmax(v, 1e-300) / fmax(fn, 1e-300)), -40.0), 20.0); };
  for (int k = 0; k < 4; ++k) f[k] = lg(a[1 + k]);
  f[4] = (float)a[5];
  f[5] = (float)log10(1.0 + a[6]);
  f[6] = (float)log2((double)max(pcount[c], 1));
  f[7] = lg(a[0]);
}

__global__ void k_leaf_err(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pstart,
                           const int* __restrict__ pcount, const int* __restrict__ perm, int outsz, int fo,
                           const double* __restrict__ F, const double* __restrict__ Fr, double* __restrict__ err) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  double e = 0.0;
  for (int j = pstart[c]; j < pstart[c] + pcount[c]; ++j) {
    const size_t i = (size_t)perm[j];
    for (int k = fo; k < outsz; ++k) {
      const double d = F[i * outsz + k] - Fr[i * outsz + k];
      e += d * d;
    }
  }
  err[li] = sqrt(e / (double)max(pcount[c], 1));
}

__global__ void k_leaf_to_cell(int nleaves, const int* __restrict__ leaves, const int* __restrict__ lq,
                               int* __restrict__ cq) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li < nleaves) cq[leaves[li]] = lq[li];
}

__global__ void k_tile_order(int ntiles, const int* __restrict__ tile_cells, const int* __restrict__ cq,
                             signed char* __restrict__ to) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int q = 0;
  for (int k = 0; k < kTile; ++k) {
    const int c = tile_cells[t * kTile + k];
    if (c >= 0) q = max(q, cq[c]);
  }
  to[t] = (signed char)q;
}

__global__ void k_raise(int nleaves, const int* __restrict__ leaves, const int* __restrict__ flag,
                        int* __restrict__ lq, int qfull) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li < nleaves && flag[leaves[li]]) lq[li] = max(lq[li], qfull);
}

__device__ __forceinline__ double macs(int q) {  // C(q + 6, 6)
  double a = 1.0;
  for (int k = 1; k <= 6; ++k) a = a * (q + k) / k;
  return a;
}

// pairs of each tile (valid group columns) times C(q + 6, 6)
__global__ void k_mac_frac(int ntiles, const int* __restrict__ goff, const int* __restrict__ gsrc,
                           const signed char* __restrict__ to, int pref, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  long long pairs = 0;
  for (long long i = (long long)goff[t] * kTile; i < (long long)goff[t + 1] * kTile; ++i) pairs += gsrc[i] >= 0;
  const int q = to ? (int)to[t] : pref;
  atomicAdd(&out[0], (double)pairs * macs(q));
  atomicAdd(&out[1], (double)pairs * macs(pref));
}

__global__ void k_class_keys(int nch, const int* __restrict__ chunk_tile, const signed char* __restrict__ to,
                             unsigned long long* __restrict__ key) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nch) key[c] = ((unsigned long long)(unsigned)to[chunk_tile[c]] << 32) | (unsigned)c;
}

__global__ void k_class_split(int nch, const unsigned long long* __restrict__ key, int* __restrict__ chunks,
                              int* __restrict__ start) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nch) return;
  chunks[i] = (int)(key[i] & 0xffffffffu);
  const int q = (int)(key[i] >> 32);
  if (i == 0 || (int)(key[i - 1] >> 32) != q) start[q] = i;
}

}  // namespace

void chunk_classes(Tree& t, hipStream_t st, int* start) {
  constexpr int NQ = M2LSchedule::kMaxOrder;
  M2LSchedule& S = t.sched;
  const int nch = S.nchunks;
  for (int q = 0; q <= NQ; ++q) start[q] = nch;
  if (nch == 0) return;
  DevBuf<unsigned long long> k0, k1;
  DevBuf<int> st_d;
  k0.alloc(nch);
  k1.alloc(nch);
  st_d.alloc(NQ + 1);
  S.class_chunks.alloc(nch);
  k_fill_i<<<1, 64, 0, st>>>(NQ + 1, st_d.get(), -1);
  k_class_keys<<<nb(nch, 256), 256, 0, st>>>(nch, S.chunk_tile.get(), t.tile_order.get(), k0.get());
  size_t bytes = 0;
  FMM_HIP(rocprim::radix_sort_keys(nullptr, bytes, k0.get(), k1.get(), (size_t)nch, 0, 40, st));
  void* tmp = t.scratch.get(bytes);
  FMM_HIP(rocprim::radix_sort_keys(tmp, bytes, k0.get(), k1.get(), (size_t)nch, 0, 40, st));
  k_class_split<<<nb(nch, 256), 256, 0, st>>>(nch, k1.get(), S.class_chunks.get(), st_d.get());
  std::vector<int> h(NQ + 1);
  FMM_HIP(hipMemcpyAsync(h.data(), st_d.get(), sizeof(int) * (NQ + 1), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  // empty classes start where the next non-empty one does
  int next = nch;
  for (int q = NQ; q >= 0; --q) {
    if (h[q] >= 0) next = h[q];
    start[q] = next;
  }
}

void leaf_features(Tree& t, int ss, int outsz, int fo, hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves;
  t.cell_qabs.alloc(nc);
  t.cell_acc.alloc((size_t)nc * kAcc);
  t.fscale.alloc(2);
  t.leaf_feat.alloc((size_t)std::max(nl, 1) * kMlpF);
  k_zero_d<<<nb(nc, 256), 256, 0, st>>>(nc, t.cell_qabs.get());
  k_zero_d<<<1, 64, 0, st>>>(2, t.fscale.get());
  if (nl == 0) return;
  k_leaf_qabs<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.pS.get(), ss,
                                          t.cell_qabs.get());
  for (int l = t.depth - 2; l >= 0; --l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_parent_sum<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cell_qabs.get());
  }
  k_cell_lists<<<nb(nc, 128), 128, 0, st>>>(nc, t.m2l_off.get(), t.m2l_src.get(), t.cx.get(), t.cy.get(), t.cz.get(),
                                           t.crmax.get(), t.cell_qabs.get(), t.cell_acc.get());
  for (int l = 1; l < t.depth; ++l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_acc_down<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.parent.get(), t.cell_acc.get());
  }
  const long long n = (long long)t.n;
  k_sumsq<<<std::min(nb(n, 256), 2048), 256, 0, st>>>(n, outsz, fo, t.Fp2p.get(), t.fscale.get());
  k_leaf_feat<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pcount.get(), t.cell_acc.get(), t.fscale.get(), n,
                                          t.leaf_feat.get());
}

void leaf_error(Tree& t, const double* F, const double* Fref, int outsz, int fo, double* err, hipStream_t st) {
  const int nl = t.nleaves;
  if (nl == 0) return;
  k_leaf_err<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.perm.get(), outsz,
                                         fo, F, Fref, err);
}

int orders_from_leaves(Tree& t, hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves;
  t.cell_q.alloc(nc);
  k_fill_i<<<nb(nc, 256), 256, 0, st>>>(nc, t.cell_q.get(), 0);
  k_leaf_to_cell<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaves.get(), t.leaf_q.get(), t.cell_q.get());
  for (int l = t.depth - 2; l >= 0; --l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_parent_max<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cell_q.get());
  }
  const int nt = t.sched.ntiles;
  t.tile_order.alloc(std::max(nt, 1));
  if (nt > 0)
    k_tile_order<<<nb(nt, 256), 256, 0, st>>>(nt, t.sched.tile_cells.get(), t.cell_q.get(), t.tile_order.get());
  int qmax = 0;
  FMM_HIP(hipMemcpyAsync(&qmax, t.cell_q.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  return qmax;
}

void raise_failing_leaves(Tree& t, int qfull, hipStream_t st) {
  const int nl = t.nleaves;
  if (nl > 0)
    k_raise<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaves.get(), t.judge_fail.get() + 1, t.leaf_q.get(), qfull);
}

void raise_all_leaves(Tree& t, int qfull, hipStream_t st) {
  const int nl = t.nleaves;
  if (nl > 0) k_fill_i<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaf_q.get(), qfull);
}

double m2l_mac_fraction(Tree& t, int pref, hipStream_t st, double* macs_out) {
  const int nt = t.sched.ntiles;
  if (macs_out) *macs_out = 0.0;
  if (nt == 0) return 1.0;
  DevBuf<double> acc;
  acc.alloc(2);
  k_zero_d<<<1, 64, 0, st>>>(2, acc.get());
  k_mac_frac<<<nb(nt, 64), 64, 0, st>>>(nt, t.sched.tile_group_off.get(), t.sched.group_src.get(),
                                        t.tile_order_ptr(), pref, acc.get());
  double h[2] = {0, 0};
  FMM_HIP(hipMemcpyAsync(h, acc.get(), sizeof(h), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  if (macs_out) *macs_out = h[0];
  return h[1] > 0 ? h[0] / h[1] : 1.0;
}

}  // namespace fmm
