// Features, order propagation and bookkeeping of the learned order
// (semantics in learned.h).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "learned.h"

namespace fmm {
namespace {

inline int nb(long long n, int bs) { return (int)((n + bs - 1) / bs); }
constexpr int kAcc = 11;  // per-cell list sums

__global__ void k_zero_d(long long n, double* __restrict__ a) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

__global__ void k_fill_i(long long n, int* __restrict__ a, int v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// sum |S| and sum |S|^2 over the particles of each leaf
__global__ void k_leaf_qabs(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pstart,
                            const int* __restrict__ pcount, const double* __restrict__ pS, int ss,
                            double* __restrict__ qabs, double* __restrict__ q2) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  double a = 0.0, b = 0.0;
  for (int j = pstart[c]; j < pstart[c] + pcount[c]; ++j) {
    double s2 = 0.0;
    for (int k = 0; k < ss; ++k) s2 += pS[(size_t)j * ss + k] * pS[(size_t)j * ss + k];
    a += sqrt(s2);
    b += s2;
  }
  qabs[c] = a;
  q2[c] = b;
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

// list sums of every target cell: acc[c] = (S0, S3, S6, S9, S12 with
// w = sum|S| / R^2; T3, T6, T9, T12 with w = sqrt(sum |S|^2) / R^2; max rho;
// count), S_k = sum w rho^k; one wavefront per cell, lanes over its list
// Per-cell inputs of k_cell_lists packed for one gather per source cell:
// geo[c] = (cx, cy, cz, rmax) in FP64, qw[c] = (float(sum |S|),
// sqrtf(float(sum S^2))).
__global__ void k_pack_cells(int ncells, const double* __restrict__ cx, const double* __restrict__ cy,
                             const double* __restrict__ cz, const double* __restrict__ crmax,
                             const double* __restrict__ qabs, const double* __restrict__ q2,
                             double4* __restrict__ geo, float2* __restrict__ qw) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  geo[c] = make_double4(cx[c], cy[c], cz[c], crmax[c]);
  qw[c] = make_float2((float)qabs[c], sqrtf((float)q2[c]));
}

__global__ __launch_bounds__(64) void k_cell_lists(int ncells, const int* __restrict__ off,
                                                   const int* __restrict__ src, const double4* __restrict__ geo,
                                                   const float2* __restrict__ qw, double* __restrict__ acc) {
  const int A = blockIdx.x;
  if (A >= ncells) return;
  // features are estimates: per-pair arithmetic in FP32 (offsets from
  // FP64 differences), sums in FP64
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  float rmx = 0.f;
  const double4 ga = geo[A];
  const double xa = ga.x, ya = ga.y, za = ga.z;
  const float ra = (float)ga.w;
  for (int i = off[A] + threadIdx.x; i < off[A + 1]; i += 64) {
    const int B = src[i];
    const double4 gb = geo[B];
    const float2 qb = qw[B];
    const float dx = (float)(xa - gb.x), dy = (float)(ya - gb.y), dz = (float)(za - gb.z);
    const float R2 = dx * dx + dy * dy + dz * dz;
    const float rho = (ra + (float)gb.w) * __frsqrt_rn(R2);
    const float w = qb.x / R2, u = qb.y / R2;
    const float r3 = rho * rho * rho, r6 = r3 * r3, r9 = r6 * r3, r12 = r6 * r6;
    v[0] += w;
    v[1] += w * r3;
    v[2] += w * r6;
    v[3] += w * r9;
    v[4] += w * r12;
    v[5] += u * r3;
    v[6] += u * r6;
    v[7] += u * r9;
    v[8] += u * r12;
    rmx = fmaxf(rmx, rho);
  }
  for (int m = 32; m > 0; m >>= 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] += __shfl_xor(v[k], m, 64);
    rmx = fmaxf(rmx, __shfl_xor(rmx, m, 64));
  }
  if (threadIdx.x == 0) {
    double* a = acc + (size_t)A * kAcc;
    for (int k = 0; k < 9; ++k) a[k] = v[k];
    a[9] = rmx;
    a[10] = (double)(off[A + 1] - off[A]);
  }
}

// cells of one level add their parent's accumulated sums (root first)
__global__ void k_acc_down(int beg, int end, const int* __restrict__ parent, double* __restrict__ acc) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  const double* p = acc + (size_t)parent[c] * kAcc;
  double* a = acc + (size_t)c * kAcc;
  for (int k = 0; k < 9; ++k) a[k] += p[k];
  a[9] = fmax(a[9], p[9]);
  a[10] += p[10];
}

__global__ void k_leaf_feat(int nleaves, const int* __restrict__ leaves, const int* __restrict__ pcount,
                            const double* __restrict__ acc, const double* __restrict__ fnp, float* __restrict__ feat) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const double fn = *fnp;
  const int c = leaves[li];
  const double* a = acc + (size_t)c * kAcc;
  float* f = feat + (size_t)li * kMlpF;
  auto lg = [&](double v) { return (float)fmin(fmax(log10(fmax(v, 1e-300) / fmax(fn, 1e-300)), -40.0), 20.0); };
  for (int k = 0; k < 4; ++k) f[k] = lg(a[1 + k]);
  f[4] = (float)a[9];
  f[5] = (float)log10(1.0 + a[10]);
  f[6] = (float)log2((double)max(pcount[c], 1));
  f[7] = lg(a[0]);
  for (int k = 0; k < 4; ++k) f[8 + k] = lg(a[5 + k]);
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

__global__ void k_tile_order(int ntiles, int ts, const int* __restrict__ tile_cells, const int* __restrict__ cq,
                             signed char* __restrict__ to) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int q = 0;
  for (int k = 0; k < ts; ++k) {
    const int c = tile_cells[(size_t)t * ts + k];
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
// The same pair in the harmonic form: rows and columns with n_z <= 1,
// sum over n of (2n + 1) (q - n + 1)^2.
__device__ __forceinline__ double macs_harmonic(int q) {
  double a = 0.0;
  for (int n = 0; n <= q; ++n) a += (double)(2 * n + 1) * (q - n + 1) * (q - n + 1);
  return a;
}

// M2L pairs of every tile (valid columns of its groups), one block per tile
__global__ void k_tile_pairs(int ntiles, const int* __restrict__ goff, const int* __restrict__ gsrc,
                             long long* __restrict__ pairs) {
  __shared__ long long red[256];
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  long long c = 0;
  for (long long i = (long long)goff[t] * kTile + threadIdx.x; i < (long long)goff[t + 1] * kTile; i += blockDim.x)
    c += gsrc[i] >= 0;
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) pairs[t] = red[0];
}

// pairs of each tile times C(q + 6, 6)
__global__ void k_mac_frac(int ntiles, const long long* __restrict__ pairs, const signed char* __restrict__ to,
                           int pref, double* __restrict__ out) {
  __shared__ double r0[256], r1[256], r2[256];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0, c = 0.0;
  if (t < ntiles) {
    const int q = to ? (int)to[t] : pref;
    a = (double)pairs[t] * macs(q);
    b = (double)pairs[t] * macs(pref);
    c = (double)pairs[t] * macs_harmonic(q);
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  r2[threadIdx.x] = c;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
      r2[threadIdx.x] += r2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], r0[0]);
    atomicAdd(&out[1], r1[0]);
    atomicAdd(&out[2], r2[0]);
  }
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

// Orders of the leaves for one Lagrange multiplier lam (10^loglam): leaf l
// takes the order q minimising C(q + 6, 6) w_l + lam n_l e_l(q)^2, with
// e_l(q) = 10^(Y_l(q) + margin) its predicted rms error relative to the rms
// field, n_l its particles (2^f6) and w_l the M2L pairs on its ancestor
// chain (10^f5 - 1), the pairs whose order it raises.  Adds sum n_l e_l^2.
__global__ void k_alloc(int nl, const float* __restrict__ feat, const float* __restrict__ Y, const double* lw,
                        float margin, int qmin, int qmax, int* __restrict__ q_out, double* err_sum, int gated) {
  __shared__ double red[256];
  if (gated && lw[kLwRealloc] == 0.0) return;  // uniform: the second allocation is not needed
  const double loglam = lw[kLwLogLam];
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (li < nl) {
    const float* f = feat + (size_t)li * kMlpF;
    const double w = fmax(exp10((double)f[5]) - 1.0, 1.0);
    const double n = exp2((double)f[6]);
    const double lam = exp10(loglam);
    double best = 0.0, be = 0.0;
    int bq = -1;
    for (int o = 0; o < kMlpO; ++o) {
      const int q = o + kMlpOrder0;
      if (q < qmin || q > qmax) continue;
      const double e2 = exp10(2.0 * ((double)Y[(size_t)li * kMlpO + o] + margin));
      const double c = macs(q) * w + lam * n * e2;
      if (bq < 0 || c < best) {
        best = c;
        bq = q;
        be = e2;
      }
    }
    if (bq < 0) bq = qmax;
    q_out[li] = bq;
    contrib = n * be;
  }
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(err_sum, red[0]);
}

// sum over target cells of (M2L pairs) x C(q_cell + 6, 6)
__global__ void k_cell_macs(int ncells, const int* __restrict__ off, const int* __restrict__ cq,
                            double* __restrict__ out) {
  __shared__ double red[256];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  red[threadIdx.x] = c < ncells ? (double)(off[c + 1] - off[c]) * macs(cq[c]) : 0.0;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// cell -> M2L tile of the cell (-1: no M2L targets), once per tree
__global__ void k_cell_tile(int ntiles, int ts, const int* __restrict__ tile_cells, int* __restrict__ cell_tile) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)ntiles * ts) return;
  const int c = tile_cells[i];
  if (c >= 0) cell_tile[c] = (int)(i / ts);
}

// effective order of every cell: the lowest tile order over its ancestor
// chain (the M2L orders its particles' far field actually ran at), one
// level at a time from the root
__global__ void k_eff_down(int beg, int end, const int* __restrict__ parent, const int* __restrict__ cell_tile,
                           const signed char* __restrict__ tile_order, int qmax, int* __restrict__ eff) {
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  int q = c == 0 ? qmax : eff[parent[c]];
  const int tl = cell_tile[c];
  if (tl >= 0) q = min(q, (int)tile_order[tl]);
  eff[c] = q;
}

// sum over leaves of n_l e_l(q_eff)^2 with e = 10^(Y + margin)
__global__ void k_eff_err(int nl, const int* __restrict__ leaves, const float* __restrict__ feat,
                          const float* __restrict__ Y, const int* __restrict__ eff, float margin,
                          double* __restrict__ out) {
  __shared__ double red[256];
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (li < nl) {
    int o = eff[leaves[li]] - kMlpOrder0;
    o = o < 0 ? 0 : (o >= kMlpO ? kMlpO - 1 : o);
    const double e2 = exp10(2.0 * ((double)Y[(size_t)li * kMlpO + o] + margin));
    v = exp2((double)feat[(size_t)li * kMlpF + 6]) * e2;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// The error sums of kLam multipliers at once: lam_k = 10^(lo + k step).
constexpr int kLam = 64;
__global__ void k_alloc_multi(int nl, const float* __restrict__ feat, const float* __restrict__ Y,
                              const double* __restrict__ lw, float margin, int qmin, int qmax,
                              double* __restrict__ sums, int gated) {
  __shared__ float sh[kLam];
  __shared__ double lamk[kLam];
  if (lw[kLwSkip] != 0.0 || (gated && lw[kLwRealloc] == 0.0)) return;  // uniform over the grid
  const double lo = lw[kLwLo], step = lw[kLwStep];
  for (int k = threadIdx.x; k < kLam; k += blockDim.x) {
    sh[k] = 0.f;
    lamk[k] = exp10(lo + k * step);
  }
  __syncthreads();
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = li < nl;
  double n = 0.0, c[kMlpO], e2[kMlpO];
  for (int o = 0; o < kMlpO; ++o) c[o] = -1.0, e2[o] = 0.0;
  if (act) {
    const float* f = feat + (size_t)li * kMlpF;
    const double w = fmax(exp10((double)f[5]) - 1.0, 1.0);
    n = exp2((double)f[6]);
    for (int o = 0; o < kMlpO; ++o) {
      const int q = o + kMlpOrder0;
      const bool ok = q >= qmin && q <= qmax;
      c[o] = ok ? macs(q) * w : -1.0;
      e2[o] = exp10(2.0 * ((double)Y[(size_t)li * kMlpO + o] + margin));
    }
  }
  // every lane of the wave runs all kLam multipliers: the wave's sum of each
  // multiplier's n_l e_l^2 is one LDS atomic
  for (int k = 0; k < kLam; ++k) {
    const double lam = lamk[k] * n;
    double best = 0.0, be = 0.0;
    bool any = false;
    for (int o = 0; o < kMlpO; ++o) {
      if (c[o] < 0.0) continue;
      const double v = c[o] + lam * e2[o];
      if (!any || v < best) {
        best = v;
        be = e2[o];
        any = true;
      }
    }
    float contrib = act ? (float)(n * be) : 0.f;
    for (int m = 32; m > 0; m >>= 1) contrib += __shfl_xor(contrib, m, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&sh[k], contrib);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kLam; k += blockDim.x) atomicAdd(&sums[k], (double)sh[k]);
}

// Start of a multiplier search: target sum n_l e_l^2 = tol^2 n, coarse
// grid 10^-10 .. 10^40; tol from lw[tol_slot].
__global__ void k_lam_init(double* __restrict__ lw, int tol_slot, double n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double tol = lw[tol_slot];
  lw[kLwTarget] = tol * tol * n;
  lw[kLwLo] = -10.0;
  lw[kLwStep] = 50.0 / (kLam - 1);
  lw[kLwSkip] = 0.0;
  lw[kLwErr] = 0.0;
}

// Coarse grid result: the first multiplier whose orders meet the target;
// a bracket inside which the fine grid runs, or the final value at either
// end of the grid (the fine sweep is skipped).
__global__ void k_lam_coarse(const double* __restrict__ sums, double* __restrict__ lw) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double target = lw[kLwTarget], lo = lw[kLwLo], step = lw[kLwStep];
  int k = 0;
  while (k < kLam && sums[k] > target) ++k;
  if (k == kLam || k == 0) {  // even the highest orders miss (best effort), or the lowest meet it
    lw[kLwLogLam] = k == 0 ? lo : lo + (kLam - 1) * step;
    lw[kLwSkip] = 1.0;
  } else {
    lw[kLwLo] = lo + (k - 1) * step;
    lw[kLwStep] = step / (kLam - 1);
  }
}

__global__ void k_lam_fine(const double* __restrict__ sums, double* __restrict__ lw) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || lw[kLwSkip] != 0.0) return;
  const double target = lw[kLwTarget];
  int j = 0;
  while (j < kLam && sums[j] > target) ++j;
  lw[kLwLogLam] = lw[kLwLo] + (j < kLam - 1 ? j : kLam - 1) * lw[kLwStep];
}

// lw[slot] = sqrt(lw[kLwErr] / n): a predicted norm-wise error
__global__ void k_err_finish(double* __restrict__ lw, int slot, double n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lw[slot] = sqrt(lw[kLwErr] / n);
}

__global__ void k_lw_zero(double* __restrict__ lw, int slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lw[slot] = 0.0;
}

__global__ void k_lw_set(double* __restrict__ lw, int slot, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lw[slot] = v;
}

// Whether to re-allocate for the slack the tile maxima leave: with e1 the
// effective error of the first allocation, tol2 = min(tol^2 / e1, 4 tol)
// when e1 < 0.9 tol.
__global__ void k_realloc_decide(double* __restrict__ lw) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double tol = lw[kLwTol], e1 = lw[kLwEff1];
  const bool re = e1 > 0.0 && e1 < 0.9 * tol;
  lw[kLwRealloc] = re ? 1.0 : 0.0;
  lw[kLwTol2] = re ? fmin(tol * tol / e1, 4.0 * tol) : tol;
}

// Keeps the second allocation (leaf, cell and tile orders) when it was made
// and its effective error still meets the tolerance; lw[kLwEffFinal] = the
// effective error of the kept one.
__global__ void k_realloc_select(int nl, int nc, int nt, int* __restrict__ lq, const int* __restrict__ lq2,
                                 int* __restrict__ cq, const int* __restrict__ cq2, signed char* __restrict__ to,
                                 const signed char* __restrict__ to2, double* lw) {
  const bool keep2 = lw[kLwRealloc] != 0.0 && lw[kLwEff2] <= lw[kLwTol];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) lw[kLwEffFinal] = keep2 ? lw[kLwEff2] : lw[kLwEff1];
  if (!keep2) return;
  if (i < nl) lq[i] = lq2[i];
  if (i < nc) cq[i] = cq2[i];
  if (i < nt) to[i] = to2[i];
}

// F_s from the judge's direct sums Fd [J][outsz]: the rms |F| (field
// components fo .. outsz-1) without the largest max(1, J/1000) of them.
// One block; the dropped samples are found by repeated arg-max, the rest
// summed directly (no cancellation against a dominant sample).
constexpr int kScaleDropMax = 16;
__global__ __launch_bounds__(256) void k_field_scale(int J, int outsz, int fo, const double* __restrict__ Fd,
                                                     double* __restrict__ out) {
  __shared__ double vb[256];
  __shared__ int vi[256];
  __shared__ int drop[kScaleDropMax];
  const int tid = threadIdx.x;
  const int k = J > 0 ? min(max(1, J / 1000), kScaleDropMax) : 0;
  auto m2 = [&](int i) {
    double a = 0.0;
    for (int c = fo; c < outsz; ++c) a += Fd[(size_t)i * outsz + c] * Fd[(size_t)i * outsz + c];
    return a;
  };
  auto dropped = [&](int i, int r) {
    for (int u = 0; u < r; ++u)
      if (drop[u] == i) return true;
    return false;
  };
  for (int r = 0; r < k; ++r) {
    double best = -1.0;
    int bi = -1;
    for (int i = tid; i < J; i += 256) {
      const double v = m2(i);
      if (!dropped(i, r) && v > best) {
        best = v;
        bi = i;
      }
    }
    vb[tid] = best;
    vi[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s && (vb[tid + s] > vb[tid] || (vb[tid + s] == vb[tid] && vi[tid + s] < vi[tid] && vi[tid + s] >= 0))) {
        vb[tid] = vb[tid + s];
        vi[tid] = vi[tid + s];
      }
      __syncthreads();
    }
    if (tid == 0) drop[r] = vi[0];
    __syncthreads();
  }
  double sum = 0.0;
  for (int i = tid; i < J; i += 256)
    if (!dropped(i, k)) sum += m2(i);
  vb[tid] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) vb[tid] += vb[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    const int keep = J - k;
    *out = keep > 0 ? sqrt(vb[0] / keep) : 0.0;
  }
}

}  // namespace

void field_scale_from_samples(Tree& t, int J, int outsz, int fo, hipStream_t st) {
  learned_scalars(t);
  k_field_scale<<<1, 256, 0, st>>>(J, outsz, fo, t.judge_Fd.get(), t.lw.get() + kLwScale);
}

void ensure_cell_tile(Tree& t, hipStream_t st) {
  const int nc = t.ncells, nt = t.sched.ntiles;
  if (t.sched.cell_tile_ready) return;
  t.sched.cell_tile.alloc(nc);
  k_fill_i<<<nb(nc, 256), 256, 0, st>>>(nc, t.sched.cell_tile.get(), -1);
  if (nt > 0)
    k_cell_tile<<<nb((long long)nt * t.sched.tile_size, 256), 256, 0, st>>>(nt, t.sched.tile_size,
                                                                           t.sched.tile_cells.get(),
                                                                           t.sched.cell_tile.get());
  t.sched.cell_tile_ready = true;
}

void learned_scalars(Tree& t) {
  if (t.lw.n < (size_t)kLwCount) {
    t.lw.alloc(kLwCount);
    FMM_HIP(hipMemset(t.lw.get(), 0, sizeof(double) * kLwCount));
  }
}

void effective_error(Tree& t, const float* Y, double margin, int qmax, const signed char* tile_order, int slot,
                     hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves;
  k_lw_zero<<<1, 64, 0, st>>>(t.lw.get(), kLwErr);
  if (nl > 0) {
    ensure_cell_tile(t, st);
    t.cell_eff.alloc(nc);
    for (int l = 0; l < t.depth; ++l) {
      const int b = t.level_off[l], e = t.level_off[l + 1];
      k_eff_down<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.parent.get(), t.sched.cell_tile.get(), tile_order, qmax,
                                                 t.cell_eff.get());
    }
    k_eff_err<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaves.get(), t.leaf_feat.get(), Y, t.cell_eff.get(),
                                           (float)margin, t.lw.get() + kLwErr);
  }
  k_err_finish<<<1, 64, 0, st>>>(t.lw.get(), slot, (double)std::max<size_t>(t.n, 1));
}

double m2l_cell_macs(Tree& t, hipStream_t st) {
  const int nc = t.ncells;
  DevBuf<double> acc;
  acc.alloc(1);
  k_zero_d<<<1, 64, 0, st>>>(1, acc.get());
  k_cell_macs<<<nb(nc, 256), 256, 0, st>>>(nc, t.m2l_off.get(), t.cell_q.get(), acc.get());
  double h = 0.0;
  FMM_HIP(hipMemcpyAsync(&h, acc.get(), sizeof(double), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  return h;
}

void allocate_orders(Tree& t, const float* Y, int tol_slot, double margin, int qmin, int qmax, int* q_out,
                     int pred_slot, bool gated, hipStream_t st) {
  const int nl = t.nleaves;
  const double n = (double)std::max<size_t>(t.n, 1);
  t.lam_sums.alloc(kLam);
  // two passes of kLam multipliers (a coarse grid over 10^-10 .. 10^40, then
  // a fine grid inside the bracket): the smallest multiplier whose orders
  // meet the target; every decision on the device
  k_lam_init<<<1, 64, 0, st>>>(t.lw.get(), tol_slot, n);
  if (nl > 0) {
    const int g = gated ? 1 : 0;
    k_zero_d<<<1, kLam, 0, st>>>(kLam, t.lam_sums.get());
    k_alloc_multi<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaf_feat.get(), Y, t.lw.get(), (float)margin, qmin, qmax,
                                               t.lam_sums.get(), g);
    k_lam_coarse<<<1, 64, 0, st>>>(t.lam_sums.get(), t.lw.get());
    k_zero_d<<<1, kLam, 0, st>>>(kLam, t.lam_sums.get());
    k_alloc_multi<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaf_feat.get(), Y, t.lw.get(), (float)margin, qmin, qmax,
                                               t.lam_sums.get(), g);
    k_lam_fine<<<1, 64, 0, st>>>(t.lam_sums.get(), t.lw.get());
    // the orders of the chosen multiplier
    k_alloc<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaf_feat.get(), Y, t.lw.get(), (float)margin, qmin, qmax, q_out,
                                         t.lw.get() + kLwErr, g);
  }
  k_err_finish<<<1, 64, 0, st>>>(t.lw.get(), pred_slot, n);
}

void chunk_classes_async(Tree& t, hipStream_t st) {
  constexpr int NQ = M2LSchedule::kMaxOrder;
  M2LSchedule& S = t.sched;
  const int nch = S.nchunks;
  S.class_start.alloc(NQ + 1);
  k_fill_i<<<1, 64, 0, st>>>(NQ + 1, S.class_start.get(), -1);
  if (nch == 0) return;
  S.class_keys.alloc(2 * (size_t)nch);
  unsigned long long* k0 = S.class_keys.get();
  unsigned long long* k1 = k0 + nch;
  S.class_chunks.alloc(nch);
  k_class_keys<<<nb(nch, 256), 256, 0, st>>>(nch, S.chunk_tile.get(), t.tile_order.get(), k0);
  size_t bytes = 0;
  FMM_HIP(rocprim::radix_sort_keys(nullptr, bytes, k0, k1, (size_t)nch, 0, 40, st));
  void* tmp = t.scratch.get(bytes);
  FMM_HIP(rocprim::radix_sort_keys(tmp, bytes, k0, k1, (size_t)nch, 0, 40, st));
  k_class_split<<<nb(nch, 256), 256, 0, st>>>(nch, k1, S.class_chunks.get(), S.class_start.get());
}

int read_orders(Tree& t, hipStream_t st, int* start) {
  constexpr int NQ = M2LSchedule::kMaxOrder;
  const int nch = t.sched.nchunks;
  std::vector<int> h(NQ + 2, -1);
  FMM_HIP(hipMemcpyAsync(h.data(), t.cell_q.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  if (nch > 0)
    FMM_HIP(hipMemcpyAsync(h.data() + 1, t.sched.class_start.get(), sizeof(int) * (NQ + 1), hipMemcpyDeviceToHost,
                           st));
  FMM_HIP(hipStreamSynchronize(st));
  // empty classes start where the next non-empty one does
  int next = nch;
  for (int q = NQ; q >= 0; --q) {
    if (h[1 + q] >= 0) next = h[1 + q];
    start[q] = next;
  }
  return h[0];
}

void leaf_features(Tree& t, int ss, hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves;
  t.cell_qabs.alloc(nc);
  t.cell_q2.alloc(nc);
  t.cell_acc.alloc((size_t)nc * kAcc);
  t.leaf_feat.alloc((size_t)std::max(nl, 1) * kMlpF);
  learned_scalars(t);
  k_zero_d<<<nb(nc, 256), 256, 0, st>>>(nc, t.cell_qabs.get());
  k_zero_d<<<nb(nc, 256), 256, 0, st>>>(nc, t.cell_q2.get());
  if (nl == 0) return;
  k_leaf_qabs<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.pS.get(), ss,
                                          t.cell_qabs.get(), t.cell_q2.get());
  for (int l = t.depth - 2; l >= 0; --l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_parent_sum<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cell_qabs.get());
    k_parent_sum<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cell_q2.get());
  }
  t.cell_geo4.alloc((size_t)nc * 4);
  t.cell_qw.alloc((size_t)nc * 2);
  double4* geo = reinterpret_cast<double4*>(t.cell_geo4.get());
  float2* qw = reinterpret_cast<float2*>(t.cell_qw.get());
  k_pack_cells<<<nb(nc, 256), 256, 0, st>>>(nc, t.cx.get(), t.cy.get(), t.cz.get(), t.crmax.get(),
                                           t.cell_qabs.get(), t.cell_q2.get(), geo, qw);
  k_cell_lists<<<nc, 64, 0, st>>>(nc, t.m2l_off.get(), t.m2l_src.get(), geo, qw, t.cell_acc.get());
  for (int l = 1; l < t.depth; ++l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_acc_down<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.parent.get(), t.cell_acc.get());
  }
  k_leaf_feat<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pcount.get(), t.cell_acc.get(),
                                          t.lw.get() + kLwScale, t.leaf_feat.get());
}

void leaf_error(Tree& t, const double* F, const double* Fref, int outsz, int fo, double* err, hipStream_t st) {
  const int nl = t.nleaves;
  if (nl == 0) return;
  k_leaf_err<<<nb(nl, 128), 128, 0, st>>>(nl, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.perm.get(), outsz,
                                         fo, F, Fref, err);
}

void orders_from_leaves(Tree& t, const int* leaf_q, int* cell_q, signed char* tile_order, hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves, nt = t.sched.ntiles;
  k_fill_i<<<nb(nc, 256), 256, 0, st>>>(nc, cell_q, 0);
  if (nl > 0) k_leaf_to_cell<<<nb(nl, 256), 256, 0, st>>>(nl, t.leaves.get(), leaf_q, cell_q);
  for (int l = t.depth - 2; l >= 0; --l) {
    const int b = t.level_off[l], e = t.level_off[l + 1];
    k_parent_max<<<nb(e - b, 256), 256, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), cell_q);
  }
  if (nt > 0)
    k_tile_order<<<nb(nt, 256), 256, 0, st>>>(nt, t.sched.tile_size, t.sched.tile_cells.get(), cell_q, tile_order);
}

void select_orders(Tree& t, const float* Y, double margin, int qmin, int qmax, bool refine, hipStream_t st) {
  const int nc = t.ncells, nl = t.nleaves, nt = t.sched.ntiles;
  t.leaf_q.alloc(std::max(nl, 1));
  t.leaf_q2.alloc(std::max(nl, 1));
  t.cell_q.alloc(nc);
  t.cell_q2a.alloc(nc);
  t.tile_order.alloc(std::max(nt, 1));
  t.tile_order2.alloc(std::max(nt, 1));
  k_lw_set<<<1, 64, 0, st>>>(t.lw.get(), kLwTol, t.tol);
  allocate_orders(t, Y, kLwTol, margin, qmin, qmax, t.leaf_q.get(), kLwPred, false, st);
  orders_from_leaves(t, t.leaf_q.get(), t.cell_q.get(), t.tile_order.get(), st);
  if (!refine) return;
  // the tile maxima leave the predicted error below the budget: allocate
  // once more for the tolerance scaled by that slack, keep it when its
  // effective error still meets the tolerance
  effective_error(t, Y, margin, qmax, t.tile_order.get(), kLwEff1, st);
  k_realloc_decide<<<1, 64, 0, st>>>(t.lw.get());
  allocate_orders(t, Y, kLwTol2, margin, qmin, qmax, t.leaf_q2.get(), kLwPred2, true, st);
  orders_from_leaves(t, t.leaf_q2.get(), t.cell_q2a.get(), t.tile_order2.get(), st);
  effective_error(t, Y, margin, qmax, t.tile_order2.get(), kLwEff2, st);
  const int m = std::max(std::max(nl, nc), nt);
  k_realloc_select<<<nb(m, 256), 256, 0, st>>>(nl, nc, nt, t.leaf_q.get(), t.leaf_q2.get(), t.cell_q.get(),
                                               t.cell_q2a.get(), t.tile_order.get(), t.tile_order2.get(),
                                               t.lw.get());
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

double m2l_mac_fraction(Tree& t, int pref, hipStream_t st, double* macs_out, double* hmacs_out) {
  const int nt = t.sched.ntiles;
  if (macs_out) *macs_out = 0.0;
  if (hmacs_out) *hmacs_out = 0.0;
  if (nt == 0) return 1.0;
  DevBuf<double> acc;
  acc.alloc(3);
  k_zero_d<<<1, 64, 0, st>>>(3, acc.get());
  if (!t.sched.tile_pairs_ready) {  // order independent: once per tile set
    t.sched.tile_pairs.alloc(nt);
    k_tile_pairs<<<nt, 256, 0, st>>>(nt, t.sched.tile_group_off.get(), t.sched.group_src.get(),
                                     t.sched.tile_pairs.get());
    t.sched.tile_pairs_ready = true;
  }
  k_mac_frac<<<nb(nt, 256), 256, 0, st>>>(nt, t.sched.tile_pairs.get(), t.tile_order_ptr(), pref, acc.get());
  double h[3] = {0, 0, 0};
  FMM_HIP(hipMemcpyAsync(h, acc.get(), sizeof(h), hipMemcpyDeviceToHost, st));
  FMM_HIP(hipStreamSynchronize(st));
  if (macs_out) *macs_out = h[0];
  if (hmacs_out) *hmacs_out = h[2];
  return h[1] > 0 ? h[0] / h[1] : 1.0;
}

}  // namespace fmm
