// Mixed-precision far field of one operator variant.
//
// Expansions are stored normalised by the cell half-width r and in FP32:
//     Mhat_n = M_n / r^|n|,            Lhat_n = L_n r^(|n|+1).
// In these variables every operator is the generated body applied to
// coordinates divided by a half-width, with power-of-two corrections:
//   P2M  Mhat = r^-so P2M_S(x/r, S)                 x = centre - particle
//   M2M  Mhat_p = M2M(x/r_p, Mhat_c * 2^-|j|)       x = parent - child
//   M2L  Lhat_A = sum Dhat Mhat_B 2^((lA-lB)|m|)    (MFMA GEMM, m2l_gemm.h)
//   L2L  Lhat_c = 2^-(|n|+1) L2L(x/r_p, Lhat_p)     x = child - parent
//   L2P  phi = phihat / r, F = Fhat / r^2           (L2P(x/r, Lhat))
// M2M and L2L of parents on the top three levels run in FP64 at order >= 8
// (the terms of the binomial shifts grow there); the near and far parts of
// each target are summed in FP64.
#pragma once

#include "m2l_gemm.h"
#include "variant_decl.h"

namespace fmm {
namespace FMM_VARIANT_NS {

constexpr int kTopLevels = 3;

template <int N>
struct DegTab {
  signed char d[N];
  constexpr DegTab() : d{} {
    for (int j = 0; j < N; ++j) d[j] = (signed char)slot_degree(j);
  }
};

// Reads of a normalised child multipole, rescaled to the parent half-width.
template <typename S, int OFF>
struct HalfScaledIn {
  const S* p;
  __device__ __forceinline__ float operator[](int j) const {
    constexpr DegTab<512> tab{};
    return ldexpf((float)p[j], -tab.d[j + OFF]);
  }
};
template <int OFF>
struct HalfScaledIn64 {
  const double* p;
  __device__ __forceinline__ double operator[](int j) const {
    constexpr DegTab<512> tab{};
    return ldexp(p[j], -tab.d[j + OFF]);
  }
};

// FP32 P2M, thread per (leaf, chunk); leaves on the top levels also keep an
// FP64 copy for the FP64 top shifts.
template <int P>
__global__ __launch_bounds__(64) void k_p2m_mixed(int nleaves, const int* __restrict__ leaves,
                                                  const int* __restrict__ pstart, const int* __restrict__ pcount,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const double* __restrict__ cz, const double* __restrict__ cr,
                                                  const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, const double* __restrict__ pS,
                                                  float* __restrict__ Mf, double* __restrict__ Mtop, int ntop) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk;
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int c = leaves[li];
  const double x0 = cx[c], y0 = cy[c], z0 = cz[c], ri = 1.0 / cr[c];
  double sscale = 1.0;
#pragma unroll
  for (int k = 0; k < SO; ++k) sscale *= ri;
  const int b = pstart[c], e = b + pcount[c];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<float, LO, HI> acc;
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) acc.v[k] = 0.f;
    for (int j = b; j < e; ++j) {
      float s[SS];
#pragma unroll
      for (int k = 0; k < SS; ++k) s[k] = (float)(pS[(size_t)j * SS + k] * sscale);
      fmmgen_ops::P2M_S_chunk<P, LO / kChunk, float>((float)((x0 - px[j]) * ri), (float)((y0 - py[j]) * ri),
                                    (float)((z0 - pz[j]) * ri), s, acc);
    }
#pragma unroll
    for (int k = 0; k < HI - LO; ++k) {
      Mf[(size_t)c * MS + LO + k] = acc.v[k];
      if (c < ntop) Mtop[(size_t)c * MS + LO + k] = acc.v[k];
    }
  });
}

// FP32 P2M of charges, one wave per leaf: the normalised multipole of a
// charge q at u = (x_centre - x_particle) / r is
// M[a,b,c] = q u^a v^b w^c / (a! b! c!) (the closed form of the generated
// P2M_S body; multi-index layout of the expansion).  Lanes own the slots lane, lane + 64, ...;
// 64 particles at a time stage their scaled powers u^k / k! (k <= P) in
// LDS, and every lane adds q X[a] Y[b] Z[c] for its slots.  Particle data
// are read coalesced once per leaf, and the lanes stay converged (the
// thread-per-(leaf, chunk) layout loops each lane over its own leaf).
template <int P>
__global__ __launch_bounds__(64) void k_p2m_wave(int nleaves, const int* __restrict__ leaves,
                                                 const int* __restrict__ pstart, const int* __restrict__ pcount,
                                                 const double* __restrict__ cx, const double* __restrict__ cy,
                                                 const double* __restrict__ cz, const double* __restrict__ cr,
                                                 const double* __restrict__ px, const double* __restrict__ py,
                                                 const double* __restrict__ pz, const double* __restrict__ pS,
                                                 float* __restrict__ Mf, double* __restrict__ Mtop, int ntop) {
  static_assert(SO == 0 || P < 0, "charges only");
  constexpr int MS = Op<P>::MS, NPL = (MS + 63) / 64, PW = 3 * (P + 1);
  __shared__ float pw[64 * PW];  // particle k: X[0..P], Y[0..P], Z[0..P]
  __shared__ float sq[64];
  const int li = blockIdx.x;
  if (li >= nleaves) return;
  const int lane = threadIdx.x;
  const int c = leaves[li];
  const double x0 = cx[c], y0 = cy[c], z0 = cz[c], ri = 1.0 / cr[c];
  int ea[NPL], eb[NPL], ec[NPL];
  float acc[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int sl = lane + 64 * i;
    const MultiIdx m = slot_multi_rt(sl < MS ? sl : 0);
    ea[i] = m.x;
    eb[i] = P + 1 + m.y;
    ec[i] = 2 * (P + 1) + m.z;
    acc[i] = 0.f;
  }
  const int b = pstart[c], e = b + pcount[c];
  for (int base = b; base < e; base += 64) {
    const int j = base + lane;
    __syncthreads();  // previous particles consumed
    {
      const bool in = j < e;
      const float u[3] = {in ? (float)((x0 - px[j]) * ri) : 0.f, in ? (float)((y0 - py[j]) * ri) : 0.f,
                          in ? (float)((z0 - pz[j]) * ri) : 0.f};
      sq[lane] = in ? (float)pS[j] : 0.f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float v = 1.f;
        pw[lane * PW + d * (P + 1)] = 1.f;
#pragma unroll
        for (int k = 1; k <= P; ++k) {
          v = v * u[d] * (1.f / (float)k);
          pw[lane * PW + d * (P + 1) + k] = v;
        }
      }
    }
    __syncthreads();
    const int nk = min(64, e - base);
    for (int k = 0; k < nk; ++k) {
      const float q = sq[k];
      const float* w = pw + k * PW;
#pragma unroll
      for (int i = 0; i < NPL; ++i) acc[i] = fmaf(q, w[ea[i]] * w[eb[i]] * w[ec[i]], acc[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int sl = lane + 64 * i;
    if (sl < MS) {
      Mf[(size_t)c * MS + sl] = acc[i];
      if (c < ntop) Mtop[(size_t)c * MS + sl] = acc[i];
    }
  }
}

// FP32 M2M of the parents of one level (below the FP64 top levels).
template <int P>
__global__ __launch_bounds__(64) void k_m2m_mixed(int beg, int end, const int* __restrict__ first_child,
                                                  const int* __restrict__ nchild, const double* __restrict__ cx,
                                                  const double* __restrict__ cy, const double* __restrict__ cz,
                                                  const double* __restrict__ cr, float* __restrict__ Mf) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk, OFF = nterms(SO - 1);
  const int p = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= end) return;
  const int k = nchild[p];
  if (k == 0) return;
  const int fc = first_child[p];
  const double ri = 1.0 / cr[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<float, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.f;
    for (int c = fc; c < fc + k; ++c)
      fmmgen_ops::M2M_chunk<P, LO / kChunk, float>((float)((cx[p] - cx[c]) * ri), (float)((cy[p] - cy[c]) * ri),
                                  (float)((cz[p] - cz[c]) * ri),
                                  HalfScaledIn<float, OFF>{Mf + (size_t)c * MS}, acc);
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) Mf[(size_t)p * MS + LO + j] = acc.v[j];
  });
}

// FP32 M2M of one level in two passes.  First one thread per (child, slot
// chunk) translates the child's multipole to its parent's centre into the
// child's row of Mpart (rows from cbeg, the level's first child); then
// k_m2m_sum adds each parent's rows in child order.  Against one thread per
// parent looping over its children (k_m2m_mixed) this gives eight times the
// threads and one child's chain of loads per thread, which pays on levels of
// few cells, bound by that chain's latency (at order 12 on 2^22 blobs: a
// level of 1344 children 89 -> 24 us); on large levels, bound by the
// generated bodies' issue at two waves per SIMD, the second pass's traffic
// makes it slower (88832 children 436 -> 482 us).
constexpr int kM2MChildMax = 32768;
template <int P>
__global__ __launch_bounds__(64) void k_m2m_child(int cbeg, int cend, const int* __restrict__ parent,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const double* __restrict__ cz, const double* __restrict__ cr,
                                                  const float* __restrict__ Mf, float* __restrict__ Mpart) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk, OFF = nterms(SO - 1);
  const int c = cbeg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cend) return;
  const int p = parent[c];
  const double ri = 1.0 / cr[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<float, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.f;
    fmmgen_ops::M2M_chunk<P, LO / kChunk, float>((float)((cx[p] - cx[c]) * ri), (float)((cy[p] - cy[c]) * ri),
                                                 (float)((cz[p] - cz[c]) * ri),
                                                 HalfScaledIn<float, OFF>{Mf + (size_t)c * MS}, acc);
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) Mpart[(size_t)(c - cbeg) * MS + LO + j] = acc.v[j];
  });
}

// Second pass: one thread per (parent, slot), the children's rows summed in
// child order.
template <int P>
__global__ void k_m2m_sum(int beg, int end, int cbeg, const int* __restrict__ first_child,
                          const int* __restrict__ nchild, const float* __restrict__ Mpart,
                          float* __restrict__ Mf) {
  constexpr int MS = Op<P>::MS;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)(end - beg) * MS) return;
  const int p = beg + (int)(i / MS), j = (int)(i % MS);
  const int k = nchild[p];
  if (k == 0) return;
  const float* row = Mpart + (size_t)(first_child[p] - cbeg) * MS + j;
  float s = 0.f;
  for (int c = 0; c < k; ++c) s += row[(size_t)c * MS];
  Mf[(size_t)p * MS + j] = s;
}

// FP64 M2M of a top-level parent; children below the top read the FP32
// array.  The result is kept in FP64 and rounded into the FP32 array for M2L.
template <int P>
__global__ __launch_bounds__(64) void k_m2m_top(int beg, int end, const int* __restrict__ first_child,
                                                const int* __restrict__ nchild, const double* __restrict__ cx,
                                                const double* __restrict__ cy, const double* __restrict__ cz,
                                                const double* __restrict__ cr, float* __restrict__ Mf,
                                                double* __restrict__ Mtop, int ntop) {
  constexpr int MS = Op<P>::MS, NCH = (MS + kChunk - 1) / kChunk, OFF = nterms(SO - 1);
  const int p = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= end) return;
  const int k = nchild[p];
  if (k == 0) return;
  const int fc = first_child[p];
  const double ri = 1.0 / cr[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < MS) ? LO + kChunk : MS;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.0;
    for (int c = fc; c < fc + k; ++c) {
      const double x = (cx[p] - cx[c]) * ri, y = (cy[p] - cy[c]) * ri, z = (cz[p] - cz[c]) * ri;
      if (c < ntop)
        fmmgen_ops::M2M_chunk<P, LO / kChunk, double>(x, y, z, HalfScaledIn64<OFF>{Mtop + (size_t)c * MS}, acc);
      else
        fmmgen_ops::M2M_chunk<P, LO / kChunk, double>(x, y, z, HalfScaledIn<float, OFF>{Mf + (size_t)c * MS}, acc);
    }
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) {
      Mtop[(size_t)p * MS + LO + j] = acc.v[j];
      Mf[(size_t)p * MS + LO + j] = (float)acc.v[j];
    }
  });
}

// FP32 L2L of the children of one level (parents below the top levels).
template <int P>
__global__ __launch_bounds__(64) void k_l2l_mixed(int beg, int end, const int* __restrict__ parent,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const double* __restrict__ cz, const double* __restrict__ cr,
                                                  float* __restrict__ Lf) {
  constexpr int LS = Op<P>::LS, NCH = (LS + kChunk - 1) / kChunk;
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  const int p = parent[c];
  const double ri = 1.0 / cr[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < LS) ? LO + kChunk : LS;
    ChunkAcc<float, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.f;
    fmmgen_ops::L2L_chunk<P, LO / kChunk, float>((float)((cx[c] - cx[p]) * ri), (float)((cy[c] - cy[p]) * ri),
                                (float)((cz[c] - cz[p]) * ri), static_cast<const float*>(Lf + (size_t)p * LS), acc);
    constexpr DegTab<512> tab{};
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) Lf[(size_t)c * LS + LO + j] += ldexpf(acc.v[j], -(tab.d[LO + j] + 1));
  });
}

// FP64 L2L from a top-level parent.  A top-level child accumulates in the
// FP64 array (seeded with its FP32 M2L result); a child below the top adds
// into the FP32 array.
template <int P>
__global__ __launch_bounds__(64) void k_l2l_top(int beg, int end, const int* __restrict__ parent,
                                                const double* __restrict__ cx, const double* __restrict__ cy,
                                                const double* __restrict__ cz, const double* __restrict__ cr,
                                                float* __restrict__ Lf, double* __restrict__ Ltop, int ntop) {
  constexpr int LS = Op<P>::LS, NCH = (LS + kChunk - 1) / kChunk;
  const int c = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= end) return;
  const int p = parent[c];
  const double ri = 1.0 / cr[p];
  dispatch_chunk<NCH>(blockIdx.y, [&](auto C) {
    constexpr int LO = C * kChunk, HI = (LO + kChunk < LS) ? LO + kChunk : LS;
    ChunkAcc<double, LO, HI> acc;
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) acc.v[j] = 0.0;
    fmmgen_ops::L2L_chunk<P, LO / kChunk, double>((cx[c] - cx[p]) * ri, (cy[c] - cy[p]) * ri, (cz[c] - cz[p]) * ri,
                                 static_cast<const double*>(Ltop + (size_t)p * LS), acc);
    constexpr DegTab<512> tab{};
#pragma unroll
    for (int j = 0; j < HI - LO; ++j) {
      const double v = ldexp(acc.v[j], -(tab.d[LO + j] + 1));
      if (c < ntop) {
        Ltop[(size_t)c * LS + LO + j] = (double)Lf[(size_t)c * LS + LO + j] + v;
        Lf[(size_t)c * LS + LO + j] = (float)Ltop[(size_t)c * LS + LO + j];
      } else {
        Lf[(size_t)c * LS + LO + j] += (float)v;
      }
    }
  });
}

// FP64 shifts of the top levels for charges, one wave per child (M2M: the
// children's shifts summed in LDS), lanes over output slots, from the
// closed forms of the generated bodies (multi-index layout, X[a] = x^a/a!):
//   M2M: Ms[n] += sum over k <= n (componentwise) of M[n - k] X Y Z[k],
//        x = (parent - child) / r_parent, M the child's multipole scaled
//        by 2^-deg (its normalisation at the parent's radius);
//   L2L: Ls[n] = sum over |n + k| <= P of L[n + k] X Y Z[k],
//        x = (child - parent) / r_parent, then scaled by 2^-(deg n + 1).
// Same results as k_m2m_top / k_l2l_top (to FP64 rounding), with the
// source expansion and the powers in LDS instead of one thread per 32-slot
// chunk running a long generated body.
template <int P>
__global__ __launch_bounds__(512) void k_m2m_top_wave(int beg, int end, const int* __restrict__ first_child,
                                                      const int* __restrict__ nchild, const double* __restrict__ cx,
                                                      const double* __restrict__ cy, const double* __restrict__ cz,
                                                      const double* __restrict__ cr, float* __restrict__ Mf,
                                                      double* __restrict__ Mtop, int ntop) {
  static_assert(SO == 0 || P < 0, "charges only");
  constexpr int MS = Op<P>::MS, NPL = (MS + 63) / 64;
  __shared__ double sM[8][MS];  // wave w: child w's scaled multipole, then its contribution
  __shared__ double sT[8][MS];  // the x- and z-pass results
  __shared__ double pw[8][3][P + 1];
  const int p = beg + blockIdx.x;
  if (p >= end) return;
  const int k = nchild[p];
  if (k == 0) return;  // uniform over the block
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, fc = first_child[p];
  const double ri = 1.0 / cr[p];
  constexpr DegTab<512> tab{};
  double acc[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) acc[i] = 0.0;
  if (w < k) {  // wave w shifts child fc + w
    const int c = fc + w;
    for (int j = lane; j < MS; j += 64)
      sM[w][j] = ldexp(c < ntop ? Mtop[(size_t)c * MS + j] : (double)Mf[(size_t)c * MS + j], -tab.d[j]);
    if (lane < 3 * (P + 1)) {
      const int d = lane / (P + 1), a = lane % (P + 1);
      const double u = d == 0 ? (cx[p] - cx[c]) * ri : (d == 1 ? (cy[p] - cy[c]) * ri : (cz[p] - cz[c]) * ri);
      double v = 1.0;
      for (int e = 1; e <= a; ++e) v = v * u / (double)e;
      pw[w][d][a] = v;
    }
  }
  // the shift M'_n = sum_{j <= n} M_j u^(n-j)/(n-j)! factors over the three
  // axes: one pass per axis, each a sum of at most P + 1 terms per slot
  MultiIdx mi[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) mi[i] = slot_multi_rt(min(lane + 64 * i, MS - 1));
  __syncthreads();
  auto pass = [&](const double* in, double* out, int d) {
    if (w < k) {
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int n = lane + 64 * i;
        if (n < MS) {
          const MultiIdx m = mi[i];
          const int e = d == 0 ? m.x : (d == 1 ? m.y : m.z);
          double sum = 0.0;
          for (int a = 0; a <= e; ++a)
            sum += pw[w][d][a] * in[multi_slot(m.x - (d == 0 ? a : 0), m.y - (d == 1 ? a : 0), m.z - (d == 2 ? a : 0))];
          acc[i] = sum;
        }
      }
    }
    __syncthreads();  // every wave done reading `in` (out may alias another pass's input)
    if (w < k) {
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int n = lane + 64 * i;
        if (n < MS) out[n] = acc[i];
      }
    }
    __syncthreads();
  };
  pass(sM[w], sT[w], 0);
  pass(sT[w], sM[w], 1);
  pass(sM[w], sT[w], 2);
  if (w < k) {
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int n = lane + 64 * i;
      if (n < MS) sM[w][n] = sT[w][n];
    }
  }
  __syncthreads();
  for (int n = threadIdx.x; n < MS; n += 512) {
    double v = 0.0;
    for (int c = 0; c < k; ++c) v += sM[c][n];
    Mtop[(size_t)p * MS + n] = v;
    Mf[(size_t)p * MS + n] = (float)v;
  }
}

template <int P>
__global__ __launch_bounds__(64) void k_l2l_top_wave(int beg, int end, const int* __restrict__ parent,
                                                     const double* __restrict__ cx, const double* __restrict__ cy,
                                                     const double* __restrict__ cz, const double* __restrict__ cr,
                                                     float* __restrict__ Lf, double* __restrict__ Ltop, int ntop) {
  static_assert(SO == 0 || P < 0, "charges only");
  constexpr int LS = Op<P>::LS, NPL = (LS + 63) / 64;
  __shared__ double sL[LS], sT[LS];
  __shared__ double pw[3][P + 1];
  const int c = beg + blockIdx.x;
  if (c >= end) return;
  const int lane = threadIdx.x, p = parent[c];
  const double ri = 1.0 / cr[p];
  for (int j = lane; j < LS; j += 64) sL[j] = Ltop[(size_t)p * LS + j];
  if (lane < 3 * (P + 1)) {
    const int d = lane / (P + 1), a = lane % (P + 1);
    const double u = d == 0 ? (cx[c] - cx[p]) * ri : (d == 1 ? (cy[c] - cy[p]) * ri : (cz[c] - cz[p]) * ri);
    double v = 1.0;
    for (int e = 1; e <= a; ++e) v = v * u / (double)e;
    pw[d][a] = v;
  }
  // the shift L'_n = sum_{j >= n, |j| <= P} L_j u^(j-n)/(j-n)! factors over
  // the three axes (each pass keeps |j| <= P): one pass per axis, at most
  // P + 1 terms per slot
  constexpr DegTab<512> tab{};
  MultiIdx mi[NPL];
  double acc[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) mi[i] = slot_multi_rt(min(lane + 64 * i, LS - 1));
  __syncthreads();
  auto pass = [&](const double* in, double* out, int d) {
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int n = lane + 64 * i;
      if (n < LS) {
        const MultiIdx m = mi[i];
        const int R = P - (m.x + m.y + m.z);
        double sum = 0.0;
        for (int a = 0; a <= R; ++a)
          sum += pw[d][a] * in[multi_slot(m.x + (d == 0 ? a : 0), m.y + (d == 1 ? a : 0), m.z + (d == 2 ? a : 0))];
        acc[i] = sum;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int n = lane + 64 * i;
      if (n < LS) out[n] = acc[i];
    }
    __syncthreads();
  };
  pass(sL, sT, 0);
  pass(sT, sL, 1);
  pass(sL, sT, 2);
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int n = lane + 64 * i;
    if (n < LS) {
      const double v = ldexp(acc[i], -(tab.d[n] + 1));
      const size_t o = (size_t)c * LS + n;
      if (c < ntop) {
        Ltop[o] = (double)Lf[o] + v;
        Lf[o] = (float)Ltop[o];
      } else {
        Lf[o] += (float)v;
      }
    }
  }
}

static __global__ void k_seed_root(int LS, const float* __restrict__ Lf, double* __restrict__ Ltop) {
  for (int j = threadIdx.x; j < LS; j += blockDim.x) Ltop[j] = (double)Lf[j];
}

// FP32 L2P on the normalised local expansion, plus the FP64 near field.
template <int P>
__global__ __launch_bounds__(64) void k_l2p_mixed(int n, const int* __restrict__ leaf_of, const int* __restrict__ perm,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const double* __restrict__ cz, const double* __restrict__ cr,
                                                  const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, const float* __restrict__ Lf,
                                                  const double* __restrict__ Fnear, int use_l2p,
                                                  double* __restrict__ F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double out[OUT];
#pragma unroll
  for (int k = 0; k < OUT; ++k) out[k] = Fnear ? Fnear[(size_t)j * OUT + k] : 0.0;
  if (use_l2p) {
    const int c = leaf_of[j];
    const double ri = 1.0 / cr[c];
    float acc[OUT];
#pragma unroll
    for (int k = 0; k < OUT; ++k) acc[k] = 0.f;
    fmmgen_ops::L2P_o<P, float>((float)((px[j] - cx[c]) * ri), (float)((py[j] - cy[c]) * ri),
                                (float)((pz[j] - cz[c]) * ri), static_cast<const float*>(Lf + (size_t)c * Op<P>::LS),
                                acc);
    // potential scales with 1/r, field with 1/r^2
    constexpr int F0 = OUT == 3 ? 0 : 1;
    if (OUT != 3) out[0] += (double)acc[0] * ri;
#pragma unroll
    for (int k = F0; k < OUT; ++k) out[k] += (double)acc[k] * ri * ri;
  }
  const int i = perm[j];
#pragma unroll
  for (int k = 0; k < OUT; ++k) F[(size_t)i * OUT + k] = out[k];
}

template <int P>
void m2l_mixed(Tree& t, int mode, hipStream_t st);

template <int P>
void far_field_mixed(Tree& t, int mode, int mask, hipStream_t st) {
  const int nc = t.ncells;
  constexpr int MS = Op<P>::MS, LS = Op<P>::LS;
  const bool top64 = P >= 8;
  const int ntop = top64 ? t.level_off[std::min(kTopLevels, t.depth)] : 0;
  t.Mf.alloc((size_t)nc * MS);
  t.Lf.alloc((size_t)nc * LS);
  t.Mtop.alloc((size_t)std::max(ntop, 1) * MS);
  t.Ltop.alloc((size_t)std::max(ntop, 1) * LS);
  // P2M writes every slot of every leaf and M2M of every parent; the
  // harmonic form's unfold stores every slot of every local expansion.  The
  // clears are needed only when a phase is masked out, for the full form
  // (its M2L adds into Lf) and when there are no M2L groups.
  if (!((mask & kPhaseP2M) && (mask & kPhaseM2M)))
    FMM_HIP(hipMemsetAsync(t.Mf.get(), 0, sizeof(float) * nc * MS, st));
  bool unfold_all = false;
  if constexpr (SO == 0) unfold_all = (mask & kPhaseM2L) && t.sched.ngroups > 0 && m2l_harmonic(t, mode);
  if (!unfold_all) FMM_HIP(hipMemsetAsync(t.Lf.get(), 0, sizeof(float) * nc * LS, st));
  if (mask & kPhaseP2M) {
    PhaseTimer pt(t, kTP2M, st);
    if constexpr (SO == 0)
      k_p2m_wave<P><<<t.nleaves, 64, 0, st>>>(t.nleaves, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.cx.get(),
                                              t.cy.get(), t.cz.get(), t.cr.get(), t.px.get(), t.py.get(),
                                              t.pz.get(), t.pS.get(), t.Mf.get(), t.Mtop.get(), ntop);
    else
      k_p2m_mixed<P><<<dim3(nblk(t.nleaves, 64), (MS + kChunk - 1) / kChunk), 64, 0, st>>>(
          t.nleaves, t.leaves.get(), t.pstart.get(), t.pcount.get(), t.cx.get(), t.cy.get(), t.cz.get(),
          t.cr.get(), t.px.get(), t.py.get(), t.pz.get(), t.pS.get(), t.Mf.get(), t.Mtop.get(), ntop);
  }
  if (mask & kPhaseM2M) {
    PhaseTimer pt(t, kTM2M, st);
    // levels whose children number at most kM2MChildMax run in two passes
    // (k_m2m_child, k_m2m_sum), larger ones with one thread per parent;
    // Mpart is sized once for the largest such child level, before any launch
    int maxc = 1;
    for (int l = t.depth - 2; l >= 0; --l) {
      const int nch = t.level_off[l + 2] - t.level_off[l + 1];
      if (!(top64 && l < kTopLevels) && nch <= kM2MChildMax) maxc = std::max(maxc, nch);
    }
    t.Mpart.alloc((size_t)maxc * MS);
    for (int l = t.depth - 2; l >= 0; --l) {
      const int b = t.level_off[l], e = t.level_off[l + 1];
      const dim3 g(nblk(e - b, 64), (MS + kChunk - 1) / kChunk);
      if (top64 && l < kTopLevels) {
        if constexpr (SO == 0)
          k_m2m_top_wave<P><<<e - b, 512, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cx.get(),
                                                  t.cy.get(), t.cz.get(), t.cr.get(), t.Mf.get(), t.Mtop.get(),
                                                  ntop);
        else
          k_m2m_top<P><<<g, 64, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cx.get(), t.cy.get(),
                                         t.cz.get(), t.cr.get(), t.Mf.get(), t.Mtop.get(), ntop);
      } else if (t.level_off[l + 2] - e > kM2MChildMax) {
        k_m2m_mixed<P><<<g, 64, 0, st>>>(b, e, t.first_child.get(), t.nchild.get(), t.cx.get(), t.cy.get(),
                                         t.cz.get(), t.cr.get(), t.Mf.get());
      } else {
        const int cb = e, ce = t.level_off[l + 2];  // the level's children: the next level
        k_m2m_child<P><<<dim3(nblk(ce - cb, 64), (MS + kChunk - 1) / kChunk), 64, 0, st>>>(
            cb, ce, t.parent.get(), t.cx.get(), t.cy.get(), t.cz.get(), t.cr.get(), t.Mf.get(), t.Mpart.get());
        k_m2m_sum<P><<<nblk((long long)(e - b) * MS, 256), 256, 0, st>>>(b, e, cb, t.first_child.get(),
                                                                         t.nchild.get(), t.Mpart.get(), t.Mf.get());
      }
    }
  }
  if (mask & kPhaseM2L) {
    PhaseTimer pt(t, kTM2L, st);
    m2l_mixed<P>(t, mode, st);
  }
  if (mask & kPhaseL2L) {
    PhaseTimer pt(t, kTL2L, st);
    if (top64) k_seed_root<<<1, 64, 0, st>>>(LS, t.Lf.get(), t.Ltop.get());
    for (int l = 1; l < t.depth; ++l) {
      const int b = t.level_off[l], e = t.level_off[l + 1];
      const dim3 g(nblk(e - b, 64), (LS + kChunk - 1) / kChunk);
      if (top64 && l - 1 < kTopLevels) {
        if constexpr (SO == 0)
          k_l2l_top_wave<P><<<e - b, 64, 0, st>>>(b, e, t.parent.get(), t.cx.get(), t.cy.get(), t.cz.get(),
                                                  t.cr.get(), t.Lf.get(), t.Ltop.get(), ntop);
        else
          k_l2l_top<P><<<g, 64, 0, st>>>(b, e, t.parent.get(), t.cx.get(), t.cy.get(), t.cz.get(), t.cr.get(),
                                         t.Lf.get(), t.Ltop.get(), ntop);
      } else
        k_l2l_mixed<P><<<g, 64, 0, st>>>(b, e, t.parent.get(), t.cx.get(), t.cy.get(), t.cz.get(), t.cr.get(),
                                         t.Lf.get());
    }
  }
}

template <int P>
void l2p_mixed(Tree& t, int mask, double* F, hipStream_t st) {
  const int n = (int)t.n;
  PhaseTimer pt(t, kTL2P, st);
  k_l2p_mixed<P><<<nblk(n, 64), 64, 0, st>>>(n, t.leaf_of.get(), t.perm.get(), t.cx.get(), t.cy.get(), t.cz.get(),
                                             t.cr.get(), t.px.get(), t.py.get(), t.pz.get(), t.Lf.get(),
                                             (mask & kPhaseP2P) ? t.Fp2p.get() : nullptr, (mask & kPhaseL2P) ? 1 : 0,
                                             F);
}

}  // namespace FMM_VARIANT_NS
}  // namespace fmm
