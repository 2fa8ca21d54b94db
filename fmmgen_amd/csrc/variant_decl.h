// Declarations shared by the base and per-order translation units of one
// operator variant.  Included after the generated device operator header
// (namespace fmmgen_ops, FMMGEN_* macros) with FMM_VARIANT_NS defined.
#pragma once

#include "evaluate.h"
#include "fmm_common.h"

namespace fmm {
namespace FMM_VARIANT_NS {

constexpr int SO = FMMGEN_SOURCEORDER;
constexpr int SS = FMMGEN_SOURCESIZE;
constexpr int OUT = FMMGEN_OUTPUTSIZE;
constexpr int PMIN = FMMGEN_MINORDER;
constexpr int PMAX = FMMGEN_MAXORDER;  // exclusive

inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

template <int P>
struct Op {
  static constexpr int MS = msize(P, SO);
  static constexpr int LS = lsize(P, SO);
};

// Per-order phases (defined in the per-order translation units).
template <int P>
void far_field_order(Tree& t, int mode, int mask, hipStream_t st);
template <int P>
void l2p_final_order(Tree& t, int mode, int mask, double* F, hipStream_t st);
template <int P>
void op_batch_order(int op, int n, const double* xyz, const double* in, int in_stride, double* out,
                    int out_stride, hipStream_t st);
template <int P>
void bh_order(Tree& t, double* F, hipStream_t st);
// M2L of the listed chunks at order Q on expansions stored at order ps.
template <int Q>
void m2l_class(Tree& t, int mode, const int* chunks, int nch, int ps, hipStream_t st);
// Whether the M2L of an evaluation in this mode runs in the harmonic layout.
inline bool m2l_harmonic(const Tree& t, int mode) {
  if (SO != 0 || t.m2l_form == kM2LFull) return false;
  return t.m2l_form == kM2LHarmonic || mode != 0;  // automatic: the mixed and learned modes
}

// Order-independent phases (defined in the base translation unit).
void near_field(Tree& t, int mode, hipStream_t st);

}  // namespace FMM_VARIANT_NS
}  // namespace fmm
