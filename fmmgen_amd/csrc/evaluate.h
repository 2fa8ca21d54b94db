// Interfaces between the C ABI and the per-variant evaluation code.
#pragma once
#include "fmm_common.h"

namespace fmm {

// Phase bits of an evaluation (all set: the whole method).
enum Phase : int {
  kPhaseP2M = 1,
  kPhaseM2M = 2,
  kPhaseM2L = 4,
  kPhaseP2P = 8,
  kPhaseL2L = 16,
  kPhaseL2P = 32,
  kPhaseAll = 63,
};

// Operator ids of the batch test entry point.
enum OpId : int {
  kOpP2M = 0,
  kOpM2M = 1,
  kOpM2L = 2,
  kOpL2L = 3,
  kOpL2P = 4,
  kOpM2P = 5,
  kOpP2P = 6,
  kOpP2MS = 7,
};

// One compiled operator variant: a source order and an output layout.
struct Variant {
  int so, outsz, min_order, max_order;  // max_order exclusive
  void (*evaluate)(Tree&, int mode, int mask, double* F, hipStream_t);
  void (*evaluate_learned)(Tree&, int mode, double* F, hipStream_t);
  void (*field_scale)(Tree&, hipStream_t);
  void (*predict_cell_orders)(Tree&, hipStream_t);
  void (*gather_sources)(Tree&, hipStream_t);
  void (*op_batch)(int op, int order, int n, const double* xyz, const double* in, int in_stride, double* out,
                   int out_stride, hipStream_t);
  void (*direct)(int n, const double* pos, const double* S, const long long* targets, int nt, double* F,
                 hipStream_t);
  void (*evaluate_bh)(Tree&, double* F, hipStream_t);
};

extern const Variant variant_so0_pf;  // charges, potential + field
extern const Variant variant_so0_p;   // charges, potential only
extern const Variant variant_so0_f;   // charges, field only
extern const Variant variant_so1_f;   // dipoles, field only
extern const Variant variant_so2_pf;  // quadrupoles, potential + field

const Variant* find_variant(int so, int outsz);

void build_tree_device(Tree& t, const double* d_pos, hipStream_t st);
void build_lists_device(Tree& t, hipStream_t st);
void build_m2l_schedule(Tree& t, hipStream_t st);

}  // namespace fmm
