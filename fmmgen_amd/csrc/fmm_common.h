// Shared definitions of the MI355X FMM library: error plumbing, the tree
// handle, and small device helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

namespace fmm {

// ------------------------------------------------------------ status codes
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kUnsupported = 2,
  kHipError = 3,
  kTreeTooDeep = 4,
  kOutOfMemory = 5,
};

void set_error(const std::string& msg);
const char* last_error();

struct Error {
  int code;
  std::string msg;
};

#define FMM_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::fmm::Error{::fmm::kHipError, std::string(#expr) + ": " + hipGetErrorString(e_) + \
                                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
  } while (0)

#define FMM_CHECK(cond, code, msg)                  \
  do {                                              \
    if (!(cond)) throw ::fmm::Error{(code), (msg)}; \
  } while (0)

// ------------------------------------------------------------ index layout
__host__ __device__ constexpr int nterms(int p) { return p < 0 ? 0 : (p + 1) * (p + 2) * (p + 3) / 6; }
__host__ __device__ constexpr int msize(int p, int so) { return so == 0 ? nterms(p) : nterms(p) - nterms(so - 1); }
__host__ __device__ constexpr int lsize(int p, int so) { return nterms(p - so); }
__host__ __device__ constexpr int sourcesize(int so) { return nterms(so) - nterms(so - 1); }

// Device buffer owned by the library (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    FMM_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  T* get() const { return p; }
};

// Scratch storage of device primitives (rocPRIM temp storage etc.).
struct Scratch {
  DevBuf<unsigned char> buf;
  void* get(size_t bytes) {
    buf.alloc(bytes);
    return buf.get();
  }
};

// ------------------------------------------------------------ phase timers
enum PhaseTimerId : int {
  kTPrep = 0,   // source gather and zeroing
  kTP2M = 1,
  kTM2M = 2,
  kTM2L = 3,
  kTP2P = 4,
  kTL2L = 5,
  kTL2P = 6,    // L2P plus the final per-target sum
  kTOrder = 7,  // learned order: features, network inference, cell and tile orders
  kTJudge = 8,  // learned order: sampled direct-sum judge and reruns
  kTBH = 9,     // Barnes-Hut walk (M2P and P2P per particle)
  kNumPhaseTimers = 10,
};

// ------------------------------------------------------------ M2L schedule
// Pairs of one target tile are grouped by (source level, quantised centre
// offset); every group shares one translation matrix.
constexpr int kTile = 16;        // target cells per M2L tile (one MFMA column block)

// Operand form of the M2L GEMM: the full graded expansions (the generated
// operator's sums), or the harmonic layout (folded multipoles, local rows
// with n_z <= 1, the same sums by the harmonicity of 1/R); automatic = the
// harmonic layout for the FP32 and BF16x3 GEMMs, the full form for FP64.
enum : int { kM2LAuto = 0, kM2LFull = 1, kM2LHarmonic = 2 };

struct M2LSchedule {
  static constexpr int tile_size = kTile;
  int ntiles = 0;
  int ngroups = 0;               // (tile, offset) groups
  int nkeys = 0;                 // distinct (target level, source level, offset) keys
  DevBuf<int> tile_cells;        // [ntiles * kTile] target cell or -1
  DevBuf<int> tile_group_off;    // [ntiles + 1] range in the group list
  DevBuf<int> group_key;         // [ngroups] key id
  DevBuf<int> group_src;         // [ngroups * kTile] source cell per tile column or -1
  DevBuf<double> key_dx;         // [nkeys * 3] centre offset target - source of one pair of the key
  DevBuf<double> key_r;          // [nkeys] target half-width (scale of the normalised operator)
  DevBuf<int> key_dlevel;        // [nkeys] target level - source level
  // per evaluation: the D vector of every key (FP64 plain, normalised FP32 mixed)
  DevBuf<double> Dtab64;
  DevBuf<float> Dtab32;
  // (step, lane) -> D slot tables of the GEMM plans, uploaded once per
  // order and block size (4 or 8 waves)
  static constexpr int kMaxOrder = 16;
  DevBuf<unsigned> idxtab[2][kMaxOrder][3];  // [layout][order][4, 8 or 2 waves]
  DevBuf<unsigned> idxtab32[2][kMaxOrder];  // [layout][order]
  bool idx_ready[2][kMaxOrder][3] = {}, idx32_ready[2][kMaxOrder] = {};
  // learned order: chunks sorted by their tile's order, and the start of
  // every order's range (kMaxOrder + 1 entries)
  DevBuf<int> class_chunks;
  DevBuf<int> class_start;               // [kMaxOrder + 1] first chunk of every order (-1: none)
  DevBuf<unsigned long long> class_keys; // sort keys of the order classes
  DevBuf<long long> tile_pairs;  // M2L pairs per tile (statistics)
  bool tile_pairs_ready = false;
  DevBuf<int> cell_tile;         // cell -> its M2L tile (-1: none)
  bool cell_tile_ready = false;
  // work items: chunks of at most kGroupsPerChunk groups of one tile
  int nchunks = 0;
  DevBuf<int> chunk_tile, chunk_g0, chunk_g1;
};
constexpr int kGroupsPerChunk = 96;  // measured: 48 -> 96 cut the learned M2L by 11%

// ------------------------------------------------------------ tree handle
struct Tree {
  // configuration (example/tree.hpp:78-95 + generated macros)
  int order = 0, so = 0, outsz = 4;
  size_t ncrit = 0;
  double theta = 0;
  size_t n = 0;

  // geometry, canonical numbering: level-major, octant-path order inside a
  // level; children of a cell are contiguous and follow their parent.
  int depth = 0;                  // number of levels
  int ncells = 0, nleaves = 0;
  std::vector<int> level_off;     // host copy, size depth + 1
  DevBuf<double> cx, cy, cz, cr, crmax;
  DevBuf<int> parent, first_child, nchild, octmask, level, pstart, pcount;
  DevBuf<int> octant;             // octant of the cell in its parent (0 for the root)
  DevBuf<int> leaves;             // leaf cell ids in canonical order

  // particles in leaf order (particles of every cell are contiguous)
  DevBuf<int> perm;               // sorted slot -> caller index
  DevBuf<int> leaf_of;            // sorted slot -> leaf cell
  DevBuf<double> px, py, pz;      // sorted positions
  DevBuf<double> pS;              // sorted source strengths [n * SOURCESIZE], per evaluation
  const double* user_S = nullptr; // borrowed, [n * SOURCESIZE]

  // interaction lists (example/tree.hpp:87-88), multiset-equal to the
  // reference; CSR by target cell.
  long long n_m2l = 0, n_p2p = 0;
  DevBuf<int> m2l_off, m2l_src;   // [ncells+1], [n_m2l]
  DevBuf<int> p2p_off, p2p_src;   // [ncells+1], [n_p2p]
  // P2P sources of each target leaf as merged particle ranges (the target
  // leaf itself excluded), sorted by start: CSR over cells
  long long n_rng = 0;
  DevBuf<int> rng_off, rng_start, rng_len;
  // P2P work items: the source sequence of a target leaf (own particles,
  // then its merged ranges) cut into segments of at most kP2PSegment
  // source particles, so that a leaf next to a dense cluster is spread
  // over many waves.  Item = (leaf, first range (rb - 1 = own leaf),
  // offset in that range, source count); split leaves accumulate with
  // atomics into a zeroed output.
  static constexpr int kP2PSegment = 8192;
  long long n_p2p_items = 0;
  int p2p_split = 0;  // some leaf has more than one item
  DevBuf<int> p2p_item_leaf, p2p_item_r, p2p_item_o, p2p_item_n;
  M2LSchedule sched;

  // learned order: per leaf features and order, per cell order (the
  // largest order of the leaves below), per M2L tile order (the largest
  // order of its 16 target cells; the tile's groups are truncated there)
  DevBuf<float> leaf_feat;         // [nleaves * kMlpF]
  DevBuf<int> leaf_q, cell_q;      // [nleaves], [ncells]
  DevBuf<int> leaf_q2, cell_q2a;   // the second allocation's orders
  DevBuf<signed char> tile_order2;
  DevBuf<double> lw;               // device scalars of the order selection (learned.h)
  DevBuf<double> lam_sums;         // multiplier search sums
  DevBuf<int> cell_eff;            // [ncells] effective order (lowest tile order on the chain)
  DevBuf<int> tile_hint;           // [ncells] order hint of the M2L tiling (set with a model)
  bool use_tile_hint = false;
  const int* tile_hint_ptr() const { return use_tile_hint ? tile_hint.get() : nullptr; }
  DevBuf<float> leaf_y;            // [nleaves * kMlpO] network outputs
  DevBuf<double> cell_qabs, cell_q2;  // [ncells] sum of |S| and of |S|^2 below the cell
  DevBuf<double> cell_geo4;           // [ncells * 4] centre and rmax, packed for the feature gather
  DevBuf<float> cell_qw;              // [ncells * 2] float(sum |S|), sqrt(float(sum S^2))
  DevBuf<double> cell_acc;            // [ncells * 11] list sums, accumulated down the tree

  DevBuf<signed char> tile_order;  // [ntiles]
  bool use_tile_order = false;
  const signed char* tile_order_ptr() const { return use_tile_order ? tile_order.get() : nullptr; }
  // M2L by tile-order classes (set by the learned evaluation): called by
  // the far field of order P after the D table is built at order P
  void (*m2l_classes)(Tree&, int mode, int P, hipStream_t) = nullptr;
  // second stream of the per-order M2L class launches (default priority,
  // created on first use) and its fork / join events
  hipStream_t side = nullptr;
  hipEvent_t side_fork = nullptr, side_join = nullptr;
  DevBuf<_Float16> mlp_half;       // network parameters in FP16
  bool has_model = false;
  double tol = 1e-6, margin = 0.0;
  // sampled judge: targets (sorted slots), their FP64 direct sums, results
  int judge_n = 0;
  DevBuf<int> judge_slot;
  DevBuf<double> judge_Fd, judge_acc;  // [judge_n * OUT], [4]
  DevBuf<int> judge_fail;              // [1 + ncells] failing count, per cell flag
  double lstats[11] = {0};
  bool lstats_ready = false;       // lstats[5..9] computed for the last learned evaluation
  int class_start_h[M2LSchedule::kMaxOrder + 1] = {0};  // host copy: first work item of every supertile order

  // expansions, [ncells * Msize] and [ncells * Lsize] in FP64; ml_order =
  // the order of the FP64 evaluation that filled them (-1: the last
  // evaluation was mixed or learned)
  DevBuf<double> M, L;
  int ml_order = -1;
  // mixed precision: normalised FP32 expansions, FP64 copies of the top levels
  DevBuf<float> Mf, Lf;
  // FP32 M2M: each child's multipole translated to its parent's centre, one
  // level's children at a time ([cells of the largest level * Msize])
  DevBuf<float> Mpart;
  // M2L operand form (kM2LAuto / kM2LFull / kM2LHarmonic) and the folded
  // multipoles and harmonic local rows of a harmonic-layout M2L
  int m2l_form = 0;
  DevBuf<float> Mhf, Lhf;
  DevBuf<double> Mh64, Lh64;
  DevBuf<double> Mtop, Ltop;
  DevBuf<double> Fp2p;            // per-target near-field result, sorted order
  Scratch scratch;

  // work counters (list build)
  long long p2p_particle_pairs = 0;  // sum over P2P pairs of nA*nB minus self pairs

  // stream-ordered phase timing (hipEvents recorded around each phase)
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;
  std::vector<hipEvent_t> event_pool;

  DevBuf<int> p2p_next;            // work counters of the persistent near-field kernel
  int num_cu = 0;                  // compute units of the device (queried once)
  double phase_ms[kNumPhaseTimers] = {0};
  long long phase_count[kNumPhaseTimers] = {0};
};

// Correctly rounded FP64 square root: the hardware sequence, then the
// neighbour with the smallest exact residual s - c*c (one fma each).
__device__ __forceinline__ double sqrt_rn(double s) {
  const double r = __builtin_sqrt(s);
  if (!(s > 0.0) || !isfinite(s) || !(r > 0.0)) return r;
  // neighbours of a positive finite double are one bit pattern away
  const long long bits = __double_as_longlong(r);
  const double lo = __longlong_as_double(bits - 1);
  const double hi = __longlong_as_double(bits + 1);
  const double er = fabs(__builtin_fma(-r, r, s));
  const double el = fabs(__builtin_fma(-lo, lo, s));
  const double eh = fabs(__builtin_fma(-hi, hi, s));
  double best = r, be = er;
  if (el < be) { best = lo; be = el; }
  if (eh < be) { best = hi; }
  return best;
}

// |(dx, dy, dz)| as the reference's host code computes it:
// sqrt(dx*dx + dy*dy + dz*dz) with every operation rounded on its own (no
// contraction) and a correctly rounded square root, so that an opening
// criterion on it decides as the reference does.
__device__ inline double norm3_rn(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return sqrt_rn(dx * dx + dy * dy + dz * dz);
}

// Blocks of the persistent near-field kernel: kP2PWavesPerSimd one-wave
// blocks per SIMD of every compute unit.
constexpr int kP2PWavesPerSimd = 8;
inline int p2p_grid(Tree& t) {
  if (t.num_cu == 0) {
    int dev = 0;
    FMM_HIP(hipGetDevice(&dev));
    FMM_HIP(hipDeviceGetAttribute(&t.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (t.num_cu <= 0) t.num_cu = 256;
  }
  return t.num_cu * 4 * kP2PWavesPerSimd;
}

// RAII helper: records a start/stop event pair for one phase when timing
// is on.
struct PhaseTimer {
  Tree& t;
  int phase;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  PhaseTimer(Tree& tree, int ph, hipStream_t s);
  ~PhaseTimer();
};
void timing_collect(Tree& t);

}  // namespace fmm
