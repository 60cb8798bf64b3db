// fmm_internal.h -- shared host/device declarations of the MI355X FMM library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "exafmm_mi355x.h"

namespace xf {

constexpr int kMaxLevel = 21;     // 63-bit octant keys
// forest of independent problems in one tree: levels 0 .. kForestLevel - 1
// hold the problem index (up to 8^kForestLevel problems), the problems'
// roots are the cells of level kForestLevel
constexpr int kForestLevel = 3;
constexpr int kForestMax = 1 << (3 * kForestLevel);
// one problem of a gathered forest: its own device array (null: host
// memory, copied by the runtime) and its first body in the forest's array
struct ForestPart {
  exafmm_body *src;
  int64_t off;
};
constexpr int kMaxP = 16;         // expansion order cap (NTERM <= 136)
constexpr int kWave = 64;

// ------------------------------------------------------------------ errors
struct Failure {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);

#define XF_HIP(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      ::xf::fail(EXAFMM_ERR_HIP, std::string(#expr) + " -> " + hipGetErrorString(e_));     \
  } while (0)

// Host wait for a stream by polling it.  A blocking wait sleeps and the
// host wakes tens of microseconds after the stream drains; an evaluation
// waits mid-pass a few times (tree level sizes, list totals, tile plan),
// and each such wake-up is time the GPU sits idle.
inline void stream_wait(hipStream_t s) {
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) XF_HIP(e);
  }
}

#define XF_LAUNCH()                                                                        \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess)                                                                  \
      ::xf::fail(EXAFMM_ERR_HIP, std::string("kernel launch at ") + __FILE__ + ":" +       \
                                     std::to_string(__LINE__) + " -> " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ device buffer
// Grow-only device allocation: kept across evaluations so the timed path
// performs no allocation once the sizes have been seen.
template <class T>
struct Buf {
  T *p = nullptr;
  size_t cap = 0;
  void ensure(size_t count) {
    if (count <= cap && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      fail(EXAFMM_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    cap = count ? count : 1;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  ~Buf() { release(); }
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
};

// Pinned host memory for the readbacks of a pass.  Device-to-host copies
// into it are asynchronous, so one wait covers every copy of a read point
// (a pageable destination turns each copy into a blocking round trip).
// Resized only while no copy into it is pending.
template <class T>
struct HostBuf {
  T *p = nullptr;
  size_t cap = 0;
  void ensure(size_t count) {
    if (count <= cap && p) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipHostMalloc(reinterpret_cast<void **>(&p), bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      fail(EXAFMM_ERR_ALLOC, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    cap = count ? count : 1;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
  ~HostBuf() { release(); }
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
};

// Slots of Ctx::pin_i (ints read back by the tree build, traversal and M2L plan)
constexpr int kPinLevelMax = 0;   // 9: flags, largest prefix population of levels 1..8
constexpr int kPinLevels = 16;    // kMaxLevel + 2 level sizes
constexpr int kPinMisc = 48;      // 19: tree flags and counters
constexpr int kPinTot = 72;       // 8: list totals and flags
constexpr int kPinPlan = 80;      // 3: M2L slots, tiles, cells in tiles
constexpr int kPinInts = 96;
// traversal flag bits that make the pass repeat: list capacity, key table full
constexpr int kListRetry = 1 | 8;
static_assert(kPinLevels + kMaxLevel + 2 <= kPinMisc && kPinMisc + 19 <= kPinTot, "pinned slots overlap");

inline int ceil_div(int64_t a, int64_t b) { return int((a + b - 1) / b); }

// ------------------------------------------------------------------ grouped launches
// One launch per stage over many independent problems (a batch): block b
// of a grouped launch belongs to the problem i with start[i] <= b <
// start[i + 1] and runs that problem's block b - start[i].
__device__ inline int group_of(const int *__restrict__ start, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The device arrays of one problem that the charge-dependent passes read
// and write (its own tree is source and target).
struct GroupTree {
  const int *leaves, *nleaves, *cbeg, *ccnt, *cchild, *cnch;
  const double4 *cx, *xs;
  const float4 *xq;
  double4 *xs_w;
  float4 *xq_w;
  const int *lbeg, *lcnt, *list;  // P2P lists
  const double *shift_T;
  double *M, *L;
  exafmm_body *bodies;
  int64_t n, ncells;
  int P;
};

// One problem's share of a per-level grouped launch (M2M or L2L): the
// cells of level l of problem `prob` form nbx x nby blocks.
struct GroupLevel {
  int prob, lstart, lcount, nbx, nby;
};

// ------------------------------------------------------------------ context
// Device-side tree in breadth-first (level-major) order; children of a
// cell are contiguous and in octant order, so each level is a Morton-
// ordered slice.  The reference (DFS-block) numbering is produced on
// export only.
struct Tree {
  int64_t n = 0;
  int depth = 0;                 // deepest level index
  int64_t ncells = 0, nleaves = 0;
  std::vector<int> lstart;       // host copy: level start, size depth+2
  double X0[3] = {0, 0, 0}, R0 = 0;
  int lroot = 0;                 // level of the (problem) roots: kForestLevel in a forest
  int ds_hint = -1;              // sort depth of the last build (-1: none yet)
  int hist7_max = 0;             // largest level-7 prefix population of the last build

  Buf<double4> cx;     // centre xyz, radius R
  Buf<int> cbeg, ccnt; // body range in tree order
  Buf<int> cchild, cnch, cparent, clevel;
  Buf<int> leaves;     // leaf cells, BFS order
  Buf<int> perm;       // tree-order position -> caller's body index
  Buf<double4> xs;     // tree-order X, q (FP64)
  Buf<float4> xq;      // tree-order X - leaf centre (FP32), q
  Buf<int> body_leaf;  // tree-order position -> leaf cell

  // build scratch
  Buf<unsigned long long> key, key2;
  Buf<int> val, val2;
  Buf<int> hist;       // dense prefix histograms, levels 1..7
  Buf<int> hist_bmax;  // per-block largest populations of levels 3..7
  Buf<int> lcount, lstart_d, tmp_nch, tmp_off, tmp_bnd, misc;
  Buf<double> bounds_part;
  Buf<int> rs_hist;
  Buf<int> scan_tmp;  // device-wide scan partials
  bool built = false;
};

struct Lists {
  bool valid = false;
  Buf<int> m2l_beg, m2l_cnt;  // per target cell
  Buf<int2> m2l;              // (source cell, key slot)
  Buf<int> p2p_beg, p2p_cnt;  // per target cell
  Buf<int> p2p;               // source cell
  Buf<int> out_beg, out_cnt;  // traversal hand-down lists
  Buf<int> out;
  Buf<int> tot;               // device totals [m2l, p2p, out], overflow flags
  int64_t cap_m2l = 0, cap_p2p = 0, cap_out = 0;
  int64_t nm2l = 0, np2p = 0;
  // translation keys
  Buf<unsigned long long> hslots;  // open-addressing key set
  int hcap = 0;
  Buf<int> slot2id;
  Buf<int> scan_tmp;       // device-wide scan partials
  Buf<int64_t> capd;       // list capacities on the device (scan overflow check)
  HostBuf<int64_t> caps_pin;
  int64_t caps_up[3] = {-1, -1, -1};  // capacities last uploaded
  Buf<unsigned long long> keys;    // distinct keys, ascending
  Buf<int> kscratch;
  int nkeys = 0;
};

struct Ctx {
  exafmm_config cfg{};
  int P = 10, NT = 55;
  hipStream_t s0 = nullptr, s1 = nullptr;
  hipEvent_t ev[16] = {};
  std::string last_error;
  HostBuf<int> pin_i;                   // kPin* slots
  HostBuf<double> pin_d;                // root cube
  HostBuf<unsigned long long> pin_keys; // M2L key set

  exafmm_body *bodies = nullptr;   // caller's array (device pointer or our mirror)
  exafmm_body *host_bodies = nullptr;
  bool host_mode = false;
  // tree_build leaves the write-back of X, q in tree order to the caller
  // (writeback_bodies), who runs it off the latency chain
  bool defer_writeback = false, writeback_pending = false;
  Buf<exafmm_body> mirror;         // device copy when the caller passed host memory

  Tree t;
  Lists lists;
  const Ctx *src = nullptr;        // source tree of the last horizontal pass
  // forest evaluation (exafmm_evaluate_forest): forest_n problems, bodies
  // forest_off[p] .. forest_off[p+1] - 1, their roots; 0 = one problem
  int forest_n = 0;
  Buf<int64_t> forest_off;
  Buf<double4> forest_roots;
  // expansions, radius-scaled real form [ncells][P^2]:
  //   M~ = (Re, Im) M_n^m / R^n,  L~ = (Re, Im) L_j^k R^(j+1)
  Buf<double> M, L;
  bool have_M = false, have_L = false;
  // graph recording (api.hip, exafmm_graph_*): while `capturing`, launches
  // are recorded into a graph (no events, no host waits, no allocation);
  // while a graph holds the context its tree, lists and tiles stay fixed
  bool capturing = false;
  int graph_locks = 0;
  Buf<double> shift_T;             // scaled M2M / L2L operators per octant
  Buf<double> dipole_part;         // dipole moment: per-workgroup partials, then the sum
  int shift_P = -1;

  // M2L fast path
  Buf<float> ops32;                // [nkeys][P^2][P^2] scaled real-form operators
  // What ops32 holds: operators are pure functions of (key, P, truncation,
  // precision) plus R0 and the cycle for periodic keys, so an unchanged key
  // set is not generated again.
  std::vector<unsigned long long> ops_keys;
  // keyed M2L tiles built by m2l_prepare for the current lists
  bool m2l_ready = false;
  int m2l_ntiles = 0;
  int m2l_tstride = 0;      // stride of the tile arrays in `tiles` (a bound on the tile count)
  int m2l_tn = 512;         // targets per M2L tile (512 or 256)
  int m2l_ncell_tiled = 0;  // cells that receive M2L (the tiles' columns)
  Buf<int> m2l_misc, m2l_rs_hist, m2l_scan;
  Buf<float> m2l_part;      // FP32 partial locals of the key slices [slice][cell][P^2]
  // variable-order M2L: order of each (tile, key), parallel to tile_keys
  bool learned = false;
  int m2l_totalU = 0;
  int m2l_ntiles_last = 0;  // tiles of the last keyed M2L (statistics)
  std::map<long long, int> gemm_tuned;  // autotuned GEMM shape (rows, keys per round, slices) per class
  Buf<double> L_tune;                   // scratch locals of the autotuning runs
  Buf<unsigned char> tile_p;
  // BF16x3 learned order: each tile's slots grouped by order (tile_perm:
  // permuted slot -> slot; tile_cls: [first | count] x order x tile stride)
  Buf<int> tile_perm, tile_cls;
  // learned order model (order.hip): FP32 master weights, tolerance
  bool have_model = false;
  float order_log_tol = -5.0f;
  Buf<float> mlp_w, mlp_part;
  Buf<float2> mshape;  // per cell: multipole shape, log10 sum|q| (order selection)
  Buf<double2> absq;  // (sum |q|, sum q^2) per cell (order selection)
  int train_n = 0;
  Buf<float> train_X, train_T;
  std::vector<int> train_pairs;  // (target cell, source cell, image) per sample, internal numbering  // scratch of m2l_prepare (may run on the side stream)
  int ops_P = -1, ops_fill = -1, ops_bf16 = -1;
  double ops_R0 = 0, ops_cycle = 0;
  Buf<float> mt32;                 // scaled multipoles, real form [ncells][P^2]
  Buf<float> lt32;                 // scaled local accumulators [ncells][P^2]
  Buf<int> tiles;                  // M2L tile descriptors
  Buf<int> tile_mat;               // per tile [U][kTileN] source indices
  Buf<int> tile_keys;
  Buf<unsigned> tile_bits;         // per tile: bitmap of its key set (k_tile_count -> k_tile_fill)
  Buf<int> misc;
  // periodic far images
  Buf<double2> far_sup, far_part;

  float stage_ms[8] = {};
};

// ------------------------------------------------------------------ stages
void tree_build(Ctx &c);
// X, q of the caller's bodies in tree order (a build with defer_writeback set)
void writeback_bodies(Ctx &c, hipStream_t s);
void tree_export_cells(Ctx &c, exafmm_cell *out);     // reference numbering
void tree_ref_order(Ctx &c, std::vector<int64_t> &bfs2ref, std::vector<int64_t> &ref2bfs);
// Dual-tree traversal into the interaction lists.  after_kernels (if set)
// runs once per pass, right after the pass's traversal kernels are enqueued
// on s0 and before the host reads the totals (flags in lists.tot[3]; a pass
// with kListRetry bits set is repeated after s1 has drained).
void traverse(Ctx &tgt, const Ctx &src, const std::function<void()> &after_kernels = {});
void upward(Ctx &c, hipStream_t s);
void m2l_apply(Ctx &tgt, const Ctx &src);
void m2l_prepare(Ctx &c, hipStream_t s);
int m2l_fp64_levels(const Ctx &c);  // levels whose M2L runs per pair in FP64 (m2l.hip)
// learned order (order.hip)
constexpr int kMlpIn = 4, kMlpH = 32, kMlpOut = 11, kMlpPmin = 4;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpOut * kMlpH + kMlpOut;
void order_select(Ctx &c, const int *tfirst, const int *tcount, const int *tbase, const int *tU, const int *sorted,
                  int ntiles, int totalU, int tn, hipStream_t s);
float order_train(Ctx &c, int nsamples, int epochs, float lr, uint64_t seed);
void mlp_eval_host(Ctx &c, const float *w, const float *X, int64_t n, float *Y);
void mlp_grad_host(Ctx &c, const float *w, const float *X, const float *T, int64_t n, float *grad, float *loss);  // keyed-M2L tiles and operators (fast precisions only)
void p2p_apply(Ctx &tgt, const Ctx &src, hipStream_t s, const int *list_flags = nullptr);
// grouped launches over the problems of a batch (exafmm_graph_create)
int p2m_group_blocks(int64_t nleaves);
int shift_group_rows(int P);
int m2m_group_rows(int P);
int p2p_group_blocks(int64_t nleaves);
void p2m_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s);
void m2m_grouped(const GroupTree *g, const GroupLevel *lv, const int *start, int n, int nblocks, hipStream_t s);
void l2l_grouped(const GroupTree *g, const GroupLevel *lv, const int *start, int n, int nblocks, hipStream_t s);
void l2p_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s);
void p2p_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s);
// keyed M2L of every problem of one order (BF16x3, fixed order): cast +
// GEMM, one launch each; the plan holds the device descriptor tables
struct GemmGroupPlan {
  int P = 0, rtb = 0, kr = 1, nprob = 0, nblocks = 0, cast_blocks = 0, tn = 512;
  size_t lds = 0;
  Buf<char> desc, cast_desc;
  Buf<int> start, cast_start;
  struct Trial {
    int rtb, kr;
    float ms;
  };
  std::vector<Trial> tuned;  // the autotuner's candidates and their times
};
void m2l_group_setup(const std::vector<Ctx *> &probs, GemmGroupPlan &plan, hipStream_t s, bool tune);
void m2l_group_launch(const GemmGroupPlan &plan, hipStream_t s);
void shift_ops(Ctx &c, hipStream_t s);
void l2l_pass(Ctx &c);
void l2p_pass(Ctx &c);
void direct_sum(int device, hipStream_t s, exafmm_body *t, int64_t nt, const exafmm_body *src, int64_t ns);
void periodic_far(Ctx &c);
void dipole_correction(Ctx &c);
void ewald_sum(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns, int ksize, double alpha,
               double sigma, double cutoff, double cycle);
// allocation-free Ewald sum: wave table in a fixed per-device array
constexpr int kMaxDevices = 64;
constexpr int kEwaldMaxWaves = 32768;  // 40 B each (1.3 MB): ksize <= 25 (32,633 waves)
void ewald_scratch_init(int device);
void ewald_sum_noalloc(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns, int ksize,
                       double alpha, double sigma, double cutoff, double cycle);

// single-operator entry points (FP64, one call == one reference call)
void op_p2p(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns);
void op_p2m(Ctx &c, const double *C, const exafmm_body *b, int64_t nb, double *M);
void op_m2m(Ctx &c, const double *Cp, const double *Cc, const double *Mc, double *Mp);
void op_m2l(Ctx &c, const double *Xi, const double *Xj, const double *Mj, double *Li);
void op_l2l(Ctx &c, const double *Cp, const double *Lp, const double *Cc, double *Lc);
void op_l2p(Ctx &c, const double *C, const double *L, exafmm_body *b, int64_t nb);

// generic helpers (scan.hip)
void scan_exclusive(const int *in, int *out, int count, int *total, Buf<int> &scratch, hipStream_t s);
void scan_exclusive_dev(const int *in, int *out, const int *count_dev, int count_bound, int *total,
                        Buf<int> &scratch, hipStream_t s);
// three arrays at once; running[k] is added to every output and receives the
// new totals; bit 1 of *flag is raised when a total exceeds cap[k]
void scan_exclusive3(const int *in0, const int *in1, const int *in2, int *out0, int *out1, int *out2, int count,
                     int *running, const int64_t *cap, int *flag, Buf<int> &scratch, hipStream_t s);
void radix_sort_pairs(unsigned long long *&k, int *&v, unsigned long long *&k2, int *&v2, int n, int lo_bit,
                      int hi_bit, Buf<int> &hist, hipStream_t s);

}  // namespace xf

// ==================================================================== device helpers
namespace xfd {

__host__ __device__ inline int nms(int n, int m) { return n * (n + 1) / 2 + m; }
// real-form index of (n, m>=0, part) with part 0 = Re, 1 = Im (m > 0 only)
__host__ __device__ inline int rf(int n, int m, int part) { return n * n + (m == 0 ? 0 : 2 * m - 1 + part); }
__host__ __device__ inline int sgn(int n) { return (n & 1) ? -1 : 1; }

template <class T>
struct cpx {
  T re, im;
};
template <class T>
__device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Regular solid harmonics Y_n^m(x,y,z) for 0<=m<=n<P in the reference's
// normalisation (3d/kernel.h:43-80 evaluates the same functions in
// spherical form): Y_n^m = (-1)^n rho^n P_n^m(cos a) e^{imb} / (n+m)!.
// Cartesian recurrences (no trigonometry, finite at r = 0):
//   Y_m^m = Y_{m-1}^{m-1} (x+iy)/(2m),  Y_{m+1}^m = -z Y_m^m,
//   Y_n^m = -((2n-1) z Y_{n-1}^m + r^2 Y_{n-2}^m) / ((n+m)(n-m)).
// Y is stored by column: Y[nms(n,m)].
template <class T>
__device__ inline void regular_all(int P, T x, T y, T z, cpx<T> *Y) {
  const T r2 = x * x + y * y + z * z;
  cpx<T> ymm = {T(1), T(0)};
  for (int m = 0; m < P; m++) {
    if (m > 0) {
      const T inv = T(1) / T(2 * m);
      ymm = {(ymm.re * x - ymm.im * y) * inv, (ymm.re * y + ymm.im * x) * inv};
    }
    Y[nms(m, m)] = ymm;
    if (m + 1 < P) Y[nms(m + 1, m)] = {-z * ymm.re, -z * ymm.im};
    for (int n = m + 2; n < P; n++) {
      const cpx<T> a = Y[nms(n - 1, m)], b = Y[nms(n - 2, m)];
      const T s = T(-1) / T((n + m) * (n - m));
      const T c = T(2 * n - 1) * z;
      Y[nms(n, m)] = {(c * a.re + r2 * b.re) * s, (c * a.im + r2 * b.im) * s};
    }
  }
}

// Singular solid harmonics I_n^m = (-1)^n (n-m)! P_n^m(cos a) e^{imb} / rho^{n+1}
// for 0<=m<=n<Pf, the functions 3d/kernel.h:83-116 evaluates:
//   I_m^m = I_{m-1}^{m-1} (2m-1)(x+iy)/r^2,   I_0^0 = 1/r,
//   I_{n+1}^m = -((2n+1) z I_n^m + (n+m)(n-m) I_{n-1}^m) / r^2.
template <class T>
__device__ inline void singular_all(int Pf, T x, T y, T z, cpx<T> *I) {
  const T r2 = x * x + y * y + z * z;
  const T inv = T(1) / r2;
  cpx<T> imm = {T(1) / sqrt(r2), T(0)};
  for (int m = 0; m < Pf; m++) {
    if (m > 0) {
      const T f = T(2 * m - 1) * inv;
      imm = {(imm.re * x - imm.im * y) * f, (imm.re * y + imm.im * x) * f};
    }
    I[nms(m, m)] = imm;
    cpx<T> a = imm, b = {T(0), T(0)};
    for (int n = m; n + 1 < Pf; n++) {
      const T c = T(2 * n + 1) * z;
      const T d = T((n + m) * (n - m));
      const cpx<T> nx = {-(c * a.re + d * b.re) * inv, -(c * a.im + d * b.im) * inv};
      I[nms(n + 1, m)] = nx;
      b = a;
      a = nx;
    }
  }
}

}  // namespace xfd
