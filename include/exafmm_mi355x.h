/*
 * exafmm_mi355x.h -- C ABI of the MI355X 3-D Laplace FMM evaluation path.
 *
 * Drop-in boundary for minimal-ExaFMM's 3d/ path: every entry point below
 * replaces one reference function (cited per declaration) and keeps its
 * argument meaning; the C++ surface with the reference's own names
 * (buildTree, upwardPass, horizontalPass, downwardPass, direct, P2P ... L2P)
 * is include/exafmm_mi355x.hpp, built on these calls.
 *
 * Conventions
 *  - All functions return 0 on success or a negative EXAFMM_ERR_* code; the
 *    message of the last failure is exafmm_last_error(ctx) (or
 *    exafmm_last_error(NULL) for failures without a context).  Nothing
 *    aborts the process.
 *  - Bodies use the reference's 64-byte layout (3d/exafmm.h:14-19).  A body
 *    pointer is a host pointer (on_device = 0) or a HIP device pointer
 *    (on_device = 1).
 *  - Expansion coefficients are NTERM = P(P+1)/2 complex doubles per cell,
 *    interleaved (re, im), index n(n+1)/2 + m for 0 <= m <= n < P
 *    (3d/kernel.h:196-197), cells in the reference's numbering.
 *  - One context per stream; calls on one context are serialised by the
 *    caller, different contexts may be used from different threads.
 *  - Device-pointer inputs must be complete when the call is made: the
 *    context's own streams do not wait on any other stream, so a caller
 *    that filled the bodies with work on its own stream synchronises that
 *    stream (or the device) first.  Results are complete on device memory
 *    after exafmm_synchronize(ctx); host-array calls return complete.
 *  - Host body arrays: a host pointer given to exafmm_build_tree (or
 *    exafmm_evaluate) is kept by the context until exafmm_downward writes p
 *    and F back into it; the array must stay valid and unmoved until then.
 */
#ifndef EXAFMM_MI355X_H
#define EXAFMM_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXAFMM_ABI_VERSION 1

enum exafmm_status {
  EXAFMM_OK = 0,
  EXAFMM_ERR_ARG = -1,      /* invalid argument (N<1, NaN coordinate, bad config) */
  EXAFMM_ERR_STATE = -2,    /* call out of order (e.g. upward before build_tree) */
  EXAFMM_ERR_DEPTH = -3,    /* tree deeper than 21 levels (> ncrit coincident bodies) */
  EXAFMM_ERR_ALLOC = -4,    /* device allocation failed */
  EXAFMM_ERR_HIP = -5,      /* HIP runtime error */
  EXAFMM_ERR_CAPACITY = -6, /* a fixed-capacity structure overflowed */
  EXAFMM_ERR_NODEVICE = -7, /* no HIP device */
  EXAFMM_ERR_UNSUPPORTED = -8
};

enum exafmm_m2l_truncation {
  EXAFMM_M2L_REFERENCE = 0, /* terms j+n <= P-1, as 3d/kernel.h:83-116,233-260 evaluate them */
  EXAFMM_M2L_FULL = 1       /* every term j+n <= 2P-2 */
};

enum exafmm_precision {
  EXAFMM_PREC_FP64 = 0,   /* FP64 everywhere (parity mode) */
  EXAFMM_PREC_FP32 = 1,   /* FP32 P2P / FP32 MFMA M2L */
  EXAFMM_PREC_BF16X3 = 2, /* M2L only: 3-term BF16 split products, FP32 accumulate */
  EXAFMM_PREC_FP32_NEWTON = 3 /* P2P only: FP32 with one Newton step on the hardware
                                 1/sqrt and Kahan-compensated partial sums */
};

enum exafmm_order_policy {
  EXAFMM_ORDER_FIXED = 0,  /* expansion order P for every M2L pair */
  EXAFMM_ORDER_LEARNED = 1 /* per-pair order from the order model (exafmm_set_order_model) */
};

/* Reference Body (3d/exafmm.h:14-19); F is +grad(phi) in the reference's sign. */
typedef struct exafmm_body {
  double X[3];
  double q;
  double p;
  double F[3];
} exafmm_body;

/* Flattened reference Cell (3d/exafmm.h:23-36): pointers become indices. */
typedef struct exafmm_cell {
  int32_t nchild;
  int32_t nbody;
  int64_t child; /* index of the first child, -1 for a leaf */
  int64_t body;  /* index of the first body in the permuted body array */
  double X[3];
  double R;
} exafmm_cell;

/* Replaces the mutable globals P, ncrit, theta (3d/exafmm.h:40-45) and the
 * periodic globals images, cycle (3dp/exafmm.h:45-47). */
typedef struct exafmm_config {
  int32_t P;              /* expansion order, 2..16 */
  int32_t ncrit;          /* max bodies per leaf, 1..1024 */
  double theta;           /* multipole acceptance criterion */
  int32_t m2l_truncation; /* enum exafmm_m2l_truncation */
  int32_t m2l_precision;  /* enum exafmm_precision */
  int32_t p2p_precision;  /* EXAFMM_PREC_FP64, EXAFMM_PREC_FP32 or EXAFMM_PREC_FP32_NEWTON */
  int32_t images;         /* 0: free space; >=1: periodic, 3dp semantics */
  double cycle;           /* periodic cycle (images >= 1) */
  int32_t order_policy;   /* enum exafmm_order_policy */
  int32_t device;         /* HIP device ordinal */
  int32_t autotune;       /* 1: time the keyed-M2L GEMM shapes once per shape class and keep the fastest */
  int32_t reserved[5];
} exafmm_config;

typedef struct exafmm_ctx exafmm_ctx;

/* Defaults: P=10, ncrit=64, theta=0.4, full truncation, BF16X3 M2L, FP32
 * P2P, free space, fixed order, device 0. */
int exafmm_config_default(exafmm_config *cfg);
int exafmm_device_count(int *count);
const char *exafmm_last_error(const exafmm_ctx *ctx);

int exafmm_create(const exafmm_config *cfg, exafmm_ctx **out);
int exafmm_destroy(exafmm_ctx *ctx);
int exafmm_synchronize(exafmm_ctx *ctx);

/* buildTree (3d/build_tree.h:88-96): octree of `bodies`; X and q are
 * permuted in place exactly as the reference permutes them, p and F are
 * left in place.  The array must stay valid until exafmm_downward. */
int exafmm_build_tree(exafmm_ctx *ctx, exafmm_body *bodies, int64_t n, int on_device);
int exafmm_num_cells(exafmm_ctx *ctx, int64_t *ncells, int32_t *nlevels);
/* The cell table in the reference's numbering (children blocks in DFS
 * preorder of their parents). */
int exafmm_get_cells(exafmm_ctx *ctx, exafmm_cell *out, int64_t capacity);

/* upwardPass (3d/traverse_eager.h:127-144): P2M at leaves, M2M upward. */
int exafmm_upward(exafmm_ctx *ctx);
/* horizontalPass (3d/traverse_eager.h:147-172, lazy form
 * 3d/traverse_lazy.h:27-56): dual-tree traversal of (target, source)
 * emitting M2L and P2P lists, then M2L into the target's locals and P2P
 * into the target's bodies.  source may equal target. */
int exafmm_horizontal(exafmm_ctx *target, exafmm_ctx *source);
/* downwardPass (3d/traverse_eager.h:175-190): L2L downward, L2P at
 * leaves; p and F are accumulated (+=) into the bodies array. */
int exafmm_downward(exafmm_ctx *ctx);
/* buildTree + upwardPass + horizontalPass(self) + downwardPass in one call
 * (3d/fmm.cxx:40-53). */
int exafmm_evaluate(exafmm_ctx *ctx, exafmm_body *bodies, int64_t n, int on_device);
/* exafmm_evaluate of nprob (<= 512) independent free-space problems in one
 * call: problem p is bodies[offsets[p] .. offsets[p+1]-1] (offsets[0] = 0,
 * offsets[nprob] = total), with its own root cube (3d/build_tree.h:7-21)
 * and tree (3d/build_tree.h:88-96) and no interaction with the others.
 * The trees hang under three levels of cells indexed by the problem number
 * (never paired by the traversal), so one tree build, one traversal and one
 * launch per pass serve all problems, and the keyed M2L packs the cells of
 * every problem into shared tiles with one operator per translation key
 * (keys are scale-free).  Each problem's X and q are permuted within its
 * own range as exafmm_evaluate permutes them.  No counterpart in the
 * reference (one problem per run, 3d/fmm.cxx:40-53). */
int exafmm_evaluate_forest(exafmm_ctx *ctx, exafmm_body *bodies, const int64_t *offsets, int32_t nprob,
                           int on_device);

/* direct (3d/traverse_eager.h:193-202): FP64 direct sum of every source on
 * every target, accumulated into the targets' p and F. */
int exafmm_direct(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources,
                  int64_t ns, int on_device);

/* Periodic contexts (config.images >= 1, 3dp semantics): build_tree,
 * upward and downward are unchanged; horizontal visits the 27 near images
 * (3dp/traverse_lazy.h:124-135) and then the images-1 far levels
 * (3dp/traverse_lazy.h:75-121).  The dipole correction of 3dp/fmm.cxx:63-75
 * is a separate call made after downward. */
int exafmm_dipole_correction(exafmm_ctx *ctx);
/* Ewald summation oracle (3dp/ewald.h:22-150) in FP64 on the device, with
 * the context's periodic cycle: p and F of the targets += Ewald sums over
 * the sources (wave part over |k| <= ksize, real part within cutoff using
 * each pair's nearest image, self term). */
int exafmm_ewald(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources, int64_t ns,
                 int ksize, double alpha, double sigma, double cutoff, int on_device);
/* exafmm_ewald without allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  targets and
 * sources are device pointers; the wave table lives in a fixed array of
 * 32,768 waves per GPU (ksize <= 25; larger ksize returns
 * EXAFMM_ERR_CAPACITY), set up by exafmm_create on the context's device.
 * Results equal exafmm_ewald(..., on_device = 1) bit for bit.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (several
 * host threads or contexts at once) race. */
int exafmm_ewald_noalloc(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources,
                         int64_t ns, int ksize, double alpha, double sigma, double cutoff);

/* Interaction lists of the last horizontal pass (Cell::listM2L /
 * Cell::listP2P, 3d/exafmm.h:30-33), reference cell numbering, CSR by
 * target cell (offsets has ncells+1 entries). */
int exafmm_list_sizes(exafmm_ctx *ctx, int64_t *m2l_pairs, int64_t *p2p_pairs);
int exafmm_get_lists(exafmm_ctx *ctx, int which, int64_t *offsets, int64_t *sources);

/* Cell::M / Cell::L (3d/exafmm.h:34-35) of every cell, reference numbering. */
int exafmm_get_multipoles(exafmm_ctx *ctx, double *out);
int exafmm_get_locals(exafmm_ctx *ctx, double *out);

/* Stage times (ms, device clock) of the last build/upward/horizontal/
 * downward calls: [tree, lists, p2m+m2m, m2l, p2p, l2l+l2p, total,
 * m2l kernel alone]. */
int exafmm_stage_times(exafmm_ctx *ctx, double *ms, int capacity);

/* Work of the last horizontal pass, for throughput and roofline figures:
 * [0] M2L pairs, [1] P2P leaf pairs, [2] P2P body-body interactions
 * (sum of ni*nj), [3] P2P source bodies read (sum of nj), [4] cells,
 * [5] leaves, [6] distinct M2L translation keys, [7] deepest level. */
int exafmm_work_counts(exafmm_ctx *ctx, int64_t *out, int capacity);

/* ---- the kernel.h operator surface, one reference call per entry ------
 * Host arrays; the work runs on the context's device in FP64 unless the
 * context's precision fields say otherwise.  Coefficient arrays are
 * NTERM complex doubles (see conventions).  */
/* P2P(Ci,Cj) (3d/kernel.h:122-150): targets += interaction with sources. */
int exafmm_p2p(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources, int64_t ns);
/* P2M(C) (3d/kernel.h:187-202): M += multipole of bodies about center. */
int exafmm_p2m(exafmm_ctx *ctx, const double center[3], const exafmm_body *bodies, int64_t nb, double *M);
/* M2M(Ci) (3d/kernel.h:204-231), one child: Mparent += shift(Mchild). */
int exafmm_m2m(exafmm_ctx *ctx, const double cparent[3], const double cchild[3], const double *Mchild,
               double *Mparent);
/* M2L(Ci,Cj) (3d/kernel.h:233-260): Li += translate(Mj, Xi - Xj). */
int exafmm_m2l(exafmm_ctx *ctx, const double Xi[3], const double Xj[3], const double *Mj, double *Li);
/* L2L(Cj) (3d/kernel.h:262-291), one child: Lchild += shift(Lparent). */
int exafmm_l2l(exafmm_ctx *ctx, const double cparent[3], const double *Lparent, const double cchild[3],
               double *Lchild);
/* L2P(Ci) (3d/kernel.h:293-322): bodies p, F += local expansion. */
int exafmm_l2p(exafmm_ctx *ctx, const double center[3], const double *L, exafmm_body *bodies, int64_t nb);

/* ---- learned per-pair expansion order (no counterpart in the reference,
 * which uses one order P for every M2L pair, 3d/kernel.h:233-260).
 * Order model: MLP 4 -> 32 -> 32 -> 11 (ReLU), EXAFMM_MLP_PARAMS FP32
 * master weights, evaluated with BF16-rounded weights and FP32
 * accumulation.  Inputs per pair: log2(Rs/Rt), log2(|d|/(Rt+Rs)),
 * level/8, log10 share of the source multipole norm in degrees >= 2.
 * Outputs: predicted log10 truncation error at orders 4..14 (normalised
 * by sum|q_source|/|d|).  With order_policy = EXAFMM_ORDER_LEARNED and a
 * model set, each M2L pair runs at the smallest order whose prediction is
 * <= log10_tol, at most the context's P (which is then the maximum order of
 * the expansions); without a model the context's P is used for every pair. */
#define EXAFMM_MLP_PARAMS 1579
/* Trains the order model on the device from the context's current
 * evaluation (P = 14, after build_tree, upward and horizontal): nsamples M2L
 * pairs are drawn from the lists, their errors at orders 4..14 measured in
 * FP64 against the direct sum, and the MLP fitted with Adam for `epochs`
 * full-batch steps.  *loss = final mean squared error (log10 units). */
int exafmm_order_model_train(exafmm_ctx *ctx, int nsamples, int epochs, double lr, uint64_t seed, double *loss);
int exafmm_set_order_model(exafmm_ctx *ctx, const float *weights, int count, double log10_tol);
int exafmm_get_order_model(exafmm_ctx *ctx, float *weights, int count);
/* The last training set: X [n][4], T [n][11] (NaN = no data) and, when
 * pairs is non-null, pairs [n][3] = (target cell, source cell, image) in
 * the numbering of exafmm_get_cells (image 13 = no periodic shift). */
int exafmm_order_training_set(exafmm_ctx *ctx, float *X, float *T, int32_t *pairs, int64_t capacity, int64_t *n);
/* Device forward pass / loss gradient of the order MLP on host arrays
 * (loss = sum over samples and non-NaN outputs of (y - t)^2 / n). */
int exafmm_mlp_forward(exafmm_ctx *ctx, const float *weights, const float *X, int64_t n, float *Y);
int exafmm_mlp_gradient(exafmm_ctx *ctx, const float *weights, const float *X, const float *T, int64_t n,
                        float *grad, float *loss);
/* Orders used by the last keyed M2L: hist[p] = number of (target, key)
 * slots of the GEMM tiles run at order p (p < capacity). */
int exafmm_order_stats(exafmm_ctx *ctx, int64_t *hist, int capacity);

/* ---- batched evaluation in a HIP graph (no counterpart in the reference,
 * whose drivers evaluate one problem per process).  A graph re-evaluates a
 * set of independent problems with one hipGraphLaunch: every context must
 * have been evaluated once with device-resident bodies (exafmm_evaluate,
 * on_device = 1; not periodic).  Creating the graph keeps each context's
 * tree, lists, keyed-M2L tiles, operators and autotuned GEMM shape, and
 * records per problem the passes that depend on the charges (P2M, M2M,
 * P2P, M2L, L2L, L2P, after zeroing p, F and the locals) on one of
 * `nstreams` (1..16) capture streams, so independent problems run as
 * parallel branches.  A launch reads the current charges q of the
 * contexts' (tree-ordered) body arrays, positions as at creation, and
 * overwrites p and F.  While a graph holds a context, building a tree on
 * it or destroying it fails with EXAFMM_ERR_STATE.  Errors of these calls
 * are reported by exafmm_last_error(NULL). */
typedef struct exafmm_graph exafmm_graph;
int exafmm_graph_create(exafmm_ctx *const *ctxs, int count, int nstreams, exafmm_graph **out);
/* The same with one launch per stage over all problems instead of per-problem
 * launches: P2M, P2P, L2P and the charge prologue once each, M2M and L2L once
 * per tree level, the keyed-M2L cast and GEMM once per expansion order (one
 * GEMM shape per order: the largest row block, one key per round, no key
 * slices).  Needs BF16x3 M2L, FP32 P2P and fixed orders. */
int exafmm_graph_create_grouped(exafmm_ctx *const *ctxs, int count, exafmm_graph **out);
int exafmm_graph_launch(exafmm_graph *graph);      /* asynchronous */
int exafmm_graph_synchronize(exafmm_graph *graph);
int exafmm_graph_nodes(exafmm_graph *graph, int64_t *nodes);
/* Keyed-GEMM shapes of a grouped graph, as picked by its autotuner (every
 * candidate timed on the batch itself when the contexts' config has
 * autotune = 1): per expansion order, shapes[3k..3k+2] = (P, operator rows
 * per workgroup, keys per LDS round); *count = number of orders. */
int exafmm_graph_gemm_shapes(exafmm_graph *graph, int32_t *shapes, int capacity, int *count);
int exafmm_graph_destroy(exafmm_graph *graph);

/* ---- batched evaluation through one context (the survey's
 * exafmm_evaluate_batch; no counterpart in the reference).  Evaluates every
 * problem (tree, lists, all passes) in sub-contexts that the context keeps
 * with its own configuration and the problem's order P.  With
 * EXAFMM_BATCH_REUSE_GEOMETRY and device-resident bodies, the call also
 * records the problems in a graph (grouped launches when the configuration
 * allows, else per-problem launches); a later call with the same problems
 * (same arrays, sizes and orders) and the flag replays that graph: the
 * positions of the previous call are kept (the arrays are in tree order
 * after the first call), the charges are read afresh, p and F overwritten.
 * Recording ends with one pass of the graph, so a recording call returns
 * the graph's results (a grouped graph's GEMM shapes sum the M2L partials
 * in another order than the eager pass: ~1e-7 relative differences). */
typedef struct exafmm_problem {
  exafmm_body *bodies;
  int64_t n;
  int32_t P;
  int32_t on_device;
} exafmm_problem;
#define EXAFMM_BATCH_REUSE_GEOMETRY 1
/* EXAFMM_BATCH_FOREST: the problems of each order are gathered into one
 * array and evaluated as one forest (exafmm_evaluate_forest), then copied
 * back: every call builds the trees and lists of all problems afresh, with
 * one launch per pass over all problems of an order. */
#define EXAFMM_BATCH_FOREST 2
int exafmm_evaluate_batch(exafmm_ctx *ctx, const exafmm_problem *problems, int count, int flags);

#ifdef __cplusplus
}
#endif
#endif
