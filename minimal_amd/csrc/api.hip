// api.hip -- the C ABI (include/exafmm_mi355x.h): context lifetime, the
// pass-level drop-in for buildTree / upwardPass / horizontalPass /
// downwardPass / direct, exports in the reference's numbering, and stage
// timing with HIP events.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <memory>

#include "fmm_internal.h"

struct exafmm_ctx {
  xf::Ctx c;
  // exafmm_evaluate_batch: one sub-context per problem slot, and the graph
  // that re-evaluates them when the caller keeps the positions
  std::vector<exafmm_ctx *> batch;
  exafmm_graph *batch_graph = nullptr;
  std::vector<std::pair<const void *, int64_t>> batch_sig;
  // EXAFMM_BATCH_FOREST: one forest context per order and its gathered bodies
  std::vector<exafmm_ctx *> forests;
  std::vector<std::unique_ptr<xf::Buf<exafmm_body>>> forest_bodies;
  std::vector<std::unique_ptr<xf::Buf<xf::ForestPart>>> forest_parts;  // gather / scatter tables
  std::vector<std::vector<xf::ForestPart>> forest_parts_host;
};

namespace xf {

void fail(int code, const std::string &msg) { throw Failure{code, msg}; }

namespace {

thread_local std::string g_error;

template <class F>
int guard(exafmm_ctx *ctx, F &&f) {
  try {
    if (ctx) XF_HIP(hipSetDevice(ctx->c.cfg.device));
    f();
    return EXAFMM_OK;
  } catch (const Failure &e) {
    (ctx ? ctx->c.last_error : g_error) = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    (ctx ? ctx->c.last_error : g_error) = "host allocation failed";
    return EXAFMM_ERR_ALLOC;
  } catch (const std::exception &e) {
    (ctx ? ctx->c.last_error : g_error) = e.what();
    return EXAFMM_ERR_ARG;
  }
}

void check_config(const exafmm_config &cfg) {
  if (cfg.P < 2 || cfg.P > kMaxP) fail(EXAFMM_ERR_ARG, "P must be in [2, 16]");
  if (cfg.ncrit < 1 || cfg.ncrit > 1024) fail(EXAFMM_ERR_ARG, "ncrit must be in [1, 1024]");
  if (!(cfg.theta > 0) || !(cfg.theta < 1.5)) fail(EXAFMM_ERR_ARG, "theta must be in (0, 1.5)");
  if (cfg.m2l_truncation != EXAFMM_M2L_REFERENCE && cfg.m2l_truncation != EXAFMM_M2L_FULL)
    fail(EXAFMM_ERR_ARG, "bad m2l_truncation");
  if (cfg.m2l_precision < EXAFMM_PREC_FP64 || cfg.m2l_precision > EXAFMM_PREC_BF16X3)
    fail(EXAFMM_ERR_ARG, "bad m2l_precision");
  if (cfg.p2p_precision != EXAFMM_PREC_FP64 && cfg.p2p_precision != EXAFMM_PREC_FP32 &&
      cfg.p2p_precision != EXAFMM_PREC_FP32_NEWTON)
    fail(EXAFMM_ERR_ARG, "bad p2p_precision");
  if (cfg.images < 0 || cfg.images > 8) fail(EXAFMM_ERR_ARG, "images must be in [0, 8]");
  if (cfg.images > 0 && !(cfg.cycle > 0)) fail(EXAFMM_ERR_ARG, "periodic cycle must be positive");
  if (cfg.order_policy != EXAFMM_ORDER_FIXED && cfg.order_policy != EXAFMM_ORDER_LEARNED)
    fail(EXAFMM_ERR_ARG, "bad order_policy");
}

void rec(Ctx &c, int i, hipStream_t s) { XF_HIP(hipEventRecord(c.ev[i], s)); }

float elapsed(Ctx &c, int a, int b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, c.ev[a], c.ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return ms;
}

void set_bodies(Ctx &c, exafmm_body *b, int64_t n, int on_device) {
  if (c.graph_locks) fail(EXAFMM_ERR_STATE, "the context is held by a graph (exafmm_graph_destroy it first)");
  if (!b || n < 1) fail(EXAFMM_ERR_ARG, "need at least one body");
  c.t.n = n;
  c.forest_n = 0;  // one problem unless exafmm_evaluate_forest says otherwise
  if (on_device) {
    c.bodies = b;
    c.host_mode = false;
    c.host_bodies = nullptr;
  } else {
    c.mirror.ensure(n);
    XF_HIP(hipMemcpyAsync(c.mirror.p, b, sizeof(exafmm_body) * n, hipMemcpyHostToDevice, c.s0));
    c.bodies = c.mirror.p;
    c.host_mode = true;
    c.host_bodies = b;
  }
}

void bodies_to_host(Ctx &c) {
  if (!c.host_mode) return;
  XF_HIP(hipMemcpyAsync(c.host_bodies, c.bodies, sizeof(exafmm_body) * c.t.n, hipMemcpyDeviceToHost, c.s0));
  XF_HIP(hipStreamSynchronize(c.s0));
}

// event slots: 0/1 build, 2/3 upward, 11/12 traversal, 9/4 M2L (s0),
// 5/6 P2P (s1), 15 end of s1's work, 10/8 downward
void do_build(Ctx &c) {
  if (c.graph_locks) fail(EXAFMM_ERR_STATE, "the context is held by a graph (exafmm_graph_destroy it first)");
  rec(c, 0, c.s0);
  tree_build(c);
  rec(c, 1, c.s0);
  c.have_M = c.have_L = false;
  c.lists.valid = false;
}

void do_upward(Ctx &c, hipStream_t s) {
  if (!c.t.built) fail(EXAFMM_ERR_STATE, "upward pass before build_tree");
  rec(c, 2, s);
  upward(c, s);
  rec(c, 3, s);
}

// Interaction lists (geometry only: no multipoles needed).
void do_lists(Ctx &tc, const Ctx &sc, const std::function<void()> &after_kernels = {}) {
  if (!tc.t.built || !sc.t.built) fail(EXAFMM_ERR_STATE, "horizontal pass before build_tree");
  if (tc.graph_locks) fail(EXAFMM_ERR_STATE, "the context is held by a graph (exafmm_graph_destroy it first)");
  if (tc.P != sc.P) fail(EXAFMM_ERR_ARG, "target and source contexts differ in P");
  if (tc.cfg.images != sc.cfg.images || (tc.cfg.images > 0 && &tc != &sc))
    fail(EXAFMM_ERR_UNSUPPORTED, "a periodic horizontal pass needs source == target");
  rec(tc, 11, tc.s0);
  traverse(tc, sc, after_kernels);
  rec(tc, 12, tc.s0);
  tc.L.ensure(size_t(tc.t.ncells) * tc.P * tc.P);
  XF_HIP(hipMemsetAsync(tc.L.p, 0, sizeof(double) * tc.t.ncells * tc.P * tc.P, tc.s0));
}

// P2P on the second stream from the current point of the main stream, so it
// overlaps whatever the main stream runs next (upward pass, M2L).
// after_lists: event 14 already marks the end of the lists on s0
void do_p2p(Ctx &tc, const Ctx &sc, bool after_lists = false) {
  if (!after_lists) rec(tc, 14, tc.s0);
  XF_HIP(hipStreamWaitEvent(tc.s1, tc.ev[14], 0));
  rec(tc, 5, tc.s1);
  p2p_apply(tc, sc, tc.s1);
  rec(tc, 6, tc.s1);
  rec(tc, 15, tc.s1);
}

// M2L (and the periodic far field) on the main stream; joins the P2P stream.
void do_far(Ctx &tc, const Ctx &sc) {
  if (!sc.have_M) fail(EXAFMM_ERR_STATE, "horizontal pass before the source's upward pass");
  rec(tc, 9, tc.s0);
  rec(tc, 13, tc.s0);
  m2l_apply(tc, sc);
  rec(tc, 4, tc.s0);
  if (tc.cfg.images > 0) periodic_far(tc);
  XF_HIP(hipStreamWaitEvent(tc.s0, tc.ev[15], 0));  // the end of the second stream's work
  tc.src = &sc;
  tc.have_L = true;
}

void do_horizontal(Ctx &tc, const Ctx &sc) {
  if (!sc.have_M) fail(EXAFMM_ERR_STATE, "horizontal pass before the source's upward pass");
  do_lists(tc, sc);
  do_p2p(tc, sc);
  do_far(tc, sc);
}

void do_downward(Ctx &c) {
  if (!c.have_L) fail(EXAFMM_ERR_STATE, "downward pass before horizontal pass");
  rec(c, 10, c.s0);
  l2l_pass(c);
  l2p_pass(c);
  rec(c, 8, c.s0);
}

// One evaluation over two streams.  P2P needs only the P2P lists: it is
// launched on s1 as soon as the traversal kernels are enqueued (before the
// host reads the list totals; a pass that has to be repeated makes it exit
// at once), and runs beside the rest of the lists, the upward pass and the
// M2L tile and operator build on s0, its long VALU-bound blocks filling the
// CUs those latency-bound chains leave idle; the keyed GEMM follows on s0.
// A periodic tree launches P2P after the lists and the upward pass instead:
// its far field adds a long latency-bound chain on s0 (the top levels' FP64
// M2L per pair, the image ladder) that P2P started earlier would starve
// (N = 2^20, images = 3: 6.5 ms against 6.9).  (Running the upward pass
// beside the traversal is slower too: its workgroups hold the CUs each
// short traversal launch waits for.)
void do_passes(Ctx &c) {
  {
    // the bodies' write-back in tree order runs on s1 after P2P, off the
    // latency chain (nothing in the evaluation reads their X, q)
    struct Defer {
      Ctx &c;
      ~Defer() { c.defer_writeback = false; }
    } defer{c};
    c.defer_writeback = true;
    do_build(c);
  }
  // P2P on s1 from event 14 (the lists' end on s0)
  auto launch_p2p = [&c](const int *list_flags) {
    XF_HIP(hipStreamWaitEvent(c.s1, c.ev[14], 0));
    rec(c, 5, c.s1);
    p2p_apply(c, c, c.s1, list_flags);
    rec(c, 6, c.s1);
    if (c.writeback_pending) writeback_bodies(c, c.s1);
    rec(c, 15, c.s1);
  };
  if (c.cfg.images > 0) {
    do_lists(c, c);
    rec(c, 14, c.s0);
    do_upward(c, c.s0);
    launch_p2p(nullptr);
  } else {
    do_lists(c, c, [&] {
      rec(c, 14, c.s0);
      launch_p2p(c.lists.tot.p + 3);
    });
    do_upward(c, c.s0);
  }
  do_far(c, c);
  do_downward(c);
}

}  // namespace
}  // namespace xf

using namespace xf;

extern "C" {

int exafmm_config_default(exafmm_config *cfg) {
  if (!cfg) return EXAFMM_ERR_ARG;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->P = 10;
  cfg->ncrit = 64;
  cfg->theta = 0.4;
  cfg->m2l_truncation = EXAFMM_M2L_FULL;
  cfg->m2l_precision = EXAFMM_PREC_BF16X3;
  cfg->p2p_precision = EXAFMM_PREC_FP32;
  cfg->images = 0;
  cfg->cycle = 0;
  cfg->order_policy = EXAFMM_ORDER_FIXED;
  cfg->autotune = 1;
  cfg->device = 0;
  return EXAFMM_OK;
}

int exafmm_device_count(int *count) {
  return guard(nullptr, [&] {
    if (!count) fail(EXAFMM_ERR_ARG, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}

const char *exafmm_last_error(const exafmm_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : g_error.c_str(); }

int exafmm_create(const exafmm_config *cfg, exafmm_ctx **out) {
  return guard(nullptr, [&] {
    if (!cfg || !out) fail(EXAFMM_ERR_ARG, "null argument");
    *out = nullptr;
    check_config(*cfg);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      fail(EXAFMM_ERR_NODEVICE, "no HIP device");
    }
    if (cfg->device < 0 || cfg->device >= ndev) fail(EXAFMM_ERR_ARG, "device ordinal out of range");
    XF_HIP(hipSetDevice(cfg->device));
    ewald_scratch_init(cfg->device);  // address of this device's fixed scratch array
    auto *h = new exafmm_ctx();
    Ctx &c = h->c;
    c.cfg = *cfg;
    c.P = cfg->P;
    c.NT = cfg->P * (cfg->P + 1) / 2;
    try {
      XF_HIP(hipStreamCreateWithFlags(&c.s0, hipStreamNonBlocking));
      // EXAFMM_SERIAL_STREAMS=1 runs P2P on the main stream (per-kernel
      // profiles without the P2P / M2L overlap)
      const char *serial = std::getenv("EXAFMM_SERIAL_STREAMS");
      if (serial && serial[0] == '1')
        c.s1 = c.s0;
      else
        XF_HIP(hipStreamCreateWithFlags(&c.s1, hipStreamNonBlocking));
      for (auto &e : c.ev) XF_HIP(hipEventCreate(&e));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int exafmm_destroy(exafmm_ctx *ctx) {
  if (!ctx) return EXAFMM_ERR_ARG;
  if (ctx->c.graph_locks) {
    ctx->c.last_error = "the context is held by a graph (exafmm_graph_destroy it first)";
    return EXAFMM_ERR_STATE;
  }
  if (ctx->batch_graph) exafmm_graph_destroy(ctx->batch_graph);
  ctx->batch_graph = nullptr;
  for (exafmm_ctx *b : ctx->batch) exafmm_destroy(b);
  ctx->batch.clear();
  for (exafmm_ctx *f : ctx->forests) exafmm_destroy(f);
  ctx->forests.clear();
  ctx->forest_bodies.clear();
  ctx->forest_parts.clear();
  hipSetDevice(ctx->c.cfg.device);
  if (ctx->c.s0) hipStreamSynchronize(ctx->c.s0);
  if (ctx->c.s1) hipStreamSynchronize(ctx->c.s1);
  for (auto &e : ctx->c.ev)
    if (e) hipEventDestroy(e);
  if (ctx->c.s0) hipStreamDestroy(ctx->c.s0);
  if (ctx->c.s1 && ctx->c.s1 != ctx->c.s0) hipStreamDestroy(ctx->c.s1);
  delete ctx;
  return EXAFMM_OK;
}

int exafmm_synchronize(exafmm_ctx *ctx) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    XF_HIP(hipStreamSynchronize(ctx->c.s1));
    XF_HIP(hipStreamSynchronize(ctx->c.s0));
  });
}

int exafmm_build_tree(exafmm_ctx *ctx, exafmm_body *bodies, int64_t n, int on_device) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    set_bodies(c, bodies, n, on_device);
    do_build(c);
    bodies_to_host(c);
  });
}

int exafmm_num_cells(exafmm_ctx *ctx, int64_t *ncells, int32_t *nlevels) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    if (!ctx->c.t.built) fail(EXAFMM_ERR_STATE, "no tree");
    if (ncells) *ncells = ctx->c.t.ncells;
    if (nlevels) *nlevels = ctx->c.t.depth + 1;
  });
}

int exafmm_get_cells(exafmm_ctx *ctx, exafmm_cell *out, int64_t capacity) {
  return guard(ctx, [&] {
    if (!ctx || !out) fail(EXAFMM_ERR_ARG, "null argument");
    if (!ctx->c.t.built) fail(EXAFMM_ERR_STATE, "no tree");
    if (capacity < ctx->c.t.ncells) fail(EXAFMM_ERR_ARG, "capacity below the number of cells");
    XF_HIP(hipStreamSynchronize(ctx->c.s0));
    tree_export_cells(ctx->c, out);
  });
}

int exafmm_upward(exafmm_ctx *ctx) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    do_upward(ctx->c, ctx->c.s0);
  });
}

int exafmm_horizontal(exafmm_ctx *target, exafmm_ctx *source) {
  return guard(target, [&] {
    if (!target || !source) fail(EXAFMM_ERR_ARG, "null context");
    if (source->c.cfg.device != target->c.cfg.device) fail(EXAFMM_ERR_ARG, "contexts on different devices");
    if (source != target) XF_HIP(hipStreamSynchronize(source->c.s0));
    do_horizontal(target->c, source->c);
  });
}

int exafmm_downward(exafmm_ctx *ctx) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    do_downward(ctx->c);
    bodies_to_host(ctx->c);
  });
}

int exafmm_evaluate(exafmm_ctx *ctx, exafmm_body *bodies, int64_t n, int on_device) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    set_bodies(c, bodies, n, on_device);
    do_passes(c);
    bodies_to_host(c);
  });
}

int exafmm_evaluate_forest(exafmm_ctx *ctx, exafmm_body *bodies, const int64_t *offsets, int32_t nprob,
                           int on_device) {
  return guard(ctx, [&] {
    if (!ctx || !offsets) fail(EXAFMM_ERR_ARG, "null argument");
    if (nprob < 1 || nprob > kForestMax) fail(EXAFMM_ERR_ARG, "number of problems must be in [1, 512]");
    if (offsets[0] != 0) fail(EXAFMM_ERR_ARG, "offsets[0] must be 0");
    for (int p = 0; p < nprob; p++)
      if (offsets[p + 1] < offsets[p]) fail(EXAFMM_ERR_ARG, "offsets must not decrease");
    Ctx &c = ctx->c;
    if (c.cfg.images > 0) fail(EXAFMM_ERR_UNSUPPORTED, "a forest evaluation is free-space only");
    set_bodies(c, bodies, offsets[nprob], on_device);
    c.forest_n = nprob;
    c.forest_off.ensure(nprob + 1);
    c.forest_roots.ensure(nprob);
    // synchronous upload: the caller's offsets need not outlive the call
    XF_HIP(hipMemcpy(c.forest_off.p, offsets, sizeof(int64_t) * (nprob + 1), hipMemcpyHostToDevice));
    do_passes(c);
    bodies_to_host(c);
  });
}

int exafmm_stage_times(exafmm_ctx *ctx, double *ms, int capacity) {
  return guard(ctx, [&] {
    if (!ctx || !ms) fail(EXAFMM_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    XF_HIP(hipStreamSynchronize(c.s1));
    XF_HIP(hipStreamSynchronize(c.s0));
    const double v[8] = {elapsed(c, 0, 1), elapsed(c, 11, 12), elapsed(c, 2, 3), elapsed(c, 9, 4),
                         elapsed(c, 5, 6), elapsed(c, 10, 8), elapsed(c, 0, 8), elapsed(c, 13, 4)};
    for (int i = 0; i < std::min(capacity, 8); i++) ms[i] = v[i];
  });
}

int exafmm_direct(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources, int64_t ns,
                  int on_device) {
  return guard(ctx, [&] {
    if (!ctx || !targets || !sources || nt < 0 || ns < 0) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    if (on_device) {
      direct_sum(c.cfg.device, c.s0, targets, nt, sources, ns);
      XF_HIP(hipStreamSynchronize(c.s0));
    } else {
      op_p2p(c, targets, nt, sources, ns);
    }
  });
}

int exafmm_dipole_correction(exafmm_ctx *ctx) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    if (c.cfg.images < 1) fail(EXAFMM_ERR_STATE, "dipole correction needs a periodic context");
    if (!c.t.built) fail(EXAFMM_ERR_STATE, "no bodies");
    dipole_correction(c);
    bodies_to_host(c);
    XF_HIP(hipStreamSynchronize(c.s0));
  });
}

int exafmm_ewald(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources, int64_t ns,
                 int ksize, double alpha, double sigma, double cutoff, int on_device) {
  return guard(ctx, [&] {
    if (!ctx || !targets || !sources || nt < 0 || ns < 0 || ksize < 0 || !(alpha > 0) || !(sigma > 0))
      fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    if (!(c.cfg.cycle > 0)) fail(EXAFMM_ERR_ARG, "Ewald sum needs the periodic cycle in the configuration");
    if (on_device) {
      ewald_sum(c, targets, nt, sources, ns, ksize, alpha, sigma, cutoff, c.cfg.cycle);
    } else {
      Buf<exafmm_body> dt, ds;
      dt.ensure(nt);
      ds.ensure(ns);
      XF_HIP(hipMemcpyAsync(dt.p, targets, sizeof(exafmm_body) * nt, hipMemcpyHostToDevice, c.s0));
      XF_HIP(hipMemcpyAsync(ds.p, sources, sizeof(exafmm_body) * ns, hipMemcpyHostToDevice, c.s0));
      ewald_sum(c, dt.p, nt, ds.p, ns, ksize, alpha, sigma, cutoff, c.cfg.cycle);
      XF_HIP(hipMemcpy(targets, dt.p, sizeof(exafmm_body) * nt, hipMemcpyDeviceToHost));
    }
  });
}

int exafmm_ewald_noalloc(exafmm_ctx *ctx, exafmm_body *targets, int64_t nt, const exafmm_body *sources,
                         int64_t ns, int ksize, double alpha, double sigma, double cutoff) {
  return guard(ctx, [&] {
    if (!ctx || !targets || !sources || nt < 0 || ns < 0 || ksize < 0 || !(alpha > 0) || !(sigma > 0))
      fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    if (!(c.cfg.cycle > 0)) fail(EXAFMM_ERR_ARG, "Ewald sum needs the periodic cycle in the configuration");
    ewald_sum_noalloc(c, targets, nt, sources, ns, ksize, alpha, sigma, cutoff, c.cfg.cycle);
  });
}

int exafmm_list_sizes(exafmm_ctx *ctx, int64_t *m2l_pairs, int64_t *p2p_pairs) {
  return guard(ctx, [&] {
    if (!ctx) fail(EXAFMM_ERR_ARG, "null context");
    if (!ctx->c.lists.valid) fail(EXAFMM_ERR_STATE, "no lists");
    if (m2l_pairs) *m2l_pairs = ctx->c.lists.nm2l;
    if (p2p_pairs) *p2p_pairs = ctx->c.lists.np2p;
  });
}

int exafmm_get_lists(exafmm_ctx *ctx, int which, int64_t *offsets, int64_t *sources) {
  return guard(ctx, [&] {
    if (!ctx || !offsets || !sources) fail(EXAFMM_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (!c.lists.valid) fail(EXAFMM_ERR_STATE, "no lists");
    XF_HIP(hipStreamSynchronize(c.s0));
    const int64_t nc = c.t.ncells;
    std::vector<int64_t> b2r, r2b, sb2r, sr2b;
    tree_ref_order(c, b2r, r2b);
    const Ctx &sc = c.src ? *c.src : c;
    if (&sc == &c) {
      sb2r = b2r;
    } else {
      tree_ref_order(const_cast<Ctx &>(sc), sb2r, sr2b);
    }
    std::vector<int> beg(nc), cnt(nc);
    const Lists &L = c.lists;
    XF_HIP(hipMemcpy(beg.data(), which == 0 ? L.m2l_beg.p : L.p2p_beg.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(cnt.data(), which == 0 ? L.m2l_cnt.p : L.p2p_cnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
    const int64_t tot = which == 0 ? L.nm2l : L.np2p;
    std::vector<int> src(std::max<int64_t>(tot, 1));
    if (which == 0) {
      std::vector<int2> e(std::max<int64_t>(tot, 1));
      if (tot) XF_HIP(hipMemcpy(e.data(), L.m2l.p, sizeof(int2) * tot, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < tot; i++) src[i] = e[i].x >> 5;
    } else if (tot) {
      XF_HIP(hipMemcpy(src.data(), L.p2p.p, sizeof(int) * tot, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < tot; i++) src[i] >>= 5;
    }
    int64_t o = 0;
    for (int64_t r = 0; r < nc; r++) {
      offsets[r] = o;
      const int64_t b = r2b[r];
      for (int k = 0; k < cnt[b]; k++) sources[o++] = sb2r[src[beg[b] + k]];
    }
    offsets[nc] = o;
  });
}

int exafmm_work_counts(exafmm_ctx *ctx, int64_t *out, int capacity) {
  return guard(ctx, [&] {
    if (!ctx || !out) fail(EXAFMM_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (!c.lists.valid) fail(EXAFMM_ERR_STATE, "no lists");
    XF_HIP(hipStreamSynchronize(c.s1));
    XF_HIP(hipStreamSynchronize(c.s0));
    const Ctx &sc = c.src ? *c.src : c;
    const int64_t nc = c.t.ncells, snc = sc.t.ncells;
    std::vector<int> beg(nc), cnt(nc), ccnt(nc), sccnt(snc), src(std::max<int64_t>(c.lists.np2p, 1));
    XF_HIP(hipMemcpy(beg.data(), c.lists.p2p_beg.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(cnt.data(), c.lists.p2p_cnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(ccnt.data(), c.t.ccnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
    XF_HIP(hipMemcpy(sccnt.data(), sc.t.ccnt.p, sizeof(int) * snc, hipMemcpyDeviceToHost));
    if (c.lists.np2p) XF_HIP(hipMemcpy(src.data(), c.lists.p2p.p, sizeof(int) * c.lists.np2p, hipMemcpyDeviceToHost));
    int64_t inter = 0, nj = 0;
    for (int64_t i = 0; i < nc; i++)
      for (int k = 0; k < cnt[i]; k++) {
        const int64_t s = sccnt[src[beg[i] + k] >> 5];
        inter += int64_t(ccnt[i]) * s;
        nj += s;
      }
    const int64_t v[8] = {c.lists.nm2l, c.lists.np2p, inter, nj, nc, c.t.nleaves, c.lists.nkeys, c.t.depth};
    for (int i = 0; i < std::min(capacity, 8); i++) out[i] = v[i];
  });
}

// Scaled real-form coefficients -> complex M_n^m = M~ R^n or L_j^k = L~ / R^(j+1),
// reference cell numbering.
static void export_coef(exafmm_ctx *ctx, const Buf<double> &B, bool locals, double *out) {
  Ctx &c = ctx->c;
  XF_HIP(hipStreamSynchronize(c.s1));
  XF_HIP(hipStreamSynchronize(c.s0));
  const int64_t nc = c.t.ncells;
  const int P = c.P, PP = P * P, NT = c.NT;
  std::vector<int64_t> b2r, r2b;
  tree_ref_order(c, b2r, r2b);
  std::vector<double> h(size_t(nc) * PP);
  std::vector<double4> cx(nc);
  XF_HIP(hipMemcpy(h.data(), B.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(cx.data(), c.t.cx.p, sizeof(double4) * nc, hipMemcpyDeviceToHost));
  for (int64_t b = 0; b < nc; b++) {
    double *o = out + 2 * NT * b2r[b];
    const double *v = h.data() + size_t(b) * PP;
    for (int n = 0; n < P; n++)
      for (int m = 0; m <= n; m++) {
        const double s = locals ? std::pow(cx[b].w, -double(n + 1)) : std::pow(cx[b].w, double(n));
        const int i = xfd::nms(n, m);
        o[2 * i] = v[xfd::rf(n, m, 0)] * s;
        o[2 * i + 1] = m ? v[xfd::rf(n, m, 1)] * s : 0.0;
      }
  }
}

int exafmm_get_multipoles(exafmm_ctx *ctx, double *out) {
  return guard(ctx, [&] {
    if (!ctx || !out) fail(EXAFMM_ERR_ARG, "null argument");
    if (!ctx->c.have_M) fail(EXAFMM_ERR_STATE, "no multipoles");
    export_coef(ctx, ctx->c.M, false, out);
  });
}

int exafmm_get_locals(exafmm_ctx *ctx, double *out) {
  return guard(ctx, [&] {
    if (!ctx || !out) fail(EXAFMM_ERR_ARG, "null argument");
    if (!ctx->c.have_L) fail(EXAFMM_ERR_STATE, "no locals");
    export_coef(ctx, ctx->c.L, true, out);
  });
}

int exafmm_order_model_train(exafmm_ctx *ctx, int nsamples, int epochs, double lr, uint64_t seed, double *loss) {
  return guard(ctx, [&] {
    if (!ctx || nsamples < 1 || epochs < 0 || !(lr > 0)) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    XF_HIP(hipStreamSynchronize(c.s1));
    const float l = order_train(c, nsamples, epochs, float(lr), seed);
    if (loss) *loss = l;
  });
}

int exafmm_set_order_model(exafmm_ctx *ctx, const float *weights, int count, double log10_tol) {
  return guard(ctx, [&] {
    if (!ctx || !weights || count != kMlpParams || !std::isfinite(log10_tol)) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    c.mlp_w.ensure(kMlpParams);
    XF_HIP(hipMemcpyAsync(c.mlp_w.p, weights, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c.s0));
    XF_HIP(hipStreamSynchronize(c.s0));
    c.order_log_tol = float(log10_tol);
    c.have_model = true;
  });
}

int exafmm_get_order_model(exafmm_ctx *ctx, float *weights, int count) {
  return guard(ctx, [&] {
    if (!ctx || !weights || count != kMlpParams) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    if (!c.have_model) fail(EXAFMM_ERR_STATE, "no order model");
    XF_HIP(hipMemcpyAsync(weights, c.mlp_w.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c.s0));
    XF_HIP(hipStreamSynchronize(c.s0));
  });
}

int exafmm_order_training_set(exafmm_ctx *ctx, float *X, float *T, int32_t *pairs, int64_t capacity, int64_t *n) {
  return guard(ctx, [&] {
    if (!ctx || !n) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    *n = c.train_n;
    if (!X || !T) return;
    if (capacity < c.train_n) fail(EXAFMM_ERR_CAPACITY, "training set larger than capacity");
    if (pairs) {
      std::vector<int64_t> b2r, r2b;
      tree_ref_order(c, b2r, r2b);
      for (int64_t i = 0; i < c.train_n; i++) {
        pairs[3 * i] = int32_t(b2r[c.train_pairs[3 * i]]);
        pairs[3 * i + 1] = int32_t(b2r[c.train_pairs[3 * i + 1]]);
        pairs[3 * i + 2] = c.train_pairs[3 * i + 2];
      }
    }
    XF_HIP(hipMemcpyAsync(X, c.train_X.p, sizeof(float) * c.train_n * kMlpIn, hipMemcpyDeviceToHost, c.s0));
    XF_HIP(hipMemcpyAsync(T, c.train_T.p, sizeof(float) * c.train_n * kMlpOut, hipMemcpyDeviceToHost, c.s0));
    XF_HIP(hipStreamSynchronize(c.s0));
  });
}

int exafmm_mlp_forward(exafmm_ctx *ctx, const float *weights, const float *X, int64_t n, float *Y) {
  return guard(ctx, [&] {
    if (!ctx || !weights || !X || !Y || n < 0) fail(EXAFMM_ERR_ARG, "bad argument");
    if (n) mlp_eval_host(ctx->c, weights, X, n, Y);
  });
}

int exafmm_mlp_gradient(exafmm_ctx *ctx, const float *weights, const float *X, const float *T, int64_t n,
                        float *grad, float *loss) {
  return guard(ctx, [&] {
    if (!ctx || !weights || !X || !T || !grad || !loss || n < 1) fail(EXAFMM_ERR_ARG, "bad argument");
    mlp_grad_host(ctx->c, weights, X, T, n, grad, loss);
  });
}

int exafmm_order_stats(exafmm_ctx *ctx, int64_t *hist, int capacity) {
  return guard(ctx, [&] {
    if (!ctx || !hist || capacity < 1) fail(EXAFMM_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    XF_HIP(hipStreamSynchronize(c.s1));
    XF_HIP(hipStreamSynchronize(c.s0));
    for (int i = 0; i < capacity; i++) hist[i] = 0;
    const int nt = c.m2l_ntiles_last, U = c.m2l_totalU;
    if (nt == 0 || U == 0) return;
    std::vector<int> tU(nt);
    XF_HIP(hipMemcpy(tU.data(), c.tiles.p + 3 * c.m2l_tstride, sizeof(int) * nt, hipMemcpyDeviceToHost));
    std::vector<unsigned char> tp(U, (unsigned char)c.P);
    if (c.learned) XF_HIP(hipMemcpy(tp.data(), c.tile_p.p, U, hipMemcpyDeviceToHost));
    std::vector<int> cnt(nt);
    XF_HIP(hipMemcpy(cnt.data(), c.tiles.p + c.m2l_tstride, sizeof(int) * nt, hipMemcpyDeviceToHost));
    int64_t u = 0;
    for (int t = 0; t < nt; t++)
      for (int k = 0; k < tU[t]; k++, u++)
        if (tp[u] < capacity) hist[tp[u]] += cnt[t];
  });
}

int exafmm_p2p(exafmm_ctx *ctx, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns) {
  return guard(ctx, [&] {
    if (!ctx || !t || !s || nt < 0 || ns < 0) fail(EXAFMM_ERR_ARG, "bad argument");
    op_p2p(ctx->c, t, nt, s, ns);
  });
}
int exafmm_p2m(exafmm_ctx *ctx, const double center[3], const exafmm_body *b, int64_t nb, double *M) {
  return guard(ctx, [&] {
    if (!ctx || !center || !b || !M || nb < 0) fail(EXAFMM_ERR_ARG, "bad argument");
    op_p2m(ctx->c, center, b, nb, M);
  });
}
int exafmm_m2m(exafmm_ctx *ctx, const double cp[3], const double cc[3], const double *Mc, double *Mp) {
  return guard(ctx, [&] {
    if (!ctx || !cp || !cc || !Mc || !Mp) fail(EXAFMM_ERR_ARG, "bad argument");
    op_m2m(ctx->c, cp, cc, Mc, Mp);
  });
}
int exafmm_m2l(exafmm_ctx *ctx, const double Xi[3], const double Xj[3], const double *Mj, double *Li) {
  return guard(ctx, [&] {
    if (!ctx || !Xi || !Xj || !Mj || !Li) fail(EXAFMM_ERR_ARG, "bad argument");
    op_m2l(ctx->c, Xi, Xj, Mj, Li);
  });
}
int exafmm_l2l(exafmm_ctx *ctx, const double cp[3], const double *Lp, const double cc[3], double *Lc) {
  return guard(ctx, [&] {
    if (!ctx || !cp || !cc || !Lp || !Lc) fail(EXAFMM_ERR_ARG, "bad argument");
    op_l2l(ctx->c, cp, Lp, cc, Lc);
  });
}
int exafmm_l2p(exafmm_ctx *ctx, const double center[3], const double *L, exafmm_body *b, int64_t nb) {
  return guard(ctx, [&] {
    if (!ctx || !center || !L || !b || nb < 0) fail(EXAFMM_ERR_ARG, "bad argument");
    op_l2p(ctx->c, center, L, b, nb);
  });
}

}  // extern "C"

// ---------------------------------------------------------------- graphs
// Many independent problems re-evaluated by one hipGraph launch.  Each
// context was evaluated once (tree, lists, keyed-M2L tiles and operators,
// autotuned GEMM shape); the graph records, per problem, the
// charge-dependent passes on one of several capture streams (independent
// problems form parallel branches of the graph):
//   charges of the tree-ordered bodies -> P2M inputs, p = F = 0, L = 0
//   P2M, M2M (per level), P2P, keyed M2L (cast, GEMM, slice sum),
//   L2L (per level), L2P
// No memset nodes (fill kernels instead), no events, no host waits.
namespace xf {
namespace {

__global__ void __launch_bounds__(256) k_replay_bodies(int64_t n, exafmm_body *__restrict__ b,
                                                       double4 *__restrict__ xs, float4 *__restrict__ xq) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double q = b[i].q;
  xs[i].w = q;
  xq[i].w = float(q);
  b[i].p = 0.0;
  b[i].F[0] = 0.0;
  b[i].F[1] = 0.0;
  b[i].F[2] = 0.0;
}

__global__ void __launch_bounds__(256) k_zero(int64_t n, double *__restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Grouped prologue: per problem, its bodies (charges into the P2M / P2P
// inputs, p = F = 0) and then its locals (= 0).
__global__ void __launch_bounds__(256) k_replay_grouped(const GroupTree *__restrict__ g, const int *__restrict__ start,
                                                        int nprob) {
  const int b = blockIdx.x, i = group_of(start, nprob, b), lb = b - start[i];
  const GroupTree &t = g[i];
  const int nbb = int((t.n + 255) / 256);
  if (lb < nbb) {
    const int64_t k = int64_t(lb) * 256 + threadIdx.x;
    if (k >= t.n) return;
    exafmm_body *bo = t.bodies + k;
    const double q = bo->q;
    t.xs_w[k].w = q;
    t.xq_w[k].w = float(q);
    bo->p = 0.0;
    bo->F[0] = 0.0;
    bo->F[1] = 0.0;
    bo->F[2] = 0.0;
  } else {
    const int64_t k = int64_t(lb - nbb) * 256 + threadIdx.x;
    if (k < t.ncells * t.P * t.P) t.L[k] = 0.0;
  }
}

// The charge-dependent passes of one problem, all on stream s.
void replay(Ctx &c, hipStream_t s) {
  const hipStream_t a = c.s0, b = c.s1;
  c.s0 = c.s1 = s;
  try {
    const int64_t n = c.t.n, nl = int64_t(c.t.ncells) * c.P * c.P;
    hipLaunchKernelGGL(k_replay_bodies, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, c.bodies, c.t.xs.p, c.t.xq.p);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_zero, dim3(ceil_div(nl, 256)), dim3(256), 0, s, nl, c.L.p);
    XF_LAUNCH();
    upward(c, s);
    p2p_apply(c, c, s);
    m2l_apply(c, c);
    l2l_pass(c);
    l2p_pass(c);
  } catch (...) {
    c.s0 = a;
    c.s1 = b;
    c.capturing = false;
    throw;
  }
  c.s0 = a;
  c.s1 = b;
}

}  // namespace
}  // namespace xf

struct exafmm_graph {
  // grouped mode: one launch per stage over all problems
  struct Stage {
    int off = 0, n = 0, blocks = 0;  // descriptor / start offsets, entries, workgroups
  };
  bool grouped = false;
  xf::Buf<xf::GroupTree> gt;
  xf::Buf<xf::GroupLevel> lv;
  xf::Buf<int> starts;
  Stage pro, p2m, p2p, l2p;
  std::vector<Stage> m2m_levels, l2l_levels;  // M2M bottom-up, L2L top-down
  std::vector<std::unique_ptr<xf::GemmGroupPlan>> gemm;  // one per order
  std::vector<int> level_soff_m2m, level_soff_l2l;       // start-array offsets of the level stages

  std::vector<exafmm_ctx *> ctxs;
  int device = 0;
  hipStream_t main = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;  // [0] fork, [1..] joins
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void graph_release(exafmm_graph *g) {
  if (!g) return;
  hipSetDevice(g->device);
  if (g->main) hipStreamSynchronize(g->main);
  for (hipStream_t s : g->streams) hipStreamSynchronize(s);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  for (hipEvent_t e : g->events)
    if (e) hipEventDestroy(e);
  for (hipStream_t s : g->streams)
    if (s) hipStreamDestroy(s);
  if (g->main) hipStreamDestroy(g->main);
  for (exafmm_ctx *x : g->ctxs) {
    x->c.graph_locks--;
    x->c.capturing = false;
    if (!x->c.graph_locks) x->c.m2l_ready = false;
  }
  delete g;
}

// Descriptor tables of a grouped graph (host-built, copied once).
void grouped_tables(exafmm_graph *g) {
  const int n = int(g->ctxs.size());
  std::vector<GroupTree> gt(n);
  std::vector<int> st;
  std::vector<GroupLevel> lv;
  int maxd = 0;
  for (int i = 0; i < n; i++) {
    const Ctx &c = g->ctxs[i]->c;
    GroupTree &t = gt[i];
    t.leaves = c.t.leaves.p;
    t.nleaves = c.t.misc.p + 17;
    t.cbeg = c.t.cbeg.p;
    t.ccnt = c.t.ccnt.p;
    t.cchild = c.t.cchild.p;
    t.cnch = c.t.cnch.p;
    t.cx = c.t.cx.p;
    t.xs = c.t.xs.p;
    t.xq = c.t.xq.p;
    t.xs_w = c.t.xs.p;
    t.xq_w = c.t.xq.p;
    t.lbeg = c.lists.p2p_beg.p;
    t.lcnt = c.lists.p2p_cnt.p;
    t.list = c.lists.p2p.p;
    t.shift_T = c.shift_T.p;
    t.M = c.M.p;
    t.L = c.L.p;
    t.bodies = c.bodies;
    t.n = c.t.n;
    t.ncells = c.t.ncells;
    t.P = c.P;
    maxd = std::max(maxd, c.t.depth);
  }
  // per-problem stages: starts = prefix of each problem's workgroups
  auto per_problem = [&](exafmm_graph::Stage &sg, auto blocks_of) {
    sg.off = int(st.size());
    sg.n = n;
    int64_t acc = 0;
    for (int i = 0; i < n; i++) {
      st.push_back(int(acc));
      acc += blocks_of(g->ctxs[i]->c);
    }
    st.push_back(int(acc));
    if (acc > (int64_t(1) << 31) - 1) fail(EXAFMM_ERR_CAPACITY, "batch too large");
    sg.blocks = int(acc);
  };
  per_problem(g->pro, [](const Ctx &c) {
    return int64_t(ceil_div(c.t.n, 256)) + ceil_div(c.t.ncells * c.P * c.P, 256);
  });
  per_problem(g->p2m, [](const Ctx &c) { return int64_t(p2m_group_blocks(c.t.nleaves)); });
  per_problem(g->p2p, [](const Ctx &c) { return int64_t(p2p_group_blocks(c.t.nleaves)); });
  per_problem(g->l2p, [](const Ctx &c) { return int64_t(c.t.nleaves); });
  // per-level stages: the problems that have parents at level l
  auto per_level = [&](int l, int mult, int (*rows)(int)) {
    exafmm_graph::Stage sg;
    sg.off = int(lv.size());
    const int soff = int(st.size());
    int64_t acc = 0;
    for (int i = 0; i < n; i++) {
      const Ctx &c = g->ctxs[i]->c;
      if (l >= c.t.depth) continue;
      const int cnt = c.t.lstart[l + 1] - c.t.lstart[l];
      if (cnt <= 0) continue;
      GroupLevel e{i, c.t.lstart[l], cnt, ceil_div(cnt, 16), rows(c.P)};
      lv.push_back(e);
      st.push_back(int(acc));
      acc += int64_t(e.nbx) * e.nby * mult;
    }
    st.push_back(int(acc));
    sg.n = int(lv.size()) - sg.off;
    sg.blocks = int(acc);
    // the level's start array sits at soff; keep it beside the descriptors
    return std::make_pair(sg, soff);
  };
  std::vector<std::pair<exafmm_graph::Stage, int>> m2m, l2l;
  for (int l = maxd - 1; l >= 0; l--) m2m.push_back(per_level(l, 1, m2m_group_rows));
  for (int l = 0; l < maxd; l++) l2l.push_back(per_level(l, 8, shift_group_rows));
  g->gt.ensure(n);
  g->lv.ensure(std::max<size_t>(lv.size(), 1));
  g->starts.ensure(st.size());
  XF_HIP(hipMemcpy(g->gt.p, gt.data(), sizeof(GroupTree) * n, hipMemcpyHostToDevice));
  if (!lv.empty()) XF_HIP(hipMemcpy(g->lv.p, lv.data(), sizeof(GroupLevel) * lv.size(), hipMemcpyHostToDevice));
  XF_HIP(hipMemcpy(g->starts.p, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
  // level stages: (descriptor offset, entries, workgroups), start-array
  // offsets in parallel lists
  g->m2m_levels.clear();
  g->l2l_levels.clear();
  g->level_soff_m2m.clear();
  g->level_soff_l2l.clear();
  for (auto &pr : m2m) {
    g->m2m_levels.push_back(pr.first);
    g->level_soff_m2m.push_back(pr.second);
  }
  for (auto &pr : l2l) {
    g->l2l_levels.push_back(pr.first);
    g->level_soff_l2l.push_back(pr.second);
  }
  // keyed GEMM: one plan per order
  std::map<int, std::vector<Ctx *>> byP;
  for (int i = 0; i < n; i++) byP[g->ctxs[i]->c.P].push_back(&g->ctxs[i]->c);
  g->gemm.clear();
  for (auto &kv : byP) {
    g->gemm.emplace_back(new GemmGroupPlan);
    m2l_group_setup(kv.second, *g->gemm.back(), g->main, kv.second[0]->cfg.autotune != 0);
    XF_HIP(hipStreamSynchronize(g->main));
  }
}

// All charge-dependent passes of every problem: one launch per stage (per
// level for M2M and L2L, per order for the keyed GEMM); P2P on the second
// stream beside the upward pass and M2L, joined before L2P.
void grouped_record(exafmm_graph *g, hipStream_t sa, hipStream_t sb, hipEvent_t fork, hipEvent_t join) {
  const int n = int(g->ctxs.size());
  const int *st = g->starts.p;
  hipLaunchKernelGGL(k_replay_grouped, dim3(g->pro.blocks), dim3(256), 0, sa, g->gt.p, st + g->pro.off, n);
  XF_LAUNCH();
  XF_HIP(hipEventRecord(fork, sa));
  XF_HIP(hipStreamWaitEvent(sb, fork, 0));
  if (g->p2p.blocks) p2p_grouped(g->gt.p, st + g->p2p.off, n, g->p2p.blocks, sb);
  XF_HIP(hipEventRecord(join, sb));
  if (g->p2m.blocks) p2m_grouped(g->gt.p, st + g->p2m.off, n, g->p2m.blocks, sa);
  for (size_t k = 0; k < g->m2m_levels.size(); k++) {
    const auto &sg = g->m2m_levels[k];
    if (sg.blocks) m2m_grouped(g->gt.p, g->lv.p + sg.off, st + g->level_soff_m2m[k], sg.n, sg.blocks, sa);
  }
  for (auto &plan : g->gemm) m2l_group_launch(*plan, sa);
  for (size_t k = 0; k < g->l2l_levels.size(); k++) {
    const auto &sg = g->l2l_levels[k];
    if (sg.blocks) l2l_grouped(g->gt.p, g->lv.p + sg.off, st + g->level_soff_l2l[k], sg.n, sg.blocks, sa);
  }
  XF_HIP(hipStreamWaitEvent(sa, join, 0));
  if (g->l2p.blocks) l2p_grouped(g->gt.p, st + g->l2p.off, n, g->l2p.blocks, sa);
}

}  // namespace

extern "C" {

int exafmm_graph_create(exafmm_ctx *const *ctxs, int count, int nstreams, exafmm_graph **out) {
  exafmm_graph *g = nullptr;
  bool capturing = false;
  hipStream_t cap = nullptr;
  const int rc = guard(nullptr, [&] {
    if (!ctxs || count < 1 || !out) fail(EXAFMM_ERR_ARG, "bad argument");
    *out = nullptr;
    const int dev = ctxs[0] ? ctxs[0]->c.cfg.device : 0;
    for (int i = 0; i < count; i++) {
      if (!ctxs[i]) fail(EXAFMM_ERR_ARG, "null context");
      const Ctx &c = ctxs[i]->c;
      if (c.cfg.device != dev) fail(EXAFMM_ERR_ARG, "contexts on different devices");
      if (!c.have_L || c.src != &c) fail(EXAFMM_ERR_STATE, "each context must have been evaluated once");
      if (c.forest_n > 0) fail(EXAFMM_ERR_UNSUPPORTED, "a forest context cannot be recorded in a graph");
      if (c.host_mode) fail(EXAFMM_ERR_UNSUPPORTED, "graphs need device-resident bodies (on_device = 1)");
      if (c.cfg.images > 0) fail(EXAFMM_ERR_UNSUPPORTED, "periodic contexts cannot be recorded in a graph");
      if (c.graph_locks) fail(EXAFMM_ERR_STATE, "a context is already held by a graph");
      for (int j = 0; j < i; j++)
        if (ctxs[j] == ctxs[i]) fail(EXAFMM_ERR_ARG, "a context appears twice");
    }
    XF_HIP(hipSetDevice(dev));
    g = new exafmm_graph;
    g->device = dev;
    const int ns = std::max(1, std::min(nstreams, 16));
    XF_HIP(hipStreamCreateWithFlags(&g->main, hipStreamNonBlocking));
    g->streams.assign(ns, nullptr);
    g->events.assign(ns + 1, nullptr);
    for (int k = 0; k < ns; k++) XF_HIP(hipStreamCreateWithFlags(&g->streams[k], hipStreamNonBlocking));
    for (auto &e : g->events) XF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (int i = 0; i < count; i++) {
      Ctx &c = ctxs[i]->c;
      XF_HIP(hipStreamSynchronize(c.s1));
      XF_HIP(hipStreamSynchronize(c.s0));
      g->ctxs.push_back(ctxs[i]);
      c.graph_locks++;
      if (c.cfg.m2l_precision != EXAFMM_PREC_FP64 && c.lists.nkeys > 0) {
        m2l_prepare(c, c.s0);  // tiles kept for the graph's lifetime
        XF_HIP(hipStreamSynchronize(c.s0));
      }
    }
    // one unrecorded pass: sizes every buffer, tunes any GEMM shape not yet
    // tuned, and leaves the results of the current charges
    for (int i = 0; i < count; i++) replay(ctxs[i]->c, g->streams[i % ns]);
    for (hipStream_t s : g->streams) XF_HIP(hipStreamSynchronize(s));
    cap = g->main;
    XF_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    capturing = true;
    XF_HIP(hipEventRecord(g->events[0], cap));
    for (int k = 0; k < ns; k++) XF_HIP(hipStreamWaitEvent(g->streams[k], g->events[0], 0));
    for (int i = 0; i < count; i++) {
      Ctx &c = ctxs[i]->c;
      c.capturing = true;
      replay(c, g->streams[i % ns]);
      c.capturing = false;
    }
    for (int k = 0; k < ns; k++) {
      XF_HIP(hipEventRecord(g->events[k + 1], g->streams[k]));
      XF_HIP(hipStreamWaitEvent(cap, g->events[k + 1], 0));
    }
    capturing = false;
    XF_HIP(hipStreamEndCapture(cap, &g->graph));
    XF_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g;
  });
  if (rc != EXAFMM_OK) {
    if (capturing) {
      hipGraph_t junk = nullptr;
      if (hipStreamEndCapture(cap, &junk) == hipSuccess && junk) hipGraphDestroy(junk);
      (void)hipGetLastError();
    }
    graph_release(g);
  }
  return rc;
}

int exafmm_graph_create_grouped(exafmm_ctx *const *ctxs, int count, exafmm_graph **out) {
  exafmm_graph *g = nullptr;
  bool capturing = false;
  hipStream_t cap = nullptr;
  const int rc = guard(nullptr, [&] {
    if (!ctxs || count < 1 || !out) fail(EXAFMM_ERR_ARG, "bad argument");
    *out = nullptr;
    const int dev = ctxs[0] ? ctxs[0]->c.cfg.device : 0;
    for (int i = 0; i < count; i++) {
      if (!ctxs[i]) fail(EXAFMM_ERR_ARG, "null context");
      const Ctx &c = ctxs[i]->c;
      if (c.cfg.device != dev) fail(EXAFMM_ERR_ARG, "contexts on different devices");
      if (!c.have_L || c.src != &c) fail(EXAFMM_ERR_STATE, "each context must have been evaluated once");
      if (c.forest_n > 0) fail(EXAFMM_ERR_UNSUPPORTED, "a forest context cannot be recorded in a graph");
      if (c.host_mode) fail(EXAFMM_ERR_UNSUPPORTED, "graphs need device-resident bodies (on_device = 1)");
      if (c.cfg.images > 0) fail(EXAFMM_ERR_UNSUPPORTED, "periodic contexts cannot be recorded in a graph");
      if (c.cfg.m2l_precision != EXAFMM_PREC_BF16X3 || c.cfg.p2p_precision != EXAFMM_PREC_FP32)
        fail(EXAFMM_ERR_UNSUPPORTED, "grouped graphs need BF16x3 M2L and FP32 P2P");
      if (c.cfg.order_policy == EXAFMM_ORDER_LEARNED && c.have_model)
        fail(EXAFMM_ERR_UNSUPPORTED, "grouped graphs run fixed orders");
      if (c.lists.nkeys < 1) fail(EXAFMM_ERR_UNSUPPORTED, "a problem without M2L work");
      if (c.graph_locks) fail(EXAFMM_ERR_STATE, "a context is already held by a graph");
      for (int j = 0; j < i; j++)
        if (ctxs[j] == ctxs[i]) fail(EXAFMM_ERR_ARG, "a context appears twice");
    }
    XF_HIP(hipSetDevice(dev));
    g = new exafmm_graph;
    g->device = dev;
    g->grouped = true;
    XF_HIP(hipStreamCreateWithFlags(&g->main, hipStreamNonBlocking));
    g->streams.assign(1, nullptr);
    g->events.assign(2, nullptr);
    XF_HIP(hipStreamCreateWithFlags(&g->streams[0], hipStreamNonBlocking));
    for (auto &e : g->events) XF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (int i = 0; i < count; i++) {
      Ctx &c = ctxs[i]->c;
      XF_HIP(hipStreamSynchronize(c.s1));
      XF_HIP(hipStreamSynchronize(c.s0));
      g->ctxs.push_back(ctxs[i]);
      c.graph_locks++;
      m2l_prepare(c, c.s0);  // tiles kept for the graph's lifetime
      XF_HIP(hipStreamSynchronize(c.s0));
      replay(c, c.s0);  // sizes every buffer (cast multipoles, shift operators)
      XF_HIP(hipStreamSynchronize(c.s0));
    }
    grouped_tables(g);
    // one unrecorded grouped pass, then the recording
    grouped_record(g, g->main, g->streams[0], g->events[0], g->events[1]);
    XF_HIP(hipStreamSynchronize(g->main));
    cap = g->main;
    XF_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    capturing = true;
    grouped_record(g, g->main, g->streams[0], g->events[0], g->events[1]);
    capturing = false;
    XF_HIP(hipStreamEndCapture(cap, &g->graph));
    XF_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g;
  });
  if (rc != EXAFMM_OK) {
    if (capturing) {
      hipGraph_t junk = nullptr;
      if (hipStreamEndCapture(cap, &junk) == hipSuccess && junk) hipGraphDestroy(junk);
      (void)hipGetLastError();
    }
    graph_release(g);
  }
  return rc;
}

int exafmm_graph_launch(exafmm_graph *g) {
  return guard(nullptr, [&] {
    if (!g || !g->exec) fail(EXAFMM_ERR_ARG, "null graph");
    XF_HIP(hipSetDevice(g->device));
    XF_HIP(hipGraphLaunch(g->exec, g->main));
  });
}

int exafmm_graph_synchronize(exafmm_graph *g) {
  return guard(nullptr, [&] {
    if (!g) fail(EXAFMM_ERR_ARG, "null graph");
    XF_HIP(hipSetDevice(g->device));
    XF_HIP(hipStreamSynchronize(g->main));
  });
}

int exafmm_graph_nodes(exafmm_graph *g, int64_t *nodes) {
  return guard(nullptr, [&] {
    if (!g || !nodes || !g->graph) fail(EXAFMM_ERR_ARG, "bad argument");
    size_t n = 0;
    XF_HIP(hipGraphGetNodes(g->graph, nullptr, &n));
    *nodes = int64_t(n);
  });
}

int exafmm_graph_gemm_shapes(exafmm_graph *g, int32_t *shapes, int capacity, int *count) {
  return guard(nullptr, [&] {
    if (!g || !count || (capacity > 0 && !shapes)) fail(EXAFMM_ERR_ARG, "bad argument");
    int k = 0;
    for (auto &pl : g->gemm) {
      if (!pl->nprob) continue;
      if (k < capacity) {
        shapes[3 * k] = pl->P;
        shapes[3 * k + 1] = pl->rtb * 16;
        shapes[3 * k + 2] = pl->kr;
      }
      k++;
    }
    *count = k;
  });
}

int exafmm_graph_destroy(exafmm_graph *g) {
  if (!g) return EXAFMM_ERR_ARG;
  graph_release(g);
  return EXAFMM_OK;
}


namespace xf {
// Device-resident problems of a forest: gathered into the forest's array and
// scattered back by one launch each (one thread per body and 8-byte word;
// a body's problem by binary search over the offsets).
__global__ void __launch_bounds__(256) k_forest_copy(const ForestPart *__restrict__ parts, int np, int64_t n,
                                                     exafmm_body *__restrict__ fb, int to_forest) {
  // the workgroup's 32 bodies lie in one problem or a few consecutive ones:
  // one search for the first body, then each thread steps forward from it
  __shared__ int first;
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (threadIdx.x == 0) {
    const int64_t i0 = (int64_t(blockIdx.x) * 256) >> 3;
    int lo = 0, hi = np - 1;  // last problem with off <= i0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (parts[mid].off <= i0)
        lo = mid;
      else
        hi = mid - 1;
    }
    first = lo;
  }
  __syncthreads();
  const int64_t i = g >> 3;
  const int w = int(g & 7);
  if (i >= n) return;
  int lo = first;
  while (lo + 1 < np && parts[lo + 1].off <= i) lo++;
  const ForestPart q = parts[lo];
  if (!q.src) return;  // a host problem: copied by the runtime
  double *f = reinterpret_cast<double *>(fb + i) + w;
  double *o = reinterpret_cast<double *>(q.src + (i - q.off)) + w;
  if (to_forest)
    *f = *o;
  else
    *o = *f;
}
}  // namespace xf

// EXAFMM_BATCH_FOREST: the problems of each order, in their given order and
// at most kForestMax per forest, are gathered into one device array and
// evaluated as one forest (exafmm_evaluate_forest: one tree build, one
// traversal and one launch per pass over all of them), then copied back.
static void batch_forests(exafmm_ctx *ctx, const exafmm_problem *problems, int count) {
  using namespace xf;
  std::vector<std::vector<int>> groups;  // problem indices, by order, chunks of kForestMax
  {
    std::vector<int> idx(count);
    for (int i = 0; i < count; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return problems[a].P < problems[b].P; });
    for (int i : idx) {
      if (groups.empty() || problems[groups.back()[0]].P != problems[i].P ||
          int(groups.back().size()) == kForestMax)
        groups.emplace_back();
      groups.back().push_back(i);
    }
  }
  while (ctx->forests.size() < groups.size()) {
    exafmm_config cfg = ctx->c.cfg;
    exafmm_ctx *sub = nullptr;
    if (exafmm_create(&cfg, &sub) != EXAFMM_OK) fail(EXAFMM_ERR_ALLOC, g_error);
    ctx->forests.push_back(sub);
  }
  while (ctx->forest_bodies.size() < ctx->forests.size()) {
    ctx->forest_bodies.push_back(std::make_unique<Buf<exafmm_body>>());
    ctx->forest_parts.push_back(std::make_unique<Buf<ForestPart>>());
    ctx->forest_parts_host.emplace_back();
  }
  for (size_t gi = 0; gi < groups.size(); gi++) {
    const std::vector<int> &g = groups[gi];
    const int P = problems[g[0]].P;
    if (ctx->forests[gi]->c.P != P) {  // another order in this slot: a new context
      exafmm_destroy(ctx->forests[gi]);
      exafmm_config cfg = ctx->c.cfg;
      cfg.P = P;
      ctx->forests[gi] = nullptr;
      if (exafmm_create(&cfg, &ctx->forests[gi]) != EXAFMM_OK) fail(EXAFMM_ERR_ALLOC, g_error);
    }
  }
  for (size_t gi = 0; gi < groups.size(); gi++) {
    const std::vector<int> &g = groups[gi];
    exafmm_ctx *fx = ctx->forests[gi];
    Ctx &fc = fx->c;
    std::vector<int64_t> off(g.size() + 1, 0);
    for (size_t k = 0; k < g.size(); k++) off[k + 1] = off[k] + problems[g[k]].n;
    Buf<exafmm_body> &buf = *ctx->forest_bodies[gi];
    buf.ensure(off.back());
    // device problems in one gather launch, host problems by copies
    std::vector<ForestPart> &ph = ctx->forest_parts_host[gi];
    ph.resize(g.size());
    for (size_t k = 0; k < g.size(); k++) {
      const exafmm_problem &q = problems[g[k]];
      ph[k] = ForestPart{q.on_device ? q.bodies : nullptr, off[k]};
      if (!q.on_device)
        XF_HIP(hipMemcpyAsync(buf.p + off[k], q.bodies, sizeof(exafmm_body) * q.n, hipMemcpyHostToDevice, fc.s0));
    }
    Buf<ForestPart> &pd = *ctx->forest_parts[gi];
    pd.ensure(g.size());
    XF_HIP(hipMemcpyAsync(pd.p, ph.data(), sizeof(ForestPart) * g.size(), hipMemcpyHostToDevice, fc.s0));
    const int nblk = ceil_div(off.back() * 8, 256);
    hipLaunchKernelGGL(k_forest_copy, dim3(nblk), dim3(256), 0, fc.s0, pd.p, int(g.size()), off.back(), buf.p, 1);
    XF_LAUNCH();
    if (exafmm_evaluate_forest(fx, buf.p, off.data(), int(g.size()), 1) != EXAFMM_OK)
      fail(EXAFMM_ERR_HIP, std::string("forest of order ") + std::to_string(fc.P) + ": " + fc.last_error);
    hipLaunchKernelGGL(k_forest_copy, dim3(nblk), dim3(256), 0, fc.s0, pd.p, int(g.size()), off.back(), buf.p, 0);
    XF_LAUNCH();
    for (size_t k = 0; k < g.size(); k++) {
      const exafmm_problem &q = problems[g[k]];
      if (!q.on_device)
        XF_HIP(hipMemcpyAsync(q.bodies, buf.p + off[k], sizeof(exafmm_body) * q.n, hipMemcpyDeviceToHost, fc.s0));
    }
  }
  for (exafmm_ctx *f : ctx->forests) {
    XF_HIP(hipStreamSynchronize(f->c.s1));
    XF_HIP(hipStreamSynchronize(f->c.s0));
  }
}

int exafmm_evaluate_batch(exafmm_ctx *ctx, const exafmm_problem *problems, int count, int flags) {
  return guard(ctx, [&] {
    if (!ctx || !problems || count < 1) fail(EXAFMM_ERR_ARG, "bad argument");
    const bool reuse = (flags & EXAFMM_BATCH_REUSE_GEOMETRY) != 0;
    std::vector<std::pair<const void *, int64_t>> sig(count);
    for (int i = 0; i < count; i++) {
      const exafmm_problem &q = problems[i];
      if (!q.bodies || q.n < 1) fail(EXAFMM_ERR_ARG, "a problem without bodies");
      if (q.P < 2 || q.P > kMaxP) fail(EXAFMM_ERR_ARG, "P must be in [2, 16]");
      sig[i] = {q.bodies, (q.n << 8) | (int64_t(q.P) << 2) | (q.on_device ? 1 : 0)};
    }
    if (flags & EXAFMM_BATCH_FOREST) {
      if (reuse) fail(EXAFMM_ERR_ARG, "EXAFMM_BATCH_FOREST and EXAFMM_BATCH_REUSE_GEOMETRY exclude each other");
      batch_forests(ctx, problems, count);
      return;
    }
    if (reuse && ctx->batch_graph && sig == ctx->batch_sig) {  // same problems, positions kept
      XF_HIP(hipSetDevice(ctx->c.cfg.device));
      if (exafmm_graph_launch(ctx->batch_graph) != EXAFMM_OK) fail(EXAFMM_ERR_HIP, g_error);
      if (exafmm_graph_synchronize(ctx->batch_graph) != EXAFMM_OK) fail(EXAFMM_ERR_HIP, g_error);
      return;
    }
    if (ctx->batch_graph) exafmm_graph_destroy(ctx->batch_graph);
    ctx->batch_graph = nullptr;
    ctx->batch_sig.clear();
    // one sub-context per slot, with this context's configuration and the
    // problem's order
    if (int(ctx->batch.size()) > count) {
      for (size_t i = count; i < ctx->batch.size(); i++) exafmm_destroy(ctx->batch[i]);
      ctx->batch.resize(count);
    }
    for (int i = 0; i < count; i++) {
      if (i < int(ctx->batch.size()) && ctx->batch[i]->c.P == problems[i].P) continue;
      exafmm_config cfg = ctx->c.cfg;
      cfg.P = problems[i].P;
      exafmm_ctx *sub = nullptr;
      if (exafmm_create(&cfg, &sub) != EXAFMM_OK) fail(EXAFMM_ERR_ALLOC, g_error);
      if (i < int(ctx->batch.size())) {
        exafmm_destroy(ctx->batch[i]);
        ctx->batch[i] = sub;
      } else {
        ctx->batch.push_back(sub);
      }
    }
    for (int i = 0; i < count; i++) {
      const exafmm_problem &q = problems[i];
      if (exafmm_evaluate(ctx->batch[i], q.bodies, q.n, q.on_device) != EXAFMM_OK)
        fail(EXAFMM_ERR_HIP, std::string("problem ") + std::to_string(i) + ": " + ctx->batch[i]->c.last_error);
    }
    for (int i = 0; i < count; i++) {
      XF_HIP(hipStreamSynchronize(ctx->batch[i]->c.s1));
      XF_HIP(hipStreamSynchronize(ctx->batch[i]->c.s0));
    }
    if (reuse) {
      for (int i = 0; i < count; i++)
        if (!problems[i].on_device) fail(EXAFMM_ERR_UNSUPPORTED, "reusing geometry needs device-resident bodies");
      exafmm_graph *g = nullptr;
      if (exafmm_graph_create_grouped(ctx->batch.data(), count, &g) != EXAFMM_OK &&
          exafmm_graph_create(ctx->batch.data(), count, 8, &g) != EXAFMM_OK)
        fail(EXAFMM_ERR_HIP, g_error);
      ctx->batch_graph = g;
      ctx->batch_sig = sig;
    }
  });
}

}  // extern "C"
