// ref_shim3d.cpp -- C entry points onto the reference's own 3d/ headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (-I<reference>/3d, lazy traversal), with
// output only into oracle/_ref/.  It is used to validate the oracle, to
// write the golden fixtures under tests/golden/, and as bench.py's
// cpu_baseline ("reference" kind).  Nothing here restates the algorithm:
// every result comes from the reference's buildTree, kernels and passes.
#define EXAFMM_LAZY 1
#include "build_tree.h"
#include "kernel.h"
#include "traverse_lazy.h"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#define API extern "C" __attribute__((visibility("default")))

using namespace exafmm;

namespace {
Bodies g_bodies;
Cells g_cells;
Bodies g_jbodies;  // source set of the two-tree evaluation
Cells g_jcells;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void load(Bodies &dst, const Body *src, int64_t n) {
  dst.resize(n);
  std::memcpy(dst.data(), src, sizeof(Body) * n);
}

void set_coef(std::vector<complex_t> &v, const double *a) {
  v.resize(NTERM);
  for (int i = 0; i < NTERM; i++) v[i] = complex_t(a[2 * i], a[2 * i + 1]);
}
void get_coef(const std::vector<complex_t> &v, double *a) {
  for (int i = 0; i < NTERM; i++) {
    a[2 * i] = v[i].real();
    a[2 * i + 1] = v[i].imag();
  }
}
}  // namespace

static_assert(sizeof(Body) == 64, "Body layout");

API void ref3d_set(int p, int nc, double th) {
  P = p;
  ncrit = nc;
  theta = th;
  initKernel();
}

// buildTree on a copy; bodies (permuted X,q) written back; returns #cells.
API int64_t ref3d_build_tree(Body *b, int64_t n) {
  load(g_bodies, b, n);
  g_cells = buildTree(g_bodies);
  std::memcpy(b, g_bodies.data(), sizeof(Body) * n);
  return int64_t(g_cells.size());
}

// Flattened cell table of the last tree: nchild, nbody, child, body, X, R.
API void ref3d_get_cells(int32_t *nchild, int32_t *nbody, int64_t *child, int64_t *body, double *XR) {
  for (size_t i = 0; i < g_cells.size(); i++) {
    const Cell &c = g_cells[i];
    nchild[i] = c.NCHILD;
    nbody[i] = c.NBODY;
    child[i] = c.NCHILD ? int64_t(c.CHILD - &g_cells[0]) : -1;
    body[i] = int64_t(c.BODY - &g_bodies[0]);
    XR[4 * i + 0] = c.X[0];
    XR[4 * i + 1] = c.X[1];
    XR[4 * i + 2] = c.X[2];
    XR[4 * i + 3] = c.R;
  }
}

// getList on the last tree; returns total M2L / P2P entries via out[2].
API void ref3d_make_lists(int64_t *totals) {
  for (auto &c : g_cells) {
    c.listM2L.clear();
    c.listP2P.clear();
  }
  getList(&g_cells[0], &g_cells[0]);
  int64_t a = 0, b = 0;
  for (auto &c : g_cells) {
    a += c.listM2L.size();
    b += c.listP2P.size();
  }
  totals[0] = a;
  totals[1] = b;
}

API void ref3d_get_lists(int which, int64_t *offsets, int64_t *sources) {
  int64_t o = 0;
  for (size_t i = 0; i < g_cells.size(); i++) {
    offsets[i] = o;
    const auto &l = which == 0 ? g_cells[i].listM2L : g_cells[i].listP2P;
    for (auto *c : l) sources[o++] = int64_t(c - &g_cells[0]);
  }
  offsets[g_cells.size()] = o;
}

// Full evaluation: buildTree, upwardPass, horizontalPass (getList +
// evaluate), downwardPass.  times[0..4] = build, upward, lists, evaluate,
// downward (seconds).  Bodies are written back (permuted, with p and F).
API int ref3d_fmm(Body *b, int64_t n, double *times, double *coefs_out_M) {
  load(g_bodies, b, n);
  double t0 = now_s();
  g_cells = buildTree(g_bodies);
  double t1 = now_s();
  upwardPass(g_cells);
  double t2 = now_s();
  getList(&g_cells[0], &g_cells[0]);
  double t3 = now_s();
  evaluate(g_cells);
  double t4 = now_s();
  downwardPass(g_cells);
  double t5 = now_s();
  if (times) {
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t3 - t2;
    times[3] = t4 - t3;
    times[4] = t5 - t4;
  }
  if (coefs_out_M) {
    for (size_t i = 0; i < g_cells.size(); i++) get_coef(g_cells[i].M, coefs_out_M + 2 * NTERM * i);
  }
  std::memcpy(b, g_bodies.data(), sizeof(Body) * n);
  return int(g_cells.size());
}

// Two-tree evaluation: buildTree of the targets and of the sources,
// upwardPass of both, horizontalPass(icells, jcells) in its lazy form
// (getList(&icells[0], &jcells[0]) then evaluate, 3d/traverse_lazy.h:59-62)
// and downwardPass of the targets.  Both body arrays are written back
// (permuted; targets with p and F).  totals[0..1] = M2L / P2P entries.
API void ref3d_fmm_two(Body *t, int64_t nt, Body *s, int64_t ns, int64_t *totals) {
  load(g_bodies, t, nt);
  load(g_jbodies, s, ns);
  g_cells = buildTree(g_bodies);
  g_jcells = buildTree(g_jbodies);
  upwardPass(g_cells);
  upwardPass(g_jcells);
  getList(&g_cells[0], &g_jcells[0]);
  evaluate(g_cells);
  downwardPass(g_cells);
  int64_t a = 0, b = 0;
  for (auto &c : g_cells) {
    a += c.listM2L.size();
    b += c.listP2P.size();
  }
  totals[0] = a;
  totals[1] = b;
  std::memcpy(t, g_bodies.data(), sizeof(Body) * nt);
  std::memcpy(s, g_jbodies.data(), sizeof(Body) * ns);
}

API int64_t ref3d_num_cells(int source_tree) { return int64_t(source_tree ? g_jcells.size() : g_cells.size()); }

// Lists of the two-tree evaluation: CSR by target cell, sources numbered in
// the source tree.
API void ref3d_get_lists_two(int which, int64_t *offsets, int64_t *sources) {
  int64_t o = 0;
  for (size_t i = 0; i < g_cells.size(); i++) {
    offsets[i] = o;
    const auto &l = which == 0 ? g_cells[i].listM2L : g_cells[i].listP2P;
    for (auto *c : l) sources[o++] = int64_t(c - &g_jcells[0]);
  }
  offsets[g_cells.size()] = o;
}

// L coefficients of the last full evaluation.
API void ref3d_get_L(double *out) {
  for (size_t i = 0; i < g_cells.size(); i++) get_coef(g_cells[i].L, out + 2 * NTERM * i);
}

API void ref3d_direct(Body *t, int64_t nt, const Body *s, int64_t ns) {
  Bodies tb, sb;
  load(tb, t, nt);
  load(sb, s, ns);
  direct(tb, sb);
  std::memcpy(t, tb.data(), sizeof(Body) * nt);
}

// ---- single operator calls on ad-hoc cells
API void ref3d_p2p(Body *t, int64_t nt, const Body *s, int64_t ns) {
  Bodies tb, sb;
  load(tb, t, nt);
  load(sb, s, ns);
  Cells c(2);
  c[0].BODY = tb.data();
  c[0].NBODY = int(nt);
  c[1].BODY = sb.data();
  c[1].NBODY = int(ns);
  P2P(&c[0], &c[1]);
  std::memcpy(t, tb.data(), sizeof(Body) * nt);
}

API void ref3d_p2m(const double *C, const Body *bodies, int64_t nb, double *M) {
  Bodies bb;
  load(bb, bodies, nb);
  Cells c(1);
  for (int d = 0; d < 3; d++) c[0].X[d] = C[d];
  c[0].BODY = bb.data();
  c[0].NBODY = int(nb);
  c[0].NCHILD = 0;
  set_coef(c[0].M, M);
  P2M(&c[0]);
  get_coef(c[0].M, M);
}

API void ref3d_m2m(const double *Cp, const double *Cc, const double *Mc, double *Mp) {
  Cells c(2);
  for (int d = 0; d < 3; d++) {
    c[0].X[d] = Cp[d];
    c[1].X[d] = Cc[d];
  }
  c[0].CHILD = &c[1];
  c[0].NCHILD = 1;
  set_coef(c[1].M, Mc);
  set_coef(c[0].M, Mp);
  M2M(&c[0]);
  get_coef(c[0].M, Mp);
}

API void ref3d_m2l(const double *Xi, const double *Xj, const double *Mj, double *Li) {
  Cells c(2);
  for (int d = 0; d < 3; d++) {
    c[0].X[d] = Xi[d];
    c[1].X[d] = Xj[d];
  }
  set_coef(c[1].M, Mj);
  set_coef(c[0].L, Li);
  M2L(&c[0], &c[1]);
  get_coef(c[0].L, Li);
}

API void ref3d_l2l(const double *Cp, const double *Lp, const double *Cc, double *Lc) {
  Cells c(2);
  for (int d = 0; d < 3; d++) {
    c[0].X[d] = Cp[d];
    c[1].X[d] = Cc[d];
  }
  c[0].CHILD = &c[1];
  c[0].NCHILD = 1;
  set_coef(c[0].L, Lp);
  set_coef(c[1].L, Lc);
  L2L(&c[0]);
  get_coef(c[1].L, Lc);
}

API void ref3d_l2p(const double *C, const double *L, Body *bodies, int64_t nb) {
  Bodies bb;
  load(bb, bodies, nb);
  Cells c(1);
  for (int d = 0; d < 3; d++) c[0].X[d] = C[d];
  c[0].BODY = bb.data();
  c[0].NBODY = int(nb);
  set_coef(c[0].L, L);
  L2P(&c[0]);
  std::memcpy(bodies, bb.data(), sizeof(Body) * nb);
}

API void ref3d_eval_local(double rho, double alpha, double beta, double *Y) {
  std::vector<complex_t> y(4 * P * P, complex_t(0, 0));
  evalLocal(rho, alpha, beta, y.data());
  for (int i = 0; i < 4 * P * P; i++) {
    Y[2 * i] = y[i].real();
    Y[2 * i + 1] = y[i].imag();
  }
}

API void ref3d_eval_multipole(double rho, double alpha, double beta, double *Y, double *Yt) {
  std::vector<complex_t> y(P * P), yt(P * P);
  evalMultipole(rho, alpha, beta, y.data(), yt.data());
  for (int i = 0; i < P * P; i++) {
    Y[2 * i] = y[i].real();
    Y[2 * i + 1] = y[i].imag();
    Yt[2 * i] = yt[i].real();
    Yt[2 * i + 1] = yt[i].imag();
  }
}
