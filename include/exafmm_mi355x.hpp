// exafmm_mi355x.hpp -- the reference's C++ surface over the MI355X C ABI.
//
// A caller written against minimal-ExaFMM's 3d/ headers (3d/exafmm.h,
// 3d/build_tree.h, 3d/kernel.h, 3d/traverse_eager.h or 3d/traverse_lazy.h)
// recompiles unchanged against this header instead:
//   globals      P, NTERM, ncrit, theta            (3d/exafmm.h:40-43)
//   types        Body, Bodies, Cell, Cells          (3d/exafmm.h:14-37)
//   tree         Cells buildTree(Bodies&)           (3d/build_tree.h:88)
//   passes       upwardPass, horizontalPass, downwardPass, direct
//                                                   (3d/traverse_eager.h:140-202)
//   kernels      initKernel, P2P, P2M, M2M, M2L, L2L, L2P (3d/kernel.h:118-322)
// The passes run on the GPU through one device context per Cells object:
// Cells is the reference's std::vector<Cell> with a shared handle on that
// context, which is destroyed with the last Cells that holds it (so a
// time-stepping driver that rebuilds its tree every step holds one context at
// a time).  Cell::M is filled from the device after upwardPass and Cell::L
// after horizontalPass (the M2L sums, as the reference's M2L writes them) and
// again after downwardPass, so code that inspects them keeps working.  The
// single-call kernels run one FP64 device call each.  The extra knobs of the
// C ABI (truncation, precision, device) are set through exafmm::options().
#ifndef EXAFMM_MI355X_HPP
#define EXAFMM_MI355X_HPP
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "exafmm_mi355x.h"

namespace exafmm {
typedef double real_t;
typedef std::complex<real_t> complex_t;

struct Body {  // 3d/exafmm.h:14-19, same 64-byte layout as exafmm_body
  real_t X[3];
  real_t q;
  real_t p;
  real_t F[3];
};
static_assert(sizeof(Body) == sizeof(exafmm_body), "Body layout");
typedef std::vector<Body> Bodies;

struct Cell {  // 3d/exafmm.h:23-36
  int NCHILD;
  int NBODY;
  Cell *CHILD;
  Body *BODY;
  real_t X[3];
  real_t R;
  std::vector<Cell *> listM2L;  // filled by horizontalPass (lazy form)
  std::vector<Cell *> listP2P;
  std::vector<complex_t> M;
  std::vector<complex_t> L;
};
namespace detail {
inline std::atomic<int> &live_contexts() {  // device contexts held by Cells objects
  static std::atomic<int> n{0};
  return n;
}
// The device context of one tree; destroyed with the last Cells holding it.
struct Handle {
  exafmm_ctx *ctx = nullptr;
  explicit Handle(exafmm_ctx *c) : ctx(c) { live_contexts()++; }
  Handle(const Handle &) = delete;
  Handle &operator=(const Handle &) = delete;
  ~Handle() {
    exafmm_destroy(ctx);
    live_contexts()--;
  }
};
}  // namespace detail

// std::vector<Cell> (3d/exafmm.h:37) plus the device context of its tree.
struct Cells : std::vector<Cell> {
  using std::vector<Cell>::vector;
  std::shared_ptr<detail::Handle> device;
};

inline int P = 10;
inline int NTERM = 55;
inline int ncrit = 64;
inline real_t theta = 0.4;

struct Options {
  int m2l_truncation = EXAFMM_M2L_FULL;
  int m2l_precision = EXAFMM_PREC_BF16X3;
  int p2p_precision = EXAFMM_PREC_FP32;
  int device = 0;
  bool fill_lists = false;  // copy the interaction lists into Cell::listM2L/listP2P
};
inline Options &options() {
  static Options o;
  return o;
}

namespace detail {
inline void check(int code, const exafmm_ctx *ctx) {
  if (code != EXAFMM_OK) throw std::runtime_error(std::string("exafmm: ") + exafmm_last_error(ctx));
}
inline exafmm_config config() {
  exafmm_config c;
  exafmm_config_default(&c);
  c.P = P;
  c.ncrit = ncrit;
  c.theta = theta;
  c.m2l_truncation = options().m2l_truncation;
  c.m2l_precision = options().m2l_precision;
  c.p2p_precision = options().p2p_precision;
  c.device = options().device;
  return c;
}
inline exafmm_ctx *bound(const Cells &cells) {
  if (!cells.device) throw std::runtime_error("exafmm: Cells not produced by buildTree");
  return cells.device->ctx;
}
inline exafmm_ctx *op_ctx() {  // context for the single-call kernels (FP64)
  static exafmm_ctx *ctx = nullptr;
  static int cP = -1, cT = -1;
  if (!ctx || cP != P || cT != options().m2l_truncation) {
    if (ctx) exafmm_destroy(ctx);
    exafmm_config c = config();
    c.m2l_precision = EXAFMM_PREC_FP64;
    c.p2p_precision = EXAFMM_PREC_FP64;
    check(exafmm_create(&c, &ctx), nullptr);
    cP = P;
    cT = options().m2l_truncation;
  }
  return ctx;
}
inline void get_coefs(exafmm_ctx *ctx, Cells &cells, bool locals) {
  std::vector<double> buf(cells.size() * 2 * NTERM);
  check(locals ? exafmm_get_locals(ctx, buf.data()) : exafmm_get_multipoles(ctx, buf.data()), ctx);
  for (size_t i = 0; i < cells.size(); i++) {
    auto &v = locals ? cells[i].L : cells[i].M;
    v.resize(NTERM);
    for (int k = 0; k < NTERM; k++) v[k] = complex_t(buf[(i * NTERM + k) * 2], buf[(i * NTERM + k) * 2 + 1]);
  }
}
inline std::vector<double> flat(const std::vector<complex_t> &v) {
  std::vector<double> o(2 * NTERM, 0.0);
  for (int k = 0; k < NTERM && k < int(v.size()); k++) {
    o[2 * k] = v[k].real();
    o[2 * k + 1] = v[k].imag();
  }
  return o;
}
inline void unflat(const std::vector<double> &o, std::vector<complex_t> &v) {
  v.resize(NTERM);
  for (int k = 0; k < NTERM; k++) v[k] = complex_t(o[2 * k], o[2 * k + 1]);
}
}  // namespace detail

// 3d/kernel.h:118-120
inline void initKernel() { NTERM = P * (P + 1) / 2; }

// 3d/build_tree.h:88-96: bodies' X and q are permuted in place; every
// Cell::BODY points into `bodies`, which must not reallocate and must stay
// alive until downwardPass has written p and F back into it.
inline Cells buildTree(Bodies &bodies) {
  initKernel();
  exafmm_ctx *raw = nullptr;
  exafmm_config c = detail::config();
  detail::check(exafmm_create(&c, &raw), nullptr);
  auto handle = std::make_shared<detail::Handle>(raw);
  exafmm_ctx *ctx = handle->ctx;
  detail::check(exafmm_build_tree(ctx, reinterpret_cast<exafmm_body *>(bodies.data()), int64_t(bodies.size()), 0),
                ctx);
  int64_t n = 0;
  int32_t lv = 0;
  detail::check(exafmm_num_cells(ctx, &n, &lv), ctx);
  std::vector<exafmm_cell> flat(n);
  detail::check(exafmm_get_cells(ctx, flat.data(), n), ctx);
  Cells cells(n);
  for (int64_t i = 0; i < n; i++) {
    Cell &o = cells[i];
    o.NCHILD = flat[i].nchild;
    o.NBODY = flat[i].nbody;
    o.CHILD = flat[i].child >= 0 ? &cells[flat[i].child] : nullptr;
    o.BODY = &bodies[flat[i].body];
    for (int d = 0; d < 3; d++) o.X[d] = flat[i].X[d];
    o.R = flat[i].R;
  }
  cells.device = std::move(handle);
  return cells;
}

// Releases the device context of a tree before the Cells object goes away
// (not in the reference, whose Cells hold no device resources).
inline void releaseTree(Cells &cells) { cells.device.reset(); }

// 3d/traverse_eager.h:127-144: M of every cell; L allocated and zeroed.
inline void upwardPass(Cells &cells) {
  exafmm_ctx *ctx = detail::bound(cells);
  detail::check(exafmm_upward(ctx), ctx);
  detail::get_coefs(ctx, cells, false);
  for (auto &c : cells) c.L.assign(NTERM, complex_t(0, 0));
}

// 3d/traverse_eager.h:147-172 (list form 3d/traverse_lazy.h:27-56): M2L
// into icells' L (read back into Cell::L) and P2P into icells' bodies.
// Two distinct trees are supported; their M2L runs per pair in FP64.
inline void horizontalPass(Cells &icells, Cells &jcells) {
  exafmm_ctx *ci = detail::bound(icells);
  exafmm_ctx *cj = detail::bound(jcells);
  detail::check(exafmm_horizontal(ci, cj), ci);
  detail::get_coefs(ci, icells, true);
  if (options().fill_lists) {
    for (int which = 0; which < 2; which++) {
      int64_t nm = 0, np = 0;
      detail::check(exafmm_list_sizes(ci, &nm, &np), ci);
      std::vector<int64_t> off(icells.size() + 1), src(std::max<int64_t>(which ? np : nm, 1));
      detail::check(exafmm_get_lists(ci, which, off.data(), src.data()), ci);
      for (size_t i = 0; i < icells.size(); i++) {
        auto &l = which ? icells[i].listP2P : icells[i].listM2L;
        l.clear();
        for (int64_t k = off[i]; k < off[i + 1]; k++) l.push_back(&jcells[src[k]]);
      }
    }
  }
}

// 3d/traverse_eager.h:175-190: p and F are accumulated into the bodies.
inline void downwardPass(Cells &cells) {
  exafmm_ctx *ctx = detail::bound(cells);
  detail::check(exafmm_downward(ctx), ctx);
  detail::get_coefs(ctx, cells, true);
}

// 3d/traverse_eager.h:193-202 (FP64 on the device)
inline void direct(Bodies &bodies, Bodies &jbodies) {
  exafmm_ctx *ctx = detail::op_ctx();
  detail::check(exafmm_direct(ctx, reinterpret_cast<exafmm_body *>(bodies.data()), int64_t(bodies.size()),
                              reinterpret_cast<const exafmm_body *>(jbodies.data()), int64_t(jbodies.size()), 0),
                ctx);
}

// ---- kernel.h single calls (3d/kernel.h:122-322), FP64 on the device
inline void P2P(Cell *Ci, Cell *Cj) {
  exafmm_ctx *ctx = detail::op_ctx();
  detail::check(exafmm_p2p(ctx, reinterpret_cast<exafmm_body *>(Ci->BODY), Ci->NBODY,
                           reinterpret_cast<const exafmm_body *>(Cj->BODY), Cj->NBODY),
                ctx);
}
inline void P2M(Cell *C) {
  exafmm_ctx *ctx = detail::op_ctx();
  auto M = detail::flat(C->M);
  detail::check(exafmm_p2m(ctx, C->X, reinterpret_cast<const exafmm_body *>(C->BODY), C->NBODY, M.data()), ctx);
  detail::unflat(M, C->M);
}
inline void M2M(Cell *Ci) {
  exafmm_ctx *ctx = detail::op_ctx();
  auto M = detail::flat(Ci->M);
  for (Cell *Cj = Ci->CHILD; Cj != Ci->CHILD + Ci->NCHILD; Cj++) {
    auto Mc = detail::flat(Cj->M);
    detail::check(exafmm_m2m(ctx, Ci->X, Cj->X, Mc.data(), M.data()), ctx);
  }
  detail::unflat(M, Ci->M);
}
inline void M2L(Cell *Ci, Cell *Cj) {
  exafmm_ctx *ctx = detail::op_ctx();
  auto L = detail::flat(Ci->L);
  auto M = detail::flat(Cj->M);
  detail::check(exafmm_m2l(ctx, Ci->X, Cj->X, M.data(), L.data()), ctx);
  detail::unflat(L, Ci->L);
}
inline void L2L(Cell *Cj) {
  exafmm_ctx *ctx = detail::op_ctx();
  auto Lp = detail::flat(Cj->L);
  for (Cell *Ci = Cj->CHILD; Ci != Cj->CHILD + Cj->NCHILD; Ci++) {
    auto L = detail::flat(Ci->L);
    detail::check(exafmm_l2l(ctx, Cj->X, Lp.data(), Ci->X, L.data()), ctx);
    detail::unflat(L, Ci->L);
  }
}
inline void L2P(Cell *Ci) {
  exafmm_ctx *ctx = detail::op_ctx();
  auto L = detail::flat(Ci->L);
  detail::check(exafmm_l2p(ctx, Ci->X, L.data(), reinterpret_cast<exafmm_body *>(Ci->BODY), Ci->NBODY), ctx);
}

}  // namespace exafmm
#endif
