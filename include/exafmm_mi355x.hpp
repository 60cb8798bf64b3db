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
// The passes run on the GPU through one device context per Cells object;
// Cell::M and Cell::L are filled from the device after the passes so code
// that inspects them keeps working.  The single-call kernels run one FP64
// device call each.  The extra knobs of the C ABI (truncation, precision,
// device) are set through exafmm::options().
#ifndef EXAFMM_MI355X_HPP
#define EXAFMM_MI355X_HPP
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <map>
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
typedef std::vector<Cell> Cells;

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
struct Bound {
  exafmm_ctx *ctx = nullptr;
  Bodies *bodies = nullptr;
};
// one device context per Cells array, keyed by its first cell
inline std::map<const Cell *, Bound> &registry() {
  static std::map<const Cell *, Bound> r;
  return r;
}
inline Bound &bound(const Cells &cells) {
  auto it = registry().find(cells.data());
  if (it == registry().end()) throw std::runtime_error("exafmm: Cells not produced by buildTree");
  return it->second;
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
// Cell::BODY points into `bodies`, which must not reallocate afterwards.
inline Cells buildTree(Bodies &bodies) {
  initKernel();
  exafmm_ctx *ctx = nullptr;
  exafmm_config c = detail::config();
  detail::check(exafmm_create(&c, &ctx), nullptr);
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
  detail::registry()[cells.data()] = detail::Bound{ctx, &bodies};
  return cells;
}

// 3d/traverse_eager.h:127-144
inline void upwardPass(Cells &cells) {
  auto &b = detail::bound(cells);
  detail::check(exafmm_upward(b.ctx), b.ctx);
  detail::get_coefs(b.ctx, cells, false);
  for (auto &c : cells) c.L.assign(NTERM, complex_t(0, 0));
}

// 3d/traverse_eager.h:147-172 (list form 3d/traverse_lazy.h:27-56)
inline void horizontalPass(Cells &icells, Cells &jcells) {
  auto &bi = detail::bound(icells);
  auto &bj = detail::bound(jcells);
  detail::check(exafmm_horizontal(bi.ctx, bj.ctx), bi.ctx);
  if (options().fill_lists) {
    for (int which = 0; which < 2; which++) {
      int64_t nm = 0, np = 0;
      detail::check(exafmm_list_sizes(bi.ctx, &nm, &np), bi.ctx);
      std::vector<int64_t> off(icells.size() + 1), src(std::max<int64_t>(which ? np : nm, 1));
      detail::check(exafmm_get_lists(bi.ctx, which, off.data(), src.data()), bi.ctx);
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
  auto &b = detail::bound(cells);
  detail::check(exafmm_downward(b.ctx), b.ctx);
  detail::get_coefs(b.ctx, cells, true);
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
