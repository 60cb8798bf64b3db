// fmm_oracle.cpp -- FP64 CPU oracle for the 3-D Laplace FMM evaluation path.
//
// TEST INFRASTRUCTURE ONLY.  This is a restatement of the minimal-ExaFMM
// algorithm (reference directory 3d/, periodic variant 3dp/) used to check
// the HIP implementation.  It is never linked into the product library.
//
// Bit-exactness: every floating-point expression that determines a result
// is evaluated with the same operations in the same order as the reference
// (std::complex<double>, no FMA contraction, glibc libm), so that the tree,
// the lists and the final p/F agree bit for bit with the reference built by
// oracle/Makefile (checked against fixtures in tests/golden/).  Control
// structure (index-based cells, level-ordered passes, explicit work stacks)
// is this file's own.
//
// Parity pins (tests/test_oracle_golden.py, tests/test_oracle_pins.py):
//   tree, lists, p/F at N=1e3/1e4      -> reference fixtures, bit-exact
//   tree digest, list counts, N=2^20   -> reference fixture
//   operators at P=4/10/14             -> reference fixtures, <=1e-14 rel
//   KAT (3d/kernel.cxx)                -> reference error values
//   P2P                                -> reference fixtures, an independent
//                                         numpy sum, Newton's third law
//   M2M/L2L                            -> exact shift identities
//   M2L full mode                      -> errors of the reference's kernels
//                                         filled to degree 2P-1 (config-1
//                                         set, KAT), convergence to the
//                                         direct sum in P
//   periodic, full mode                -> the reference's 3dp run at P=18
//   Ewald                              -> the reference's Ewald fixture and
//                                         independence of alpha
#include "oracle.h"

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;
const cplx kI(0., 1.);

// |X|^2 evaluated as the reference's `norm` (3d/kernel.h:9-11).
inline double sq3(const double *X) { return X[0] * X[0] + X[1] * X[1] + X[2] * X[2]; }
// (-1)^n (3d/kernel.h:14-16)
inline int parity(int n) { return ((n & 1) == 1) ? -1 : 1; }
// i^{2n} helper used by M2M (3d/kernel.h:19-21)
inline int ipow2n(int n) { return (n >= 0) ? 1 : parity(n); }

// Cartesian -> spherical (3d/kernel.h:24-28).
inline void to_sph(const double *d, double &r, double &th, double &ph) {
  r = std::sqrt(sq3(d));
  th = r == 0 ? 0 : std::acos(d[2] / r);
  ph = std::atan2(d[1], d[0]);
}

// Spherical gradient -> Cartesian gradient (3d/kernel.h:31-40).
inline void grad_to_cart(double r, double th, double ph, const double *s, double *c) {
  c[0] = std::sin(th) * std::cos(ph) * s[0] + std::cos(th) * std::cos(ph) / r * s[1] -
         std::sin(ph) / r / std::sin(th) * s[2];
  c[1] = std::sin(th) * std::sin(ph) * s[0] + std::cos(th) * std::sin(ph) / r * s[1] +
         std::cos(ph) / r / std::sin(th) * s[2];
  c[2] = std::cos(th) * s[0] - std::sin(th) / r * s[1];
}

// Regular solid harmonics and their theta derivative, degrees < P
// (3d/kernel.h:43-80).  Y[n*n+n+m], -n<=m<=n.
void regular_harmonics(int P, double rho, double alpha, double beta, cplx *Y, cplx *Yth) {
  const double x = std::cos(alpha);
  const double y = std::sin(alpha);
  const double invY = y == 0 ? 0 : 1 / y;
  double fact = 1, pn = 1, rhom = 1;
  const cplx ei = std::exp(kI * beta);
  cplx eim = 1.0;
  for (int m = 0; m < P; m++) {
    double p = pn;
    const int ipos = m * m + 2 * m, ineg = m * m;
    Y[ipos] = rhom * p * eim;
    Y[ineg] = std::conj(Y[ipos]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    Yth[ipos] = rhom * (p - (m + 1) * x * p1) * invY * eim;
    rhom *= rho;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      const int a = n * n + n + m, b = n * n + n - m;
      rhon /= -(n + m);
      Y[a] = rhon * p * eim;
      Y[b] = std::conj(Y[a]);
      const double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      Yth[a] = rhon * ((n - m + 1) * p - (n + 1) * x * p1) * invY * eim;
      rhon *= rho;
    }
    rhom /= -(2 * m + 2) * (2 * m + 1);
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

// Singular harmonics rho^{-n-1} Y_n^m for degrees < fillP (3d/kernel.h:83-116).
// The reference fills only fillP == P ("reference" truncation); fillP == 2P
// gives the "full" M2L that keeps every j+n <= 2P-2 term.
void singular_harmonics(int fillP, double rho, double alpha, double beta, cplx *Y) {
  const double x = std::cos(alpha);
  const double y = std::sin(alpha);
  double fact = 1, pn = 1;
  const double invR = -1.0 / rho;
  double rhom = -invR;
  const cplx ei = std::exp(kI * beta);
  cplx eim = 1.0;
  for (int m = 0; m < fillP; m++) {
    double p = pn;
    const int ipos = m * m + 2 * m, ineg = m * m;
    Y[ipos] = rhom * p * eim;
    Y[ineg] = std::conj(Y[ipos]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    rhom *= invR;
    double rhon = rhom;
    for (int n = m + 1; n < fillP; n++) {
      const int a = n * n + n + m, b = n * n + n - m;
      Y[a] = rhon * p * eim;
      Y[b] = std::conj(Y[a]);
      const double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhon *= invR * (n - m + 1);
    }
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

// ---------------------------------------------------------------- operators

// Body-body interaction (3d/kernel.h:122-150; 3dp/kernel.h:122-150 adds the
// image shift).  Targets accumulate p += q/r and F -= q*dX/r^3.
void p2p_block(orc_body *Bi, int64_t ni, const orc_body *Bj, int64_t nj, const double *shift) {
  for (int64_t i = 0; i < ni; i++) {
    double pot = 0, ax = 0, ay = 0, az = 0;
    for (int64_t j = 0; j < nj; j++) {
      double d[3];
      if (shift) {
        for (int k = 0; k < 3; k++) d[k] = Bi[i].X[k] - Bj[j].X[k] - shift[k];
      } else {
        for (int k = 0; k < 3; k++) d[k] = Bi[i].X[k] - Bj[j].X[k];
      }
      const double R2 = sq3(d);
      if (R2 != 0) {
        const double invR2 = 1.0 / R2;
        const double invR = Bj[j].q * std::sqrt(invR2);
        for (int k = 0; k < 3; k++) d[k] *= invR2 * invR;
        pot += invR;
        ax += d[0];
        ay += d[1];
        az += d[2];
      }
    }
    Bi[i].p += pot;
    Bi[i].F[0] -= ax;
    Bi[i].F[1] -= ay;
    Bi[i].F[2] -= az;
  }
}

// Particle -> multipole (3d/kernel.h:187-202).
void p2m(int P, const double *C, const orc_body *B, int64_t nb, cplx *M) {
  std::vector<cplx> Y(P * P), Yth(P * P);
  for (int64_t b = 0; b < nb; b++) {
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = B[b].X[k] - C[k];
    double rho, alpha, beta;
    to_sph(d, rho, alpha, beta);
    regular_harmonics(P, rho, alpha, -beta, Y.data(), Yth.data());
    for (int n = 0; n < P; n++)
      for (int m = 0; m <= n; m++) M[n * (n + 1) / 2 + m] += B[b].q * Y[n * n + n + m];
  }
}

// Multipole shift child -> parent (3d/kernel.h:204-231), one child.
void m2m(int P, const double *Cp, const double *Cc, const cplx *Mc, cplx *Mp) {
  std::vector<cplx> Y(P * P), Yth(P * P);
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = Cp[k] - Cc[k];
  double rho, alpha, beta;
  to_sph(d, rho, alpha, beta);
  regular_harmonics(P, rho, alpha, beta, Y.data(), Yth.data());
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      cplx acc = 0;
      for (int n = 0; n <= j; n++) {
        for (int m = std::max(-n, -j + k + n); m <= std::min(k - 1, n); m++) {
          const int src = (j - n) * (j - n + 1) / 2 + k - m;
          acc += Mc[src] * Y[n * n + n - m] * double(ipow2n(m) * parity(n));
        }
        for (int m = k; m <= std::min(n, j + k - n); m++) {
          const int src = (j - n) * (j - n + 1) / 2 - k + m;
          acc += std::conj(Mc[src]) * Y[n * n + n - m] * double(parity(k + n + m));
        }
      }
      Mp[j * (j + 1) / 2 + k] += acc;
    }
  }
}

// Multipole -> local (3d/kernel.h:233-260).  fillP selects the truncation.
void m2l(int P, int full, const double *d, const cplx *Mj, cplx *Li) {
  std::vector<cplx> Y(4 * P * P, cplx(0.0, 0.0));
  double rho, alpha, beta;
  to_sph(d, rho, alpha, beta);
  singular_harmonics(full ? 2 * P : P, rho, alpha, beta, Y.data());
  for (int j = 0; j < P; j++) {
    const double cj = parity(j);
    for (int k = 0; k <= j; k++) {
      cplx acc = 0;
      for (int n = 0; n < P; n++) {
        for (int m = -n; m < 0; m++) {
          const int src = n * (n + 1) / 2 - m;
          const int yi = (j + n) * (j + n) + j + n + m - k;
          acc += std::conj(Mj[src]) * cj * Y[yi];
        }
        for (int m = 0; m <= n; m++) {
          const int src = n * (n + 1) / 2 + m;
          const int yi = (j + n) * (j + n) + j + n + m - k;
          const double c2 = cj * parity((k - m) * (k < m) + m);
          acc += Mj[src] * c2 * Y[yi];
        }
      }
      Li[j * (j + 1) / 2 + k] += acc;
    }
  }
}

// Local shift parent -> child (3d/kernel.h:262-291), one child.
void l2l(int P, const double *Cp, const cplx *Lp, const double *Cc, cplx *Lc) {
  std::vector<cplx> Y(P * P), Yth(P * P);
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = Cc[k] - Cp[k];
  double rho, alpha, beta;
  to_sph(d, rho, alpha, beta);
  regular_harmonics(P, rho, alpha, beta, Y.data(), Yth.data());
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      cplx acc = 0;
      for (int n = j; n < P; n++) {
        for (int m = j + k - n; m < 0; m++) {
          const int yi = (n - j) * (n - j) + n - j + m - k;
          const int src = n * (n + 1) / 2 - m;
          acc += std::conj(Lp[src]) * Y[yi] * double(parity(k));
        }
        for (int m = 0; m <= n; m++) {
          if (n - j >= std::abs(m - k)) {
            const int yi = (n - j) * (n - j) + n - j + m - k;
            const int src = n * (n + 1) / 2 + m;
            acc += Lp[src] * Y[yi] * double(parity((m - k) * (m < k)));
          }
        }
      }
      Lc[j * (j + 1) / 2 + k] += acc;
    }
  }
}

// Local -> particle (3d/kernel.h:293-322).
void l2p(int P, const double *C, const cplx *L, orc_body *B, int64_t nb) {
  std::vector<cplx> Y(P * P), Yth(P * P);
  for (int64_t b = 0; b < nb; b++) {
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = B[b].X[k] - C[k];
    double sph[3] = {0, 0, 0}, cart[3] = {0, 0, 0};
    double r, th, ph;
    to_sph(d, r, th, ph);
    regular_harmonics(P, r, th, ph, Y.data(), Yth.data());
    for (int n = 0; n < P; n++) {
      int nm = n * n + n;
      int nms = n * (n + 1) / 2;
      B[b].p += std::real(L[nms] * Y[nm]);
      sph[0] += std::real(L[nms] * Y[nm]) / r * n;
      sph[1] += std::real(L[nms] * Yth[nm]);
      for (int m = 1; m <= n; m++) {
        nm = n * n + n + m;
        nms = n * (n + 1) / 2 + m;
        B[b].p += 2 * std::real(L[nms] * Y[nm]);
        sph[0] += 2 * std::real(L[nms] * Y[nm]) / r * n;
        sph[1] += 2 * std::real(L[nms] * Yth[nm]);
        sph[2] += 2 * std::real(L[nms] * Y[nm] * kI) * m;
      }
    }
    grad_to_cart(r, th, ph, sph, cart);
    for (int k = 0; k < 3; k++) B[b].F[k] += cart[k];
  }
}

// ---------------------------------------------------------------- tree

struct Node {
  int32_t nchild, nbody;
  int64_t child, body;
  double X[3];
  double R;
  int32_t level;
};

// Top-down octree of 3d/build_tree.h:7-96, restated with an index
// permutation instead of ping-pong body buffers.  The resulting cell table
// (children blocks appended in DFS preorder of their parents, stable
// octant partition, strict '>' against the FP64 centre sequence) and body
// order are identical to the reference's.
struct TreeBuild {
  const orc_body *in;
  int ncrit;
  double R0;
  std::vector<int64_t> perm, tmp;
  std::vector<Node> &cells;
  int status = ORC_OK;
  static const int kMaxLevel = 21;

  TreeBuild(const orc_body *b, int nc, std::vector<Node> &c) : in(b), ncrit(nc), cells(c) {}

  void build(int64_t begin, int64_t end, int64_t ci, const double *X, int level) {
    Node &c = cells[ci];
    c.body = begin;
    c.nbody = int32_t(end - begin);
    c.nchild = 0;
    c.child = -1;
    for (int d = 0; d < 3; d++) c.X[d] = X[d];
    c.R = R0 / (1 << level);
    c.level = level;
    if (end - begin <= ncrit) return;
    if (level >= kMaxLevel) {
      status = ORC_ERR_DEPTH;
      return;
    }
    int64_t size[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = begin; i < end; i++) {
      const double *x = in[perm[i]].X;
      const int oct = (x[0] > X[0]) + ((x[1] > X[1]) << 1) + ((x[2] > X[2]) << 2);
      size[oct]++;
    }
    int64_t start[8], fill[8];
    int64_t off = begin;
    int nchild = 0;
    for (int o = 0; o < 8; o++) {
      start[o] = off;
      off += size[o];
      if (size[o]) nchild++;
    }
    for (int o = 0; o < 8; o++) fill[o] = start[o];
    for (int64_t i = begin; i < end; i++) {
      const double *x = in[perm[i]].X;
      const int oct = (x[0] > X[0]) + ((x[1] > X[1]) << 1) + ((x[2] > X[2]) << 2);
      tmp[fill[oct]++] = perm[i];
    }
    std::copy(tmp.begin() + begin, tmp.begin() + end, perm.begin() + begin);
    const int64_t first = int64_t(cells.size());
    cells.resize(cells.size() + nchild);
    cells[ci].nchild = nchild;
    cells[ci].child = first;
    const double r = R0 / (1 << (level + 1));
    int c_idx = 0;
    for (int o = 0; o < 8; o++) {
      double Xc[3];
      for (int d = 0; d < 3; d++) Xc[d] = X[d];
      for (int d = 0; d < 3; d++) Xc[d] += r * (((o & 1 << d) >> d) * 2 - 1);
      if (size[o]) {
        build(start[o], start[o] + size[o], first + c_idx, Xc, level + 1);
        if (status != ORC_OK) return;
        c_idx++;
      }
    }
  }
};

// Bounding box (3d/build_tree.h:7-21).
void bounds(const orc_body *b, int64_t n, double &R0, double *X0) {
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) lo[d] = hi[d] = b[0].X[d];
  for (int64_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) lo[d] = std::fmin(b[i].X[d], lo[d]);
    for (int d = 0; d < 3; d++) hi[d] = std::fmax(b[i].X[d], hi[d]);
  }
  for (int d = 0; d < 3; d++) X0[d] = (hi[d] + lo[d]) / 2;
  R0 = 0;
  for (int d = 0; d < 3; d++) {
    R0 = std::fmax(X0[d] - lo[d], R0);
    R0 = std::fmax(hi[d] - X0[d], R0);
  }
  R0 *= 1.00001;
}

// ---------------------------------------------------------------- session

struct Session {
  int P, NTERM, ncrit;
  double theta;
  int full;
  std::vector<Node> cells;
  std::vector<cplx> M, L;
  std::vector<std::vector<int64_t>> lm2l, lp2p;    // per-target lists
  std::vector<std::vector<int32_t>> m2l_img, p2p_img;  // periodic image index
  std::vector<std::vector<int64_t>> by_level;
  // periodic state
  int images = 0;
  double cycle = 0;
  bool have_tree = false, have_lists = false;

  Session(int P_, int nc, double th, int fu)
      : P(P_), NTERM(P_ * (P_ + 1) / 2), ncrit(nc), theta(th), full(fu) {}

  int build(orc_body *b, int64_t n) {
    if (n <= 0 || ncrit < 1) return ORC_ERR_ARG;
    for (int64_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++)
        if (!std::isfinite(b[i].X[d])) return ORC_ERR_ARG;
    double R0, X0[3];
    bounds(b, n, R0, X0);
    cells.clear();
    cells.resize(1);
    TreeBuild tb(b, ncrit, cells);
    tb.R0 = R0;
    tb.perm.resize(n);
    tb.tmp.resize(n);
    for (int64_t i = 0; i < n; i++) tb.perm[i] = i;
    tb.build(0, n, 0, X0, 0);
    if (tb.status != ORC_OK) return tb.status;
    // apply permutation to X and q only (3d/build_tree.h:64-65)
    std::vector<double> xq(4 * n);
    for (int64_t i = 0; i < n; i++) {
      const orc_body &s = b[tb.perm[i]];
      xq[4 * i + 0] = s.X[0];
      xq[4 * i + 1] = s.X[1];
      xq[4 * i + 2] = s.X[2];
      xq[4 * i + 3] = s.q;
    }
    for (int64_t i = 0; i < n; i++) {
      b[i].X[0] = xq[4 * i + 0];
      b[i].X[1] = xq[4 * i + 1];
      b[i].X[2] = xq[4 * i + 2];
      b[i].q = xq[4 * i + 3];
    }
    int maxl = 0;
    for (auto &c : cells) maxl = std::max(maxl, c.level);
    by_level.assign(maxl + 1, {});
    for (int64_t i = 0; i < int64_t(cells.size()); i++) by_level[cells[i].level].push_back(i);
    M.assign(cells.size() * NTERM, cplx(0.0, 0.0));
    L.assign(cells.size() * NTERM, cplx(0.0, 0.0));
    have_tree = true;
    have_lists = false;
    return ORC_OK;
  }

  // Upward pass (3d/traverse_lazy.h:7-24): P2M at leaves, M2M child->parent,
  // children complete before their parent.  Done level by level from the
  // deepest; every cell's M is the same ordered sum as in the recursion.
  int upward(const orc_body *b) {
    if (!have_tree) return ORC_ERR_STATE;
    std::fill(M.begin(), M.end(), cplx(0.0, 0.0));
    std::fill(L.begin(), L.end(), cplx(0.0, 0.0));
    for (int lev = int(by_level.size()) - 1; lev >= 0; lev--) {
      const auto &ids = by_level[lev];
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t t = 0; t < int64_t(ids.size()); t++) {
        const Node &c = cells[ids[t]];
        cplx *Mc = &M[ids[t] * NTERM];
        if (c.nchild == 0) p2m(P, c.X, b + c.body, c.nbody, Mc);
        for (int k = 0; k < c.nchild; k++) {
          const Node &ch = cells[c.child + k];
          m2m(P, c.X, ch.X, &M[(c.child + k) * NTERM], Mc);
        }
      }
    }
    return ORC_OK;
  }

  // Dual-tree list construction (3d/traverse_lazy.h:27-43; 3dp adds the
  // image shift).  Explicit stack in the same visiting order as the
  // recursion so that each target's list order matches the reference.
  void lists_from(int64_t i0, int64_t j0, const double *shift, int img) {
    struct Pair { int64_t i, j; };
    std::vector<Pair> stack;
    stack.push_back({i0, j0});
    while (!stack.empty()) {
      Pair pr = stack.back();
      stack.pop_back();
      const Node &Ci = cells[pr.i];
      const Node &Cj = cells[pr.j];
      double d[3];
      if (shift) {
        for (int k = 0; k < 3; k++) d[k] = Ci.X[k] - Cj.X[k] - shift[k];
      } else {
        for (int k = 0; k < 3; k++) d[k] = Ci.X[k] - Cj.X[k];
      }
      const double R2 = sq3(d) * theta * theta;
      if (R2 > (Ci.R + Cj.R) * (Ci.R + Cj.R)) {
        lm2l[pr.i].push_back(pr.j);
        if (img >= 0) m2l_img[pr.i].push_back(img);
      } else if (Ci.nchild == 0 && Cj.nchild == 0) {
        lp2p[pr.i].push_back(pr.j);
        if (img >= 0) p2p_img[pr.i].push_back(img);
      } else if (Cj.nchild == 0 || (Ci.R >= Cj.R && Ci.nchild != 0)) {
        for (int k = Ci.nchild - 1; k >= 0; k--) stack.push_back({Ci.child + k, pr.j});
      } else {
        for (int k = Cj.nchild - 1; k >= 0; k--) stack.push_back({pr.i, Cj.child + k});
      }
    }
  }

  int make_lists() {
    if (!have_tree) return ORC_ERR_STATE;
    lm2l.assign(cells.size(), {});
    lp2p.assign(cells.size(), {});
    m2l_img.assign(cells.size(), {});
    p2p_img.assign(cells.size(), {});
    if (images == 0) {
      lists_from(0, 0, nullptr, -1);
    } else {
      // 27 near images, x outermost (3dp/traverse_lazy.h:124-135)
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            const double s[3] = {ix * cycle, iy * cycle, iz * cycle};
            const int img = ix + 1 + 3 * (iy + 1) + 9 * (iz + 1);
            lists_from(0, 0, s, img);
          }
    }
    have_lists = true;
    return ORC_OK;
  }

  void image_shift(int img, double *s) const {
    const int ix = (img % 3) - 1, iy = ((img / 3) % 3) - 1, iz = (img / 9) - 1;
    s[0] = ix * cycle;
    s[1] = iy * cycle;
    s[2] = iz * cycle;
  }

  // List evaluation (3d/traverse_lazy.h:46-56): per target cell, M2L list
  // then P2P list, in list order.
  int evaluate(orc_body *b) {
    if (!have_lists) return ORC_ERR_STATE;
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < int64_t(cells.size()); i++) {
      const Node &Ci = cells[i];
      for (size_t k = 0; k < lm2l[i].size(); k++) {
        const Node &Cj = cells[lm2l[i][k]];
        double d[3];
        if (images) {
          double s[3];
          image_shift(m2l_img[i][k], s);
          for (int q = 0; q < 3; q++) d[q] = Ci.X[q] - Cj.X[q] - s[q];
        } else {
          for (int q = 0; q < 3; q++) d[q] = Ci.X[q] - Cj.X[q];
        }
        m2l(P, full, d, &M[lm2l[i][k] * NTERM], &L[i * NTERM]);
      }
      for (size_t k = 0; k < lp2p[i].size(); k++) {
        const Node &Cj = cells[lp2p[i][k]];
        if (images) {
          double s[3];
          image_shift(p2p_img[i][k], s);
          p2p_block(b + Ci.body, Ci.nbody, b + Cj.body, Cj.nbody, s);
        } else {
          p2p_block(b + Ci.body, Ci.nbody, b + Cj.body, Cj.nbody, nullptr);
        }
      }
    }
    return ORC_OK;
  }

  // Far periodic images (3dp/traverse_lazy.h:75-121): a super-cell whose
  // multipole grows by 3x per level; M2L from the 26x27 shifted copies into
  // the root's local expansion, then M2M of the 26 neighbours into the
  // super-cell.
  void periodic_far() {
    const Node &root = cells[0];
    std::vector<cplx> Msup(M.begin(), M.begin() + NTERM);
    double Xs[3] = {root.X[0], root.X[1], root.X[2]};
    double cyc = cycle;
    for (int lev = 0; lev < images - 1; lev++) {
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            if (ix == 0 && iy == 0 && iz == 0) continue;
            for (int cx = -1; cx <= 1; cx++)
              for (int cy = -1; cy <= 1; cy++)
                for (int cz = -1; cz <= 1; cz++) {
                  const int sh[3] = {ix * 3 + cx, iy * 3 + cy, iz * 3 + cz};
                  double d[3];
                  for (int q = 0; q < 3; q++) d[q] = root.X[q] - Xs[q] - sh[q] * cyc;
                  m2l(P, full, d, Msup.data(), &L[0]);
                }
          }
      std::vector<cplx> Mcopy = Msup;
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            if (ix == 0 && iy == 0 && iz == 0) continue;
            const double Xc[3] = {Xs[0] + ix * cyc, Xs[1] + iy * cyc, Xs[2] + iz * cyc};
            m2m(P, Xs, Xc, Mcopy.data(), Msup.data());
          }
      cyc *= 3;
    }
  }

  // Downward pass (3d/traverse_lazy.h:65-80): L2L parent->children, L2P at
  // leaves; level by level from the root.
  int downward(orc_body *b) {
    if (!have_tree) return ORC_ERR_STATE;
    for (size_t lev = 0; lev < by_level.size(); lev++) {
      const auto &ids = by_level[lev];
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t t = 0; t < int64_t(ids.size()); t++) {
        const Node &c = cells[ids[t]];
        const cplx *Lc = &L[ids[t] * NTERM];
        for (int k = 0; k < c.nchild; k++) {
          const Node &ch = cells[c.child + k];
          l2l(P, c.X, Lc, ch.X, &L[(c.child + k) * NTERM]);
        }
        if (c.nchild == 0) l2p(P, c.X, Lc, b + c.body, c.nbody);
      }
    }
    return ORC_OK;
  }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void to_cplx(const double *in, cplx *out, int n) {
  for (int i = 0; i < n; i++) out[i] = cplx(in[2 * i], in[2 * i + 1]);
}
void from_cplx(const cplx *in, double *out, int n) {
  for (int i = 0; i < n; i++) {
    out[2 * i] = in[i].real();
    out[2 * i + 1] = in[i].imag();
  }
}

}  // namespace

// ================================================================ C API

extern "C" {

void orc_eval_multipole(int P, double rho, double alpha, double beta, double *Ynm, double *YnmTheta) {
  std::vector<cplx> Y(P * P), Yt(P * P);
  regular_harmonics(P, rho, alpha, beta, Y.data(), Yt.data());
  from_cplx(Y.data(), Ynm, P * P);
  from_cplx(Yt.data(), YnmTheta, P * P);
}

void orc_eval_local(int P, int fillP, double rho, double alpha, double beta, double *Ynm) {
  const int cap = std::max(4 * P * P, fillP * fillP);
  std::vector<cplx> Y(cap, cplx(0.0, 0.0));
  singular_harmonics(fillP, rho, alpha, beta, Y.data());
  from_cplx(Y.data(), Ynm, 4 * P * P);
}

void orc_p2p(orc_body *t, int64_t nt, const orc_body *s, int64_t ns) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; i++) p2p_block(t + i, 1, s, ns, nullptr);
}

void orc_p2p_shift(orc_body *t, int64_t nt, const orc_body *s, int64_t ns, const double shift[3]) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; i++) p2p_block(t + i, 1, s, ns, shift);
}

void orc_p2m(int P, const double C[3], const orc_body *b, int64_t nb, double *M) {
  const int NT = P * (P + 1) / 2;
  std::vector<cplx> Mc(NT);
  to_cplx(M, Mc.data(), NT);
  p2m(P, C, b, nb, Mc.data());
  from_cplx(Mc.data(), M, NT);
}

void orc_m2m(int P, const double Cp[3], const double Cc[3], const double *Mchild, double *Mparent) {
  const int NT = P * (P + 1) / 2;
  std::vector<cplx> a(NT), b(NT);
  to_cplx(Mchild, a.data(), NT);
  to_cplx(Mparent, b.data(), NT);
  m2m(P, Cp, Cc, a.data(), b.data());
  from_cplx(b.data(), Mparent, NT);
}

void orc_m2l(int P, int full, const double Xi[3], const double Xj[3], const double *Mj, double *Li) {
  const int NT = P * (P + 1) / 2;
  std::vector<cplx> a(NT), b(NT);
  to_cplx(Mj, a.data(), NT);
  to_cplx(Li, b.data(), NT);
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = Xi[k] - Xj[k];
  m2l(P, full, d, a.data(), b.data());
  from_cplx(b.data(), Li, NT);
}

void orc_l2l(int P, const double Cp[3], const double *Lp, const double Cc[3], double *Lc) {
  const int NT = P * (P + 1) / 2;
  std::vector<cplx> a(NT), b(NT);
  to_cplx(Lp, a.data(), NT);
  to_cplx(Lc, b.data(), NT);
  l2l(P, Cp, a.data(), Cc, b.data());
  from_cplx(b.data(), Lc, NT);
}

void orc_l2p(int P, const double C[3], const double *L, orc_body *b, int64_t nb) {
  const int NT = P * (P + 1) / 2;
  std::vector<cplx> a(NT);
  to_cplx(L, a.data(), NT);
  l2p(P, C, a.data(), b, nb);
}

// 3d/kernel.cxx:4-90: one source, chain P2M -> M2M -> M2L -> L2L -> L2P,
// compared with P2P.
void orc_kat(int P, int full, double *out) {
  const int NT = P * (P + 1) / 2;
  orc_body src{};
  src.X[0] = src.X[1] = src.X[2] = 2;
  src.q = 1;
  const double Xj[3] = {3, 1, 1}, XJ[3] = {4, 0, 0}, XI[3] = {-4, 0, 0}, Xi[3] = {-3, 1, 1};
  std::vector<cplx> Mj(NT, 0.0), MJ(NT, 0.0), LI(NT, 0.0), Li(NT, 0.0);
  p2m(P, Xj, &src, 1, Mj.data());
  m2m(P, XJ, Xj, Mj.data(), MJ.data());
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = XI[k] - XJ[k];
  m2l(P, full, d, MJ.data(), LI.data());
  l2l(P, XI, LI.data(), Xi, Li.data());
  orc_body t{};
  t.X[0] = -2;
  t.X[1] = 2;
  t.X[2] = 2;
  t.q = 1;
  l2p(P, Xi, Li.data(), &t, 1);
  orc_body t2 = t;
  t2.p = 0;
  t2.F[0] = t2.F[1] = t2.F[2] = 0;
  p2p_block(&t2, 1, &src, 1, nullptr);
  const double pd = (t.p - t2.p) * (t.p - t2.p), pn = t.p * t.p;
  double fd = 0, fn = 0;
  for (int k = 0; k < 3; k++) {
    fd += (t.F[k] - t2.F[k]) * (t.F[k] - t2.F[k]);
    fn += t.F[k] * t.F[k];
  }
  out[0] = std::sqrt(pd / pn);
  out[1] = std::sqrt(fd / fn);
  out[2] = t.p;
  out[3] = t.F[0];
  out[4] = t.F[1];
  out[5] = t.F[2];
  out[6] = t2.p;
  out[7] = t2.F[0];
  out[8] = t2.F[1];
  out[9] = t2.F[2];
}

void *orc_create(int P, int ncrit, double theta, int m2l_full) {
  if (P < 1 || P > 40 || ncrit < 1 || !(theta > 0)) return nullptr;
  return new Session(P, ncrit, theta, m2l_full);
}
void orc_destroy(void *h) { delete static_cast<Session *>(h); }
int orc_build_tree(void *h, orc_body *b, int64_t n) { return static_cast<Session *>(h)->build(b, n); }
int64_t orc_ncells(void *h) { return int64_t(static_cast<Session *>(h)->cells.size()); }
void orc_get_cells(void *h, orc_cell *out) {
  auto *s = static_cast<Session *>(h);
  for (size_t i = 0; i < s->cells.size(); i++) {
    const Node &c = s->cells[i];
    out[i].nchild = c.nchild;
    out[i].nbody = c.nbody;
    out[i].child = c.child;
    out[i].body = c.body;
    for (int d = 0; d < 3; d++) out[i].X[d] = c.X[d];
    out[i].R = c.R;
  }
}
void orc_get_levels(void *h, int32_t *out) {
  auto *s = static_cast<Session *>(h);
  for (size_t i = 0; i < s->cells.size(); i++) out[i] = s->cells[i].level;
}
int orc_upward(void *h, const orc_body *b) { return static_cast<Session *>(h)->upward(b); }
int orc_make_lists(void *h) { return static_cast<Session *>(h)->make_lists(); }
int64_t orc_list_total(void *h, int which) {
  auto *s = static_cast<Session *>(h);
  const auto &ls = which == 0 ? s->lm2l : s->lp2p;
  int64_t t = 0;
  for (auto &v : ls) t += int64_t(v.size());
  return t;
}
void orc_get_lists(void *h, int which, int64_t *offsets, int64_t *sources) {
  auto *s = static_cast<Session *>(h);
  const auto &ls = which == 0 ? s->lm2l : s->lp2p;
  int64_t o = 0;
  for (size_t i = 0; i < ls.size(); i++) {
    offsets[i] = o;
    for (auto v : ls[i]) sources[o++] = v;
  }
  offsets[ls.size()] = o;
}
int orc_evaluate(void *h, orc_body *b) { return static_cast<Session *>(h)->evaluate(b); }
int orc_downward(void *h, orc_body *b) { return static_cast<Session *>(h)->downward(b); }
void orc_get_M(void *h, double *out) {
  auto *s = static_cast<Session *>(h);
  from_cplx(s->M.data(), out, int(s->M.size()));
}
void orc_get_L(void *h, double *out) {
  auto *s = static_cast<Session *>(h);
  from_cplx(s->L.data(), out, int(s->L.size()));
}
void orc_set_M(void *h, const double *in) {
  auto *s = static_cast<Session *>(h);
  to_cplx(in, s->M.data(), int(s->M.size()));
}
void orc_set_L(void *h, const double *in) {
  auto *s = static_cast<Session *>(h);
  to_cplx(in, s->L.data(), int(s->L.size()));
}

int orc_fmm(orc_body *b, int64_t n, int P, int ncrit, double theta, int m2l_full, double *times) {
  Session s(P, ncrit, theta, m2l_full);
  double t0 = now_s();
  int st = s.build(b, n);
  if (st) return st;
  double t1 = now_s();
  s.upward(b);
  double t2 = now_s();
  s.make_lists();
  double t3 = now_s();
  s.evaluate(b);
  double t4 = now_s();
  s.downward(b);
  double t5 = now_s();
  if (times) {
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t3 - t2;
    times[3] = t4 - t3;
    times[4] = t5 - t4;
  }
  return ORC_OK;
}

int orc_fmm_periodic(orc_body *b, int64_t n, int P, int ncrit, double theta, int m2l_full, int images,
                     double cycle, double *times) {
  if (images < 1) return orc_fmm(b, n, P, ncrit, theta, m2l_full, times);
  Session s(P, ncrit, theta, m2l_full);
  s.images = images;
  s.cycle = cycle;
  double t0 = now_s();
  int st = s.build(b, n);
  if (st) return st;
  double t1 = now_s();
  s.upward(b);
  double t2 = now_s();
  s.make_lists();
  double t3 = now_s();
  s.evaluate(b);
  s.periodic_far();
  double t4 = now_s();
  s.downward(b);
  double t5 = now_s();
  if (times) {
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    times[2] = t3 - t2;
    times[3] = t4 - t3;
    times[4] = t5 - t4;
  }
  return ORC_OK;
}

// Periodic direct sum over (2*prange+1)^3 images, prange = sum_{i<images} 3^i
// (3dp/traverse_lazy.h:162-185); per target the image loop is sequential.
void orc_direct_periodic(orc_body *t, int64_t nt, const orc_body *s, int64_t ns, int images, double cycle) {
  int prange = 0;
  for (int i = 0; i < images; i++) prange += int(std::pow(3.0, i));
#pragma omp parallel for schedule(dynamic)
  for (int64_t i = 0; i < nt; i++) {
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          const double sh[3] = {ix * cycle, iy * cycle, iz * cycle};
          p2p_block(t + i, 1, s, ns, sh);
        }
  }
}

// Ewald sum (3dp/ewald.h:22-150): wave part by DFT/IDFT over the half-space
// of |k| <= ksize, real part with erfc over pairs within the cutoff, self
// term.  Adds into the targets' p and F.  The reference finds real-space
// pairs through a tree and picks the image per cell pair
// (3dp/ewald.h:104-119); here each body pair uses its nearest image, which
// differs only for pairs at distance ~cutoff, where erfc(alpha*cutoff) is
// below 1e-12 for the configurations used.
void orc_ewald(orc_body *t, int64_t nt, const orc_body *s, int64_t ns, int ksize, double alpha, double sigma,
               double cutoff, double cycle, int /*ncrit*/) {
  double scale[3];
  for (int d = 0; d < 3; d++) scale[d] = 2 * M_PI / cycle;
  struct Wave { double K[3]; double re, im; };
  std::vector<Wave> waves;
  const int kmaxsq = ksize * ksize, kmax = ksize;
  for (int l = 0; l <= kmax; l++) {
    const int mmin = l == 0 ? 0 : -kmax;
    for (int m = mmin; m <= kmax; m++) {
      const int nmin = (l == 0 && m == 0) ? 1 : -kmax;
      for (int n = nmin; n <= kmax; n++) {
        const double ksq = l * l + m * m + n * n;
        if (ksq <= kmaxsq) waves.push_back({{double(l), double(m), double(n)}, 0, 0});
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < int64_t(waves.size()); w++) {
    double re = 0, im = 0;
    for (int64_t b = 0; b < ns; b++) {
      double th = 0;
      for (int d = 0; d < 3; d++) th += waves[w].K[d] * s[b].X[d] * scale[d];
      re += s[b].q * std::cos(th);
      im += s[b].q * std::sin(th);
    }
    waves[w].re = re;
    waves[w].im = im;
  }
  const double coef = 2 / sigma / cycle / cycle / cycle;
  const double coef2 = 1 / (4 * alpha * alpha);
  for (auto &w : waves) {
    double K[3];
    for (int d = 0; d < 3; d++) K[d] = w.K[d] * scale[d];
    const double K2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    const double factor = coef * std::exp(-K2 * coef2) / K2;
    w.re *= factor;
    w.im *= factor;
  }
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nt; b++) {
    double p = 0, F[3] = {0, 0, 0};
    for (const auto &w : waves) {
      double th = 0;
      for (int d = 0; d < 3; d++) th += w.K[d] * t[b].X[d] * scale[d];
      const double dtmp = w.re * std::sin(th) - w.im * std::cos(th);
      p += w.re * std::cos(th) + w.im * std::sin(th);
      for (int d = 0; d < 3; d++) F[d] -= dtmp * w.K[d];
    }
    for (int d = 0; d < 3; d++) F[d] *= scale[d];
    t[b].p += p;
    for (int d = 0; d < 3; d++) t[b].F[d] += F[d];
  }
  // real part
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; i++) {
    for (int64_t j = 0; j < ns; j++) {
      double d[3];
      for (int k = 0; k < 3; k++) {
        d[k] = t[i].X[k] - s[j].X[k];
        double sh = 0;
        if (d[k] < -cycle / 2) sh = -cycle;
        if (d[k] > cycle / 2) sh = cycle;
        d[k] -= sh;
      }
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (0 < R2 && R2 < cutoff * cutoff) {
        const double R2s = R2 * alpha * alpha;
        const double Rs = std::sqrt(R2s);
        const double invRs = 1 / Rs;
        const double invR2s = invRs * invRs;
        const double invR3s = invR2s * invRs;
        double dtmp = s[j].q * (M_2_SQRTPI * std::exp(-R2s) * invR2s + std::erfc(Rs) * invR3s);
        dtmp *= alpha * alpha * alpha;
        t[i].p += s[j].q * std::erfc(Rs) * invRs * alpha;
        t[i].F[0] -= d[0] * dtmp;
        t[i].F[1] -= d[1] * dtmp;
        t[i].F[2] -= d[2] * dtmp;
      }
    }
    t[i].p -= M_2_SQRTPI * t[i].q * alpha;
  }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
