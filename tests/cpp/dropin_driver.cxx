// A driver written against the reference's 3d/ interface (the call sequence
// of 3d/fmm.cxx:40-85), compiled against include/exafmm_mi355x.hpp.
// Prints the relative L2 errors of p and F against the direct sum.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "exafmm_mi355x.hpp"
using namespace exafmm;

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 10000;
  P = 10;
  ncrit = 64;
  theta = 0.4;
  Bodies bodies(n);
  srand48(0);
  double avg = 0;
  for (auto &b : bodies) {
    for (int d = 0; d < 3; d++) b.X[d] = drand48() * 2 * M_PI - M_PI;
    b.q = drand48() - .5;
    avg += b.q;
    b.p = 0;
    for (int d = 0; d < 3; d++) b.F[d] = 0;
  }
  avg /= n;
  for (auto &b : bodies) b.q -= avg;
  try {
    Cells cells = buildTree(bodies);
    initKernel();
    upwardPass(cells);
    horizontalPass(cells, cells);
    downwardPass(cells);
    const int nt = 100;
    Bodies jbodies = bodies;
    const int stride = n / nt;
    Bodies t(nt), fmm(nt);
    for (int i = 0; i < nt; i++) {
      fmm[i] = bodies[i * stride];
      t[i] = fmm[i];
      t[i].p = 0;
      t[i].F[0] = t[i].F[1] = t[i].F[2] = 0;
    }
    direct(t, jbodies);
    double pd = 0, pn = 0, fd = 0, fn = 0;
    for (int i = 0; i < nt; i++) {
      pd += (t[i].p - fmm[i].p) * (t[i].p - fmm[i].p);
      pn += t[i].p * t[i].p;
      for (int d = 0; d < 3; d++) {
        fd += (t[i].F[d] - fmm[i].F[d]) * (t[i].F[d] - fmm[i].F[d]);
        fn += t[i].F[d] * t[i].F[d];
      }
    }
    std::printf("cells %zu  M[0][0] %.6e  err_p %.6e  err_F %.6e\n", cells.size(), cells[0].M[0].real(),
                std::sqrt(pd / pn), std::sqrt(fd / fn));
    // A time-stepping driver rebuilds its tree every step: each step's Cells
    // releases its device context, and Cell::L holds the M2L sums as soon as
    // horizontalPass returns (the reference's M2L writes Ci->L there).
    double lsum = 0;
    for (int step = 0; step < 4; step++) {
      Cells c2 = buildTree(bodies);
      upwardPass(c2);
      horizontalPass(c2, c2);
      for (const Cell &c : c2)
        for (const auto &v : c.L) lsum += std::abs(v);
      downwardPass(c2);
    }
    std::printf("live_contexts %d  L_after_horizontal %.6e\n", int(detail::live_contexts()), lsum);
    // Two trees: the first half of the bodies as targets of all the bodies.
    Bodies ib(bodies.begin(), bodies.begin() + n / 2), jb = bodies;
    for (auto &b : ib) {
      b.p = 0;
      b.F[0] = b.F[1] = b.F[2] = 0;
    }
    Cells ic = buildTree(ib), jc = buildTree(jb);
    upwardPass(ic);
    upwardPass(jc);
    horizontalPass(ic, jc);
    downwardPass(ic);
    Bodies it(ib.begin(), ib.begin() + 50), id = it;
    for (auto &b : id) {
      b.p = 0;
      b.F[0] = b.F[1] = b.F[2] = 0;
    }
    direct(id, jb);
    double qd = 0, qn = 0;
    for (int i = 0; i < 50; i++) {
      qd += (id[i].p - it[i].p) * (id[i].p - it[i].p);
      qn += id[i].p * id[i].p;
    }
    std::printf("two_trees_err_p %.6e\n", std::sqrt(qd / qn));
  } catch (const std::exception &e) {
    std::printf("error: %s\n", e.what());
    return 3;
  }
  return 0;
}
