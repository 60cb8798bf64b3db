// ref_shim3dp.cpp -- C entry points onto the reference's own 3dp/ headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (-I<reference>/3dp, lazy traversal),
// output only into oracle/_ref/.  Every result comes from the reference's
// buildTree, kernels, periodic passes and Ewald routines.
#define EXAFMM_LAZY 1
#include "build_tree.h"
#include "kernel.h"
#include "ewald.h"
#include "traverse_lazy.h"

#include <cstdint>
#include <cstring>

#define API extern "C" __attribute__((visibility("default")))

using namespace exafmm;

static_assert(sizeof(Body) == 64, "Body layout");

// Periodic FMM as 3dp/fmm.cxx:46-61 (no dipole correction).  Bodies are
// permuted in place and receive p, F.
API int ref3dp_fmm(Body *b, int64_t n, int p, int nc, double th, int img, double cyc) {
  P = p;
  ncrit = nc;
  theta = th;
  images = img;
  cycle = cyc;
  Bodies bodies(n);
  std::memcpy(bodies.data(), b, sizeof(Body) * n);
  Cells cells = buildTree(bodies);
  initKernel();
  upwardPass(cells);
  horizontalPass(cells, cells);
  downwardPass(cells);
  std::memcpy(b, bodies.data(), sizeof(Body) * n);
  return int(cells.size());
}

// Ewald sum as 3dp/fmm.cxx:79-94: targets and sources are the same body
// set (already in tree order); p, F of `b` are overwritten.
API void ref3dp_ewald(Body *b, int64_t n, int nc, int ks, double al, double sg, double cut, double cyc) {
  ncrit = nc;
  ksize = ks;
  alpha = al;
  sigma = sg;
  cutoff = cut;
  cycle = cyc;
  Bodies bodies(n);
  std::memcpy(bodies.data(), b, sizeof(Body) * n);
  for (auto &B : bodies) {
    B.p = 0;
    B.F[0] = B.F[1] = B.F[2] = 0;
  }
  Bodies jbodies = bodies;
  Cells cells = buildTree(bodies);
  Cells jcells = buildTree(jbodies);
  wavePart(bodies, jbodies);
  realPart(&cells[0], &jcells[0]);
  selfTerm(bodies);
  std::memcpy(b, bodies.data(), sizeof(Body) * n);
}
