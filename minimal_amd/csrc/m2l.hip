// m2l.hip -- multipole-to-local translation.
//
// Reference: M2L 3d/kernel.h:233-260 with singular harmonics 3d/kernel.h:83-116,
// evaluated over each target's M2L list (3d/traverse_lazy.h:46-56).
//
// Two paths:
//  * exact  (m2l_precision = FP64): one wavefront per target cell, lanes =
//    output coefficients (j,k), the reference's double sum per pair on the
//    FP64 singular harmonics of the true offset.  Parity path.
//  * fast   (FP32 / BF16X3): pairs grouped by translation key; every key's
//    radius-scaled P^2 x P^2 real operator is applied to tiles of targets
//    with MFMA (m2l_mfma.hip).
// Truncation: "reference" keeps the terms the reference evaluates (its
// harmonic table is zero above degree P-1, so j+n <= P-1); "full" keeps
// every j+n <= 2P-2.
#include <algorithm>

#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

// Expansions are stored radius-scaled in real form; the pair sums run on the
// unscaled complex coefficients (M = M~ Rs^n in, L~ += L Rc^(j+1) out).
__global__ void __launch_bounds__(64) k_m2l_fp64(int P, int NT, int fillP, int ncells, const double4 *__restrict__ cx,
                                                 const double4 *__restrict__ scx, const int *__restrict__ beg,
                                                 const int *__restrict__ cnt, const int2 *__restrict__ list,
                                                 const double *__restrict__ M, double *__restrict__ L,
                                                 double cycle) {
  extern __shared__ double2 sm[];
  const int c = blockIdx.x;
  if (c >= ncells) return;
  const int n_pairs = cnt[c];
  if (n_pairs == 0) return;
  const int NF = fillP * (fillP + 1) / 2;
  cpx<double> *I = reinterpret_cast<cpx<double> *>(sm);  // [NF]
  double2 *Ms = sm + NF;                                 // [NT]
  const int lane = threadIdx.x;
  const double4 Xc = cx[c];
  const int b0 = beg[c];
  double accr[3] = {0, 0, 0}, acci[3] = {0, 0, 0};
  for (int p = 0; p < n_pairs; p++) {
    const int2 e = list[b0 + p];
    const int src = e.x >> 5, img = e.x & 31;
    const double4 Xs = scx[src];
    // 3dp/kernel.h:200: dX = Xi - Xj - iX * cycle
    const double dx = Xc.x - Xs.x - ((img % 3) - 1) * cycle;
    const double dy = Xc.y - Xs.y - (((img / 3) % 3) - 1) * cycle;
    const double dz = Xc.z - Xs.z - ((img / 9) - 1) * cycle;
    __syncthreads();
    if (lane < fillP) singular_column(lane, fillP, dx, dy, dz, I);
    for (int i = lane; i < NT; i += 64) {
      int n, m;
      jk_of(i, n, m);
      const double *Mc = M + int64_t(src) * P * P;
      const double s = pow(Xs.w, double(n));
      Ms[i] = make_double2(Mc[rf(n, m, 0)] * s, m ? Mc[rf(n, m, 1)] * s : 0.0);
    }
    __syncthreads();
    for (int r = 0; r * 64 + lane < NT; r++) {
      int j, k;
      jk_of(r * 64 + lane, j, k);
      const double2 v = m2l_coef(P, fillP, j, k, Ms, I);
      accr[r] += v.x;
      acci[r] += v.y;
    }
  }
  for (int r = 0; r * 64 + lane < NT; r++) {
    int j, k;
    jk_of(r * 64 + lane, j, k);
    const double s = pow(Xc.w, double(j + 1));
    double *o = L + int64_t(c) * P * P;
    o[rf(j, k, 0)] += accr[r] * s;
    if (k > 0) o[rf(j, k, 1)] += acci[r] * s;
  }
}

void m2l_fast(Ctx &tc, const Ctx &sc);  // m2l_mfma.hip

// Levels whose M2L runs per pair in FP64 beside the keyed GEMM.  Periodic:
// levels 0..2, whose cells collect hundreds of image translations that add
// up to a small difference of large terms (the neutral periodic potential),
// where FP32 accumulation would leave an offset of ~1e-5 relative.
int m2l_fp64_levels(const Ctx &c) { return c.cfg.images > 0 ? std::min(3, c.t.depth + 1) : 0; }

void m2l_apply(Ctx &tc, const Ctx &sc) {
  const int P = tc.P, NT = tc.NT;
  const int fillP = tc.cfg.m2l_truncation == EXAFMM_M2L_FULL ? 2 * P - 1 : P;
  Tree &t = tc.t;
  const size_t lds = sizeof(double2) * (fillP * (fillP + 1) / 2 + NT);
  const double cyc = tc.cfg.images > 0 ? tc.cfg.cycle : 0.0;
  if (tc.cfg.m2l_precision != EXAFMM_PREC_FP64 && tc.lists.nkeys > 0) {
    const int lf = m2l_fp64_levels(tc);
    if (lf > 0) {  // cells of levels < lf come first in breadth-first order
      const int nc = t.lstart[lf];
      hipLaunchKernelGGL(k_m2l_fp64, dim3(nc), dim3(64), lds, tc.s0, P, NT, fillP, nc, t.cx.p, sc.t.cx.p,
                         tc.lists.m2l_beg.p, tc.lists.m2l_cnt.p, tc.lists.m2l.p, sc.M.p, tc.L.p, cyc);
      XF_LAUNCH();
    }
    m2l_fast(tc, sc);
    return;
  }
  hipLaunchKernelGGL(k_m2l_fp64, dim3(t.ncells), dim3(64), lds, tc.s0, P, NT, fillP, int(t.ncells), t.cx.p,
                     sc.t.cx.p, tc.lists.m2l_beg.p, tc.lists.m2l_cnt.p, tc.lists.m2l.p, sc.M.p, tc.L.p, cyc);
  XF_LAUNCH();
}

}  // namespace xf
