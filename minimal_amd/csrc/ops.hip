// ops.hip -- the kernel.h operator surface: one reference kernel call per
// entry (P2P, P2M, M2M, M2L, L2L, L2P of 3d/kernel.h:122-322), in FP64 on
// the device, for callers that drive the operators themselves.
#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

namespace {

template <class T>
struct DevCopy {  // scoped device copy of a small host array
  Buf<T> b;
  DevCopy(const T *h, size_t n, hipStream_t s) {
    b.ensure(n);
    if (h && n) XF_HIP(hipMemcpyAsync(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
  }
  void back(T *h, size_t n, hipStream_t s) {
    XF_HIP(hipMemcpyAsync(h, b.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    XF_HIP(hipStreamSynchronize(s));
  }
};

// P2M (3d/kernel.h:187-202) on one cell: thread m builds column m of the
// conjugated regular harmonics of every body and accumulates q * Y.
__global__ void k_op_p2m(int P, double cx, double cy, double cz, const exafmm_body *b, int64_t nb, double2 *M) {
  const int m = threadIdx.x;
  if (m >= P) return;
  double ar[kMaxP], ai[kMaxP];
  for (int n = 0; n < P; n++) ar[n] = ai[n] = 0;
  for (int64_t i = 0; i < nb; i++) {
    const double x = b[i].X[0] - cx, y = -(b[i].X[1] - cy), z = b[i].X[2] - cz;
    cpx<double> Y[kMaxP * (kMaxP + 1) / 2];
    regular_all<double>(P, x, y, z, Y);
    for (int n = m; n < P; n++) {
      ar[n] += b[i].q * Y[nms(n, m)].re;
      ai[n] += b[i].q * Y[nms(n, m)].im;
    }
  }
  for (int n = m; n < P; n++) {
    M[nms(n, m)].x += ar[n];
    M[nms(n, m)].y += ai[n];
  }
}

__global__ void k_op_shift(int kind, int P, int NT, double dx, double dy, double dz, int fillP, const double2 *in,
                           double2 *out) {
  __shared__ cpx<double> Y[(2 * kMaxP) * (2 * kMaxP + 1) / 2];
  if (threadIdx.x == 0) {
    if (kind == 2)
      for (int m = 0; m < fillP; m++) singular_column(m, fillP, dx, dy, dz, Y);
    else
      regular_all<double>(P, dx, dy, dz, Y);
  }
  __syncthreads();
  for (int jk = threadIdx.x; jk < NT; jk += blockDim.x) {
    int j, k;
    jk_of(jk, j, k);
    double2 v;
    if (kind == 1)
      v = m2m_coef(j, k, in, Y);
    else if (kind == 2)
      v = m2l_coef(P, fillP, j, k, in, Y);
    else
      v = l2l_coef(P, j, k, in, Y);
    out[jk].x += v.x;
    out[jk].y += v.y;
  }
}

// L2P (3d/kernel.h:293-322) with the FP64 spherical form of the reference
// (one thread per body); used only by the single-call surface.
__global__ void k_op_l2p(int P, double cx, double cy, double cz, const double2 *L, exafmm_body *b, int64_t nb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double x = b[i].X[0] - cx, y = b[i].X[1] - cy, z = b[i].X[2] - cz;
  cpx<double> Y[kMaxP * (kMaxP + 1) / 2];
  regular_all<double>(P, x, y, z, Y);
  double pot = 0, Ar = 0, Ai = 0, Br = 0, Bi = 0, Dz = 0;
  for (int n = 0; n < P; n++)
    for (int m = 0; m <= n; m++) {
      const double w = m ? 2.0 : 1.0;
      const cpx<double> yv = Y[nms(n, m)];
      const double2 l0 = L[nms(n, m)];
      pot += w * (l0.x * yv.re - l0.y * yv.im);
      if (n + 1 < P) {
        const double2 lz = L[nms(n + 1, m)];
        Dz -= w * (lz.x * yv.re - lz.y * yv.im);
        if (m >= 1) {
          const double wa = (m - 1) ? 2.0 : 1.0;
          const double2 la = L[nms(n + 1, m - 1)];
          Ar -= wa * (la.x * yv.re - la.y * yv.im);
          Ai -= wa * (la.x * yv.im + la.y * yv.re);
        }
        const double2 lb = L[nms(n + 1, m + 1)];
        Br += 2.0 * (lb.x * yv.re - lb.y * yv.im);
        Bi += 2.0 * (lb.x * yv.im + lb.y * yv.re);
        if (m == 1) {
          const double2 l00 = L[nms(n + 1, 0)];
          Br -= l00.x * yv.re + l00.y * yv.im;
          Bi -= l00.y * yv.re - l00.x * yv.im;
        }
      }
    }
  b[i].p += pot;
  b[i].F[0] += 0.5 * (Ar + Br);
  b[i].F[1] += 0.5 * (Ai - Bi);
  b[i].F[2] += Dz;
}

}  // namespace

void op_p2p(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns) {
  DevCopy<exafmm_body> dt(t, nt, c.s0), ds(s, ns, c.s0);
  direct_sum(c.cfg.device, c.s0, dt.b.p, nt, ds.b.p, ns);
  dt.back(t, nt, c.s0);
}

void op_p2m(Ctx &c, const double *C, const exafmm_body *b, int64_t nb, double *M) {
  DevCopy<exafmm_body> db(b, nb, c.s0);
  DevCopy<double2> dm(reinterpret_cast<const double2 *>(M), c.NT, c.s0);
  hipLaunchKernelGGL(k_op_p2m, dim3(1), dim3(64), 0, c.s0, c.P, C[0], C[1], C[2], db.b.p, nb, dm.b.p);
  XF_LAUNCH();
  dm.back(reinterpret_cast<double2 *>(M), c.NT, c.s0);
}

static void shift_op(Ctx &c, int kind, const double *d, int fillP, const double *in, double *out) {
  DevCopy<double2> di(reinterpret_cast<const double2 *>(in), c.NT, c.s0);
  DevCopy<double2> dout(reinterpret_cast<const double2 *>(out), c.NT, c.s0);
  hipLaunchKernelGGL(k_op_shift, dim3(1), dim3(128), 0, c.s0, kind, c.P, c.NT, d[0], d[1], d[2], fillP, di.b.p,
                     dout.b.p);
  XF_LAUNCH();
  dout.back(reinterpret_cast<double2 *>(out), c.NT, c.s0);
}

void op_m2m(Ctx &c, const double *Cp, const double *Cc, const double *Mc, double *Mp) {
  const double d[3] = {Cp[0] - Cc[0], Cp[1] - Cc[1], Cp[2] - Cc[2]};
  shift_op(c, 1, d, 0, Mc, Mp);
}

void op_m2l(Ctx &c, const double *Xi, const double *Xj, const double *Mj, double *Li) {
  const double d[3] = {Xi[0] - Xj[0], Xi[1] - Xj[1], Xi[2] - Xj[2]};
  const int fillP = c.cfg.m2l_truncation == EXAFMM_M2L_FULL ? 2 * c.P - 1 : c.P;
  shift_op(c, 2, d, fillP, Mj, Li);
}

void op_l2l(Ctx &c, const double *Cp, const double *Lp, const double *Cc, double *Lc) {
  const double d[3] = {Cc[0] - Cp[0], Cc[1] - Cp[1], Cc[2] - Cp[2]};
  shift_op(c, 3, d, 0, Lp, Lc);
}

void op_l2p(Ctx &c, const double *C, const double *L, exafmm_body *b, int64_t nb) {
  DevCopy<exafmm_body> db(b, nb, c.s0);
  DevCopy<double2> dl(reinterpret_cast<const double2 *>(L), c.NT, c.s0);
  hipLaunchKernelGGL(k_op_l2p, dim3(ceil_div(nb, 64)), dim3(64), 0, c.s0, c.P, C[0], C[1], C[2], dl.b.p, db.b.p,
                     nb);
  XF_LAUNCH();
  db.back(b, nb, c.s0);
}

}  // namespace xf
