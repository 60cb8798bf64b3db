// expansions.hip -- P2M, M2M (upward), L2L, L2P (downward).
//
// Reference operators: P2M 3d/kernel.h:187-202, M2M :204-231, L2L :262-291,
// L2P :293-322; passes 3d/traverse_lazy.h:7-24 (upward) and :65-80
// (downward).  Harmonics are evaluated with the trig-free Cartesian
// recurrences of fmm_internal.h (same functions, same normalisation).
// P2M, M2M and L2L run in FP64; L2P evaluates the local expansion in FP32
// on radius-scaled coordinates and adds into the FP64 body accumulators,
// with the force taken from Cartesian ladder identities of the solid
// harmonics (finite at r = 0 and on the z axis, where the reference's
// spherical form divides by zero, 3d/kernel.h:31-40).
#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

// ---------------------------------------------------------------- P2M
// One thread per (leaf, order m): the column n = m..P-1 of
// conj(Y_n^m(x_b - X_c)) summed over the leaf's bodies with weight q
// (3d/kernel.h:193 evaluates Y at -beta, i.e. the conjugate).
__global__ void __launch_bounds__(256) k_p2m(int P, const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                             const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                             const double4 *__restrict__ cx, const double4 *__restrict__ xs,
                                             double2 *__restrict__ M, int NT) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const int li = g / P, m = g % P;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const int beg = cbeg[c], cnt = ccnt[c];
  const double4 C = cx[c];
  double ar[kMaxP], ai[kMaxP];
  for (int n = 0; n < P; n++) ar[n] = ai[n] = 0;
  for (int b = beg; b < beg + cnt; b++) {
    const double4 B = xs[b];
    const double x = B.x - C.x, y = -(B.y - C.y), z = B.z - C.z;  // conjugate: y -> -y
    const double r2 = x * x + y * y + z * z;
    double yr = 1, yi = 0;
    for (int k = 1; k <= m; k++) {
      const double inv = 1.0 / (2 * k);
      const double tr = (yr * x - yi * y) * inv, ti = (yr * y + yi * x) * inv;
      yr = tr;
      yi = ti;
    }
    double pr = 0, pi = 0;  // Y_{n-2}
    double cr = yr, ci = yi;
    ar[m] += B.w * cr;
    ai[m] += B.w * ci;
    for (int n = m + 1; n < P; n++) {
      double nr, ni;
      if (n == m + 1) {
        nr = -z * cr;
        ni = -z * ci;
      } else {
        const double s = -1.0 / double((n + m) * (n - m));
        const double cz = double(2 * n - 1) * z;
        nr = (cz * cr + r2 * pr) * s;
        ni = (cz * ci + r2 * pi) * s;
      }
      pr = cr;
      pi = ci;
      cr = nr;
      ci = ni;
      ar[n] += B.w * cr;
      ai[n] += B.w * ci;
    }
  }
  double2 *Mc = M + int64_t(c) * NT;
  for (int n = m; n < P; n++) Mc[nms(n, m)] = make_double2(ar[n], ai[n]);
}

// Regular harmonics of the 8 parent<->child shift vectors of one level
// into LDS: Y[o][nms], shift = sgn * (+/-r) per axis by octant o.
__device__ inline void octant_tables(int P, int NT, double r, double sgn, cpx<double> *Y) {
  const int t = threadIdx.x;
  if (t < 8) {
    const double x = sgn * ((t & 1) ? r : -r), y = sgn * ((t & 2) ? r : -r), z = sgn * ((t & 4) ? r : -r);
    regular_all<double>(P, x, y, z, Y + t * NT);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- M2M
// One workgroup per parent at level l; thread = coefficient (j,k).  The
// sum over children in child order and the (n,m) loops are the
// reference's (3d/kernel.h:204-231) with dX = Xparent - Xchild.
__global__ void __launch_bounds__(64) k_m2m(int P, int NT, int lstart, int lcount, const int *__restrict__ cchild,
                                            const int *__restrict__ cnch, const double4 *__restrict__ cx,
                                            double2 *__restrict__ M) {
  extern __shared__ cpx<double> Ytab[];  // [8][NT]
  const int pidx = lstart + blockIdx.x;
  if (blockIdx.x >= lcount) return;
  const int nch = cnch[pidx];
  if (nch == 0) return;
  const double4 Cp = cx[pidx];
  const int first = cchild[pidx];
  octant_tables(P, NT, cx[first].w, -1.0, Ytab);
  for (int jk = threadIdx.x; jk < NT; jk += 64) {
    int j, k;
    jk_of(jk, j, k);
    double accr = 0, acci = 0;
    for (int ch = first; ch < first + nch; ch++) {
      const double4 Cc = cx[ch];
      const int o = (Cc.x > Cp.x) + ((Cc.y > Cp.y) << 1) + ((Cc.z > Cp.z) << 2);
      const double2 v = m2m_coef(j, k, M + int64_t(ch) * NT, Ytab + o * NT);
      accr += v.x;
      acci += v.y;
    }
    M[int64_t(pidx) * NT + jk] = make_double2(accr, acci);
  }
}

// ---------------------------------------------------------------- L2L
// One workgroup per parent at level l; every child gets
// L_child[j,k] += sum over n>=j of the parent's terms (3d/kernel.h:262-291)
// with dX = Xchild - Xparent.
__global__ void __launch_bounds__(64) k_l2l(int P, int NT, int lstart, int lcount, const int *__restrict__ cchild,
                                            const int *__restrict__ cnch, const double4 *__restrict__ cx,
                                            double2 *__restrict__ L) {
  extern __shared__ cpx<double> Ytab[];
  const int pidx = lstart + blockIdx.x;
  if (blockIdx.x >= lcount) return;
  const int nch = cnch[pidx];
  if (nch == 0) return;
  const double4 Cp = cx[pidx];
  const int first = cchild[pidx];
  octant_tables(P, NT, cx[first].w, 1.0, Ytab);
  const double2 *Lp = L + int64_t(pidx) * NT;
  for (int e = threadIdx.x; e < nch * NT; e += 64) {
    const int ch = first + e / NT;
    const int jk = e % NT;
    int j, k;
    jk_of(jk, j, k);
    const double4 Cc = cx[ch];
    const int o = (Cc.x > Cp.x) + ((Cc.y > Cp.y) << 1) + ((Cc.z > Cp.z) << 2);
    const double2 v = l2l_coef(P, j, k, Lp, Ytab + o * NT);
    const double accr = v.x, acci = v.y;
    double2 *Lc = L + int64_t(ch) * NT + jk;
    *Lc = make_double2(Lc->x + accr, Lc->y + acci);
  }
}

// ---------------------------------------------------------------- L2P
// One workgroup per leaf, one thread per body.  With u = (x - Xc)/R and
// Lh_n^m = L_n^m R^n, phi = Re sum_{n, m>=0} w_m Lh_n^m Y_n^m(u)
// (w_0 = 1, w_m = 2, 3d/kernel.h:302-316).  The gradient uses
//   dz Y_n^m = -Y_{n-1}^m,  (dx + i dy) Y_n^m = -Y_{n-1}^{m+1},
//   (dx - i dy) Y_n^m = Y_{n-1}^{m-1} (m >= 1), = -conj(Y_{n-1}^1) (m = 0),
// collected per harmonic Y_n^m so each column m is generated once.
template <int MAXP>
__global__ void __launch_bounds__(64) k_l2p(int P, int NT, const int *__restrict__ leaves,
                                            const int *__restrict__ nleaves, const int *__restrict__ cbeg,
                                            const int *__restrict__ ccnt, const double4 *__restrict__ cx,
                                            const double2 *__restrict__ L, const double4 *__restrict__ xs,
                                            exafmm_body *__restrict__ bodies) {
  __shared__ float2 Lh[MAXP * (MAXP + 1) / 2];
  const int li = blockIdx.x;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const double4 C = cx[c];
  for (int e = threadIdx.x; e < NT; e += 64) {
    int n = 0;
    while (nms(n + 1, 0) <= e) n++;
    const double s = pow(C.w, double(n)) * (e == nms(n, 0) ? 1.0 : 2.0);  // fold w_m
    const double2 v = L[int64_t(c) * NT + e];
    Lh[e] = make_float2(float(v.x * s), float(v.y * s));
  }
  __syncthreads();
  const int beg = cbeg[c], cnt = ccnt[c];
  const double invR = 1.0 / C.w;
  for (int b = beg + threadIdx.x; b < beg + cnt; b += 64) {
    const double4 B = xs[b];
    const float x = float((B.x - C.x) * invR), y = float((B.y - C.y) * invR), z = float((B.z - C.z) * invR);
    const float r2 = x * x + y * y + z * z;
    float pot = 0, Ar = 0, Ai = 0, Br = 0, Bi = 0, Dz = 0;
    float ymr = 1, ymi = 0;  // Y_m^m
    for (int m = 0; m < P; m++) {
      if (m > 0) {
        const float inv = 1.0f / float(2 * m);
        const float tr = (ymr * x - ymi * y) * inv, ti = (ymr * y + ymi * x) * inv;
        ymr = tr;
        ymi = ti;
      }
      float pr = 0, pi = 0, cr = ymr, ci = ymi;
      for (int n = m; n < P; n++) {
        if (n > m) {
          float nr, ni;
          if (n == m + 1) {
            nr = -z * cr;
            ni = -z * ci;
          } else {
            const float s = -1.0f / float((n + m) * (n - m));
            const float cz = float(2 * n - 1) * z;
            nr = (cz * cr + r2 * pr) * s;
            ni = (cz * ci + r2 * pi) * s;
          }
          pr = cr;
          pi = ci;
          cr = nr;
          ci = ni;
        }
        // potential: Re(Lh_n^m Y_n^m)  (weights folded into Lh)
        const float2 l0 = Lh[nms(n, m)];
        pot += l0.x * cr - l0.y * ci;
        if (n + 1 < P) {
          // dz:  -Lh_{n+1}^m Y_n^m
          const float2 lz = Lh[nms(n + 1, m)];
          Dz -= lz.x * cr - lz.y * ci;
          // (dx+i dy): -Lh_{n+1}^{m-1} Y_n^m  (m >= 1)
          if (m >= 1) {
            const float2 la = Lh[nms(n + 1, m - 1)];
            Ar -= la.x * cr - la.y * ci;
            Ai -= la.x * ci + la.y * cr;
          }
          // (dx-i dy): Lh_{n+1}^{m+1} Y_n^m
          const float2 lb = Lh[nms(n + 1, m + 1)];
          Br += lb.x * cr - lb.y * ci;
          Bi += lb.x * ci + lb.y * cr;
          // m = 1 also feeds (dx-i dy) of the m = 0 terms: -Lh_{n+1}^0 conj(Y_n^1)
          if (m == 1) {
            const float2 l00 = Lh[nms(n + 1, 0)];
            Br -= l00.x * cr + l00.y * ci;
            Bi -= l00.y * cr - l00.x * ci;
          }
        }
      }
    }
    // weights w_m were folded into Lh; the m = 0 terms of Lh carry w = 1
    const double Fx = 0.5 * double(Ar + Br) * invR;
    const double Fy = 0.5 * double(Ai - Bi) * invR;
    const double Fz = double(Dz) * invR;
    exafmm_body *o = bodies + b;
    o->p += double(pot);
    o->F[0] += Fx;
    o->F[1] += Fy;
    o->F[2] += Fz;
  }
}

void upward(Ctx &c) {
  Tree &t = c.t;
  hipStream_t s = c.s0;
  const int P = c.P, NT = c.NT;
  c.M.ensure(size_t(t.ncells) * NT);
  const int64_t nl = t.nleaves;
  hipLaunchKernelGGL(k_p2m, dim3(ceil_div(nl * P, 256)), dim3(256), 0, s, P, t.leaves.p, t.misc.p + 17, t.cbeg.p,
                     t.ccnt.p, t.cx.p, t.xs.p, c.M.p, NT);
  XF_LAUNCH();
  for (int l = t.depth - 1; l >= 0; l--) {
    const int cnt = t.lstart[l + 1] - t.lstart[l];
    hipLaunchKernelGGL(k_m2m, dim3(cnt), dim3(64), sizeof(cpx<double>) * 8 * NT, s, P, NT, t.lstart[l], cnt,
                       t.cchild.p, t.cnch.p, t.cx.p, c.M.p);
    XF_LAUNCH();
  }
  c.have_M = true;
}

void l2l_pass(Ctx &c) {
  Tree &t = c.t;
  hipStream_t s = c.s0;
  const int P = c.P, NT = c.NT;
  for (int l = 0; l < t.depth; l++) {
    const int cnt = t.lstart[l + 1] - t.lstart[l];
    hipLaunchKernelGGL(k_l2l, dim3(cnt), dim3(64), sizeof(cpx<double>) * 8 * NT, s, P, NT, t.lstart[l], cnt,
                       t.cchild.p, t.cnch.p, t.cx.p, c.L.p);
    XF_LAUNCH();
  }
}

void l2p_pass(Ctx &c) {
  Tree &t = c.t;
  hipStream_t s = c.s0;
  const int P = c.P, NT = c.NT;
  hipLaunchKernelGGL(k_l2p<kMaxP>, dim3(t.nleaves), dim3(64), 0, s, P, NT, t.leaves.p, t.misc.p + 17, t.cbeg.p,
                     t.ccnt.p, t.cx.p, c.L.p, t.xs.p, c.bodies);
  XF_LAUNCH();
}

}  // namespace xf
