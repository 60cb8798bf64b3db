// fmm_ops.h -- per-coefficient FP64 translation sums shared by the pass
// kernels and the single-call operator entry points.
#pragma once
#include "fmm_internal.h"

namespace xfd {

__device__ inline void jk_of(int jk, int &j, int &k) {
  j = 0;
  while (nms(j + 1, 0) <= jk) j++;
  k = jk - nms(j, 0);
}

__device__ inline cpx<double> yget(const cpx<double> *Y, int n, int m) {  // Y_n^m, any sign of m
  if (m >= 0) return Y[nms(n, m)];
  const cpx<double> v = Y[nms(n, -m)];
  return {v.re, -v.im};
}

// M2M term (3d/kernel.h:211-228) for output (j,k): child multipole Mc,
// Y = regular harmonics of Xparent - Xchild.
__device__ inline double2 m2m_coef(int j, int k, const double2 *Mc, const cpx<double> *Y) {
  double accr = 0, acci = 0;
  for (int n = 0; n <= j; n++) {
    for (int m = max(-n, -j + k + n); m <= min(k - 1, n); m++) {
      const double2 a = Mc[nms(j - n, 0) + k - m];
      const cpx<double> y = yget(Y, n, -m);
      const double s = double((m >= 0 ? 1 : sgn(m)) * sgn(n));
      accr += (a.x * y.re - a.y * y.im) * s;
      acci += (a.x * y.im + a.y * y.re) * s;
    }
    for (int m = k; m <= min(n, j + k - n); m++) {
      const double2 a = Mc[nms(j - n, 0) - k + m];
      const cpx<double> y = yget(Y, n, -m);
      const double s = double(sgn(k + n + m));
      accr += (a.x * y.re + a.y * y.im) * s;  // conj(a) * y
      acci += (a.x * y.im - a.y * y.re) * s;
    }
  }
  return make_double2(accr, acci);
}

// L2L term (3d/kernel.h:269-288) for output (j,k): parent local Lp,
// Y = regular harmonics of Xchild - Xparent.
__device__ inline double2 l2l_coef(int P, int j, int k, const double2 *Lp, const cpx<double> *Y) {
  double accr = 0, acci = 0;
  for (int n = j; n < P; n++) {
    for (int m = j + k - n; m < 0; m++) {
      const double2 a = Lp[nms(n, 0) - m];
      const cpx<double> y = yget(Y, n - j, m - k);
      const double s = double(sgn(k));
      accr += (a.x * y.re + a.y * y.im) * s;  // conj(a) * y
      acci += (a.x * y.im - a.y * y.re) * s;
    }
    for (int m = 0; m <= n; m++) {
      if (n - j >= abs(m - k)) {
        const double2 a = Lp[nms(n, 0) + m];
        const cpx<double> y = yget(Y, n - j, m - k);
        const double s = double(sgn((m - k) * (m < k)));
        accr += (a.x * y.re - a.y * y.im) * s;
        acci += (a.x * y.im + a.y * y.re) * s;
      }
    }
  }
  return make_double2(accr, acci);
}

// M2L term (3d/kernel.h:239-258) for output (j,k): source multipole Mj,
// I = singular harmonics of Xi - Xj filled for degrees < fillP.
__device__ inline double2 m2l_coef(int P, int fillP, int j, int k, const double2 *Mj, const cpx<double> *I) {
  const double cj = sgn(j);
  double sr = 0, si = 0;
  for (int n = 0; n < P; n++) {
    if (j + n >= fillP) break;
    for (int m = -n; m <= n; m++) {
      const int ord = m - k;
      cpx<double> y = I[nms(j + n, abs(ord))];
      if (ord < 0) y.im = -y.im;
      double2 a = Mj[nms(n, abs(m))];
      double s;
      if (m < 0) {
        a.y = -a.y;  // conj(M_n^{|m|})
        s = cj;
      } else {
        s = cj * sgn((k - m) * (k < m) + m);
      }
      sr += (a.x * y.re - a.y * y.im) * s;
      si += (a.x * y.im + a.y * y.re) * s;
    }
  }
  return make_double2(sr, si);
}

// Column m of the singular harmonics of (dx,dy,dz) for degrees m..fillP-1.
__device__ inline void singular_column(int m, int fillP, double dx, double dy, double dz, cpx<double> *I) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double inv = 1.0 / r2;
  double ir = 1.0 / sqrt(r2), ii = 0;
  for (int k = 1; k <= m; k++) {
    const double f = double(2 * k - 1) * inv;
    const double tr = (ir * dx - ii * dy) * f, ti = (ir * dy + ii * dx) * f;
    ir = tr;
    ii = ti;
  }
  I[nms(m, m)] = {ir, ii};
  double ar = ir, ai = ii, br = 0, bi = 0;
  for (int n = m; n + 1 < fillP; n++) {
    const double cz = double(2 * n + 1) * dz, d = double((n + m) * (n - m));
    const double nr = -(cz * ar + d * br) * inv, ni = -(cz * ai + d * bi) * inv;
    I[nms(n + 1, m)] = {nr, ni};
    br = ar;
    bi = ai;
    ar = nr;
    ai = ni;
  }
}

}  // namespace xfd
