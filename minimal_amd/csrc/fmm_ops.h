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

// Geometry of an M2L translation key: offset d in units of the finer
// radius rf (periodic image applied with the cycle), and the radius
// exponents a, b with R_target = rf 2^a, R_source = rf 2^b.  Key layout:
// (dx+32768)<<48 | (dy+32768)<<32 | (dz+32768)<<16 | image<<11 | level<<6 |
// (level difference + 32).
__device__ inline void key_at(const unsigned long long *keys, int id, double R0, double cycle, double &dx,
                              double &dy, double &dz, int &a, int &b) {
  const unsigned long long k = keys[id];
  dx = double((long long)((k >> 48) & 0xFFFF) - 32768);
  dy = double((long long)((k >> 32) & 0xFFFF) - 32768);
  dz = double((long long)((k >> 16) & 0xFFFF) - 32768);
  const int img = int((k >> 11) & 31), lf = int((k >> 6) & 31);
  const int ld = int(k & 63) - 32;  // level(target) - level(source)
  a = ld < 0 ? -ld : 0;
  b = ld > 0 ? ld : 0;
  if (img != 13) {
    const double rf = R0 / (1 << lf);
    dx -= ((img % 3) - 1) * cycle / rf;
    dy -= (((img / 3) % 3) - 1) * cycle / rf;
    dz -= ((img / 9) - 1) * cycle / rf;
  }
}

// Operator entry T[(j,k,pr)][(n,m,pc)] of one key (harmonics I in LDS).
__device__ inline double op_entry(int P, int fillP, int row, int col, int a, int b, const cpx<double> *I) {
  const int PP = P * P;
  if (row >= PP || col >= PP) return 0.0;
  int j = 0;
  while ((j + 1) * (j + 1) <= row) j++;
  int rem = row - j * j;
  const int k = (rem + 1) / 2;
  const int pr = (rem > 0 && rem % 2 == 0) ? 1 : 0;
  int n = 0;
  while ((n + 1) * (n + 1) <= col) n++;
  rem = col - n * n;
  const int m = (rem + 1) / 2;
  const int pc = (rem > 0 && rem % 2 == 0) ? 1 : 0;
  if (j + n >= fillP) return 0.0;
  const double s = ldexp(1.0, a * (j + 1) + b * n);
  const double cj = sgn(j);
  const double cp = cj * sgn(((k - m) * (k < m)) + m);
  // Gp = I_{j+n}^{m-k} (conjugated for negative order)
  cpx<double> gp = I[nms(j + n, abs(m - k))];
  if (m - k < 0) gp.im = -gp.im;
  gp.re *= s;
  gp.im *= s;
  if (m == 0) {
    if (pc == 1) return 0.0;
    return pr == 0 ? cp * gp.re : cp * gp.im;
  }
  // Gm = conj(I_{j+n}^{m+k})
  cpx<double> gm = I[nms(j + n, m + k)];
  gm.im = -gm.im;
  gm.re *= s;
  gm.im *= s;
  if (pr == 0 && pc == 0) return cp * gp.re + cj * gm.re;
  if (pr == 0 && pc == 1) return -cp * gp.im + cj * gm.im;
  if (pr == 1 && pc == 0) return cp * gp.im + cj * gm.im;
  return cp * gp.re - cj * gm.re;
}

}  // namespace xfd
