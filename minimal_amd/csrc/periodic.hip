// periodic.hip -- far periodic images, dipole correction, Ewald oracle.
//
// Far images (3dp/traverse_lazy.h:75-121, eager 3dp/traverse_eager.h:46-92):
// a super-cell starts as the root (centre X0, multipole M_root); at each of
// images-1 levels the root's local expansion receives M2L from the 26 x 27
// copies at offsets (3i + c) * cycle (i != 0), then the super-cell absorbs
// its 26 neighbours by M2M and cycle triples.  Each level's M2L is one
// launch of 702 workgroups writing partial locals, summed in a fixed order;
// the M2M is one workgroup per neighbour, summed in neighbour order.
//
// Dipole correction (3dp/fmm.cxx:63-75): p -= 4 pi/(3 cycle^3) |D|^2 / N / q,
// F -= 4 pi/(3 cycle^3) D, D = sum q X.  Bodies with q = 0 keep p (the
// reference divides by q).
//
// Ewald sum (3dp/ewald.h:22-150) in FP64 on the device: wave part by
// DFT/IDFT over the half space |k| <= ksize, real part with erfc over each
// body pair's nearest image within the cutoff, self term.
#include <cmath>

#include <mutex>
#include <string>

#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

__global__ void __launch_bounds__(64) k_far_m2l(int P, int NT, int fillP, double cyc, const double2 *__restrict__ Msup,
                                                double2 *__restrict__ part) {
  __shared__ cpx<double> I[(2 * kMaxP) * (2 * kMaxP + 1) / 2];
  __shared__ double2 Ms[kMaxP * (kMaxP + 1) / 2];
  const int e = blockIdx.x;  // 26 neighbour blocks x 27 sub-images, reference loop order
  const int ib = e / 27, cidx = e % 27;
  const int nb = ib < 13 ? ib : ib + 1;  // skip the centre block
  const int ix = nb / 9 - 1, iy = (nb / 3) % 3 - 1, iz = nb % 3 - 1;
  const int cx = cidx / 9 - 1, cy = (cidx / 3) % 3 - 1, cz = cidx % 3 - 1;
  // dX = X_root - X_super - iX * cycle, with X_super = X_root
  const double dx = -(ix * 3 + cx) * cyc, dy = -(iy * 3 + cy) * cyc, dz = -(iz * 3 + cz) * cyc;
  if (threadIdx.x < fillP) singular_column(threadIdx.x, fillP, dx, dy, dz, I);
  for (int i = threadIdx.x; i < NT; i += 64) Ms[i] = Msup[i];
  __syncthreads();
  for (int jk = threadIdx.x; jk < NT; jk += 64) {
    int j, k;
    jk_of(jk, j, k);
    part[int64_t(e) * NT + jk] = m2l_coef(P, fillP, j, k, Ms, I);
  }
}

// Sum of the partial locals, added to the root's scaled L~: one workgroup
// per coefficient, the parts added by a fixed-order tree.
__global__ void __launch_bounds__(256) k_far_sum(int P, int nparts, double R0, const double2 *__restrict__ part,
                                                 double *__restrict__ Lroot) {
  __shared__ double sx[256], sy[256];
  const int NT = P * (P + 1) / 2, jk = blockIdx.x, t = threadIdx.x;
  double ax = 0, ay = 0;
  for (int e = t; e < nparts; e += 256) {
    ax += part[int64_t(e) * NT + jk].x;
    ay += part[int64_t(e) * NT + jk].y;
  }
  sx[t] = ax;
  sy[t] = ay;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      sx[t] += sx[t + o];
      sy[t] += sy[t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    int j, k;
    jk_of(jk, j, k);
    const double s = pow(R0, double(j + 1));
    Lroot[rf(j, k, 0)] += sx[0] * s;
    if (k > 0) Lroot[rf(j, k, 1)] += sy[0] * s;
  }
}

// Root multipole M = M~ R0^n (complex) as the first super-cell.
__global__ void __launch_bounds__(64) k_far_init(int P, double R0, const double *__restrict__ Mroot,
                                                 double2 *__restrict__ sup) {
  const int NT = P * (P + 1) / 2;
  for (int jk = threadIdx.x; jk < NT; jk += 64) {
    int n, m;
    jk_of(jk, n, m);
    const double s = pow(R0, double(n));
    sup[jk] = make_double2(Mroot[rf(n, m, 0)] * s, m ? Mroot[rf(n, m, 1)] * s : 0.0);
  }
}

// Msup += sum over the 26 neighbours (child centre X + i*cycle) of
// M2M(Msup): one workgroup per neighbour writes its term, then the terms
// are added in neighbour order.
__global__ void __launch_bounds__(64) k_far_m2m(int P, int NT, double cyc, const double2 *__restrict__ Msup,
                                                double2 *__restrict__ part) {
  __shared__ cpx<double> Y[kMaxP * (kMaxP + 1) / 2];
  __shared__ double2 Mc[kMaxP * (kMaxP + 1) / 2];
  const int b = blockIdx.x, nb = b < 13 ? b : b + 1;  // skip the centre
  const int ix = nb / 9 - 1, iy = (nb / 3) % 3 - 1, iz = nb % 3 - 1;
  for (int i = threadIdx.x; i < NT; i += 64) Mc[i] = Msup[i];
  if (threadIdx.x == 0) regular_all<double>(P, -ix * cyc, -iy * cyc, -iz * cyc, Y);  // Xparent - Xchild
  __syncthreads();
  for (int jk = threadIdx.x; jk < NT; jk += 64) {
    int j, k;
    jk_of(jk, j, k);
    part[int64_t(b) * NT + jk] = m2m_coef(j, k, Mc, Y);
  }
}

__global__ void __launch_bounds__(64) k_far_m2m_sum(int NT, const double2 *__restrict__ part,
                                                    double2 *__restrict__ Msup) {
  for (int jk = threadIdx.x; jk < NT; jk += 64) {
    double2 acc = make_double2(0, 0);
    for (int b = 0; b < 26; b++) {
      acc.x += part[int64_t(b) * NT + jk].x;
      acc.y += part[int64_t(b) * NT + jk].y;
    }
    Msup[jk].x += acc.x;
    Msup[jk].y += acc.y;
  }
}

void periodic_far(Ctx &c) {
  const int P = c.P, NT = c.NT;
  const int fillP = c.cfg.m2l_truncation == EXAFMM_M2L_FULL ? 2 * P - 1 : P;
  if (c.cfg.images < 2) return;
  Buf<double2> &sup = c.far_sup, &part = c.far_part;
  sup.ensure(NT);
  part.ensure(size_t(702) * NT);
  hipLaunchKernelGGL(k_far_init, dim3(1), dim3(64), 0, c.s0, P, c.t.R0, c.M.p, sup.p);
  XF_LAUNCH();
  double cyc = c.cfg.cycle;
  for (int lev = 0; lev < c.cfg.images - 1; lev++) {
    hipLaunchKernelGGL(k_far_m2l, dim3(702), dim3(64), 0, c.s0, P, NT, fillP, cyc, sup.p, part.p);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_far_sum, dim3(NT), dim3(256), 0, c.s0, P, 702, c.t.R0, part.p, c.L.p);
    XF_LAUNCH();
    // the 26 neighbour terms reuse the partial-local buffer (read above)
    hipLaunchKernelGGL(k_far_m2m, dim3(26), dim3(64), 0, c.s0, P, NT, cyc, sup.p, part.p);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_far_m2m_sum, dim3(1), dim3(64), 0, c.s0, NT, part.p, sup.p);
    XF_LAUNCH();
    cyc *= 3;
  }
}

// ---------------------------------------------------------------- dipole
// Dipole moment sum_i q_i x_i: kDipoleBlocks workgroups sum grid-strided
// ranges, then one workgroup adds their partials in a fixed order.
constexpr int kDipoleBlocks = 256;
__device__ inline void block_sum3(double (&d)[3], double (*s)[256]) {
  for (int k = 0; k < 3; k++) s[k][threadIdx.x] = d[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_dipole(int64_t n, const exafmm_body *__restrict__ b, double *__restrict__ part) {
  __shared__ double s[3][256];
  double d[3] = {0, 0, 0};
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    for (int k = 0; k < 3; k++) d[k] += b[i].X[k] * b[i].q;
  block_sum3(d, s);
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_dipole_final(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double s[3][256];
  double d[3] = {0, 0, 0};
  for (int k = 0; k < 3; k++) d[k] = part[threadIdx.x * 3 + k];
  block_sum3(d, s);
  if (threadIdx.x < 3) out[threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_dipole_apply(int64_t n, double coef, const double *__restrict__ D,
                                                      exafmm_body *__restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double dn = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  if (b[i].q != 0) b[i].p -= coef * dn / double(n) / b[i].q;
  for (int k = 0; k < 3; k++) b[i].F[k] -= coef * D[k];
}

void dipole_correction(Ctx &c) {
  c.dipole_part.ensure(3 * kDipoleBlocks + 3);
  double *D = c.dipole_part.p + 3 * kDipoleBlocks;
  hipLaunchKernelGGL(k_dipole, dim3(kDipoleBlocks), dim3(256), 0, c.s0, c.t.n, c.bodies, c.dipole_part.p);
  XF_LAUNCH();
  hipLaunchKernelGGL(k_dipole_final, dim3(1), dim3(kDipoleBlocks), 0, c.s0, c.dipole_part.p, D);
  XF_LAUNCH();
  const double cyc = c.cfg.cycle;
  const double coef = 4 * M_PI / (3 * cyc * cyc * cyc);
  hipLaunchKernelGGL(k_dipole_apply, dim3(ceil_div(c.t.n, 256)), dim3(256), 0, c.s0, c.t.n, coef, D, c.bodies);
  XF_LAUNCH();
}

// ---------------------------------------------------------------- Ewald
struct Wave {
  double K[3];
  double re, im;
};

__global__ void __launch_bounds__(256) k_ewald_dft(int nw, Wave *__restrict__ waves, const exafmm_body *__restrict__ s,
                                                   int64_t ns, double scale, double coef, double coef2) {
  __shared__ double red[2][256];
  const int w = blockIdx.x;
  if (w >= nw) return;
  const double K0 = waves[w].K[0], K1 = waves[w].K[1], K2 = waves[w].K[2];
  double re = 0, im = 0;
  for (int64_t b = threadIdx.x; b < ns; b += 256) {
    const double th = K0 * s[b].X[0] * scale + K1 * s[b].X[1] * scale + K2 * s[b].X[2] * scale;
    double sn, cs;
    sincos(th, &sn, &cs);
    re += s[b].q * cs;
    im += s[b].q * sn;
  }
  red[0][threadIdx.x] = re;
  red[1][threadIdx.x] = im;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double k0 = K0 * scale, k1 = K1 * scale, k2 = K2 * scale;
    const double KK = k0 * k0 + k1 * k1 + k2 * k2;
    const double factor = coef * exp(-KK * coef2) / KK;
    waves[w].re = red[0][0] * factor;
    waves[w].im = red[1][0] * factor;
  }
}

__global__ void __launch_bounds__(256) k_ewald_target(int nw, const Wave *__restrict__ waves, exafmm_body *__restrict__ t,
                                                      int64_t nt, const exafmm_body *__restrict__ s, int64_t ns,
                                                      double scale, double alpha, double cutoff, double cyc) {
  __shared__ double red[4][256];
  const int64_t ti = blockIdx.x;
  if (ti >= nt) return;
  const double x = t[ti].X[0], y = t[ti].X[1], z = t[ti].X[2];
  double p = 0, F[3] = {0, 0, 0};
  // wave part (3dp/ewald.h:36-51)
  for (int w = threadIdx.x; w < nw; w += 256) {
    const double th = waves[w].K[0] * x * scale + waves[w].K[1] * y * scale + waves[w].K[2] * z * scale;
    double sn, cs;
    sincos(th, &sn, &cs);
    const double dtmp = waves[w].re * sn - waves[w].im * cs;
    p += waves[w].re * cs + waves[w].im * sn;
    for (int d = 0; d < 3; d++) F[d] -= dtmp * waves[w].K[d] * scale;
  }
  // real part (3dp/ewald.h:82-102), nearest image of each pair
  for (int64_t j = threadIdx.x; j < ns; j += 256) {
    double d[3] = {x - s[j].X[0], y - s[j].X[1], z - s[j].X[2]};
    for (int k = 0; k < 3; k++) {
      if (d[k] < -cyc / 2) d[k] += cyc;
      if (d[k] > cyc / 2) d[k] -= cyc;
    }
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (0 < R2 && R2 < cutoff * cutoff) {
      const double R2s = R2 * alpha * alpha;
      const double Rs = sqrt(R2s);
      const double invRs = 1 / Rs;
      const double invR2s = invRs * invRs;
      const double invR3s = invR2s * invRs;
      const double ec = erfc(Rs);
      double dtmp = s[j].q * (M_2_SQRTPI * exp(-R2s) * invR2s + ec * invR3s);
      dtmp *= alpha * alpha * alpha;
      p += s[j].q * ec * invRs * alpha;
      for (int k = 0; k < 3; k++) F[k] -= d[k] * dtmp;
    }
  }
  red[0][threadIdx.x] = p;
  for (int k = 0; k < 3; k++) red[1 + k][threadIdx.x] = F[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    t[ti].p += red[0][0] - M_2_SQRTPI * t[ti].q * alpha;  // self term (3dp/ewald.h:146-150)
    for (int k = 0; k < 3; k++) t[ti].F[k] += red[1 + k][0];
  }
}

// Wave vectors k in the half space (3dp/ewald.h:54-79): l = 0..kmax,
// m = -kmax..kmax (0..kmax at l = 0), n = -kmax..kmax (1..kmax at l = m = 0),
// l^2 + m^2 + n^2 <= kmax^2, in that nested order.
static int ewald_wave_count(int kmax) {
  int nw = 0;
  const int kmaxsq = kmax * kmax;
  for (int l = 0; l <= kmax; l++) {
    const int mmin = l == 0 ? 0 : -kmax;
    for (int m = mmin; m <= kmax; m++) {
      const int nmin = (l == 0 && m == 0) ? 1 : -kmax;
      for (int n = nmin; n <= kmax; n++)
        if (l * l + m * m + n * n <= kmaxsq) nw++;
    }
  }
  return nw;
}

static void ewald_launch(Ctx &c, Wave *w, int nw, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns,
                         double alpha, double sigma, double cutoff, double cyc) {
  const double scale = 2 * M_PI / cyc;
  const double coef = 2 / sigma / cyc / cyc / cyc;
  const double coef2 = 1 / (4 * alpha * alpha);
  if (nw) {
    hipLaunchKernelGGL(k_ewald_dft, dim3(nw), dim3(256), 0, c.s0, nw, w, s, ns, scale, coef, coef2);
    XF_LAUNCH();
  }
  if (nt) {
    hipLaunchKernelGGL(k_ewald_target, dim3(nt), dim3(256), 0, c.s0, nw, w, t, nt, s, ns, scale, alpha, cutoff, cyc);
    XF_LAUNCH();
  }
  XF_HIP(hipStreamSynchronize(c.s0));
}

void ewald_sum(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns, int ksize, double alpha,
               double sigma, double cutoff, double cyc) {
  std::vector<Wave> w;
  const int kmax = ksize, kmaxsq = ksize * ksize;
  for (int l = 0; l <= kmax; l++) {
    const int mmin = l == 0 ? 0 : -kmax;
    for (int m = mmin; m <= kmax; m++) {
      const int nmin = (l == 0 && m == 0) ? 1 : -kmax;
      for (int n = nmin; n <= kmax; n++)
        if (l * l + m * m + n * n <= kmaxsq) w.push_back({{double(l), double(m), double(n)}, 0, 0});
    }
  }
  const int nw = int(w.size());
  Buf<Wave> dw;
  dw.ensure(std::max(nw, 1));
  if (nw) XF_HIP(hipMemcpyAsync(dw.p, w.data(), sizeof(Wave) * nw, hipMemcpyHostToDevice, c.s0));
  ewald_launch(c, dw.p, nw, t, nt, s, ns, alpha, sigma, cutoff, cyc);
}

// ---------------------------------------------------------------- Ewald without allocation
// The wave table of exafmm_ewald_noalloc lives in this fixed device array
// (one copy per GPU, kept until the process ends).  Host code never passes
// the array's name to a kernel: it uses the address hipGetSymbolAddress
// returned for the current device (ewald_scratch).
static __device__ __align__(256) Wave scratch[kEwaldMaxWaves];

namespace {
std::mutex g_scratch_mu;
Wave *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void ewald_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) fail(EXAFMM_ERR_ARG, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  XF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<Wave *>(p);
}

static Wave *ewald_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Wave *p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
  if (!p) fail(EXAFMM_ERR_STATE, "scratch of the device not initialised (no context created on it)");
  return p;
}

// The wave table on the device, in ewald_wave_count's order: one
// workgroup walks the candidate box (l, m, n) 256 at a time and compacts
// the valid vectors with a ballot prefix.
__global__ void __launch_bounds__(256) k_ewald_waves(int kmax, Wave *__restrict__ w) {
  __shared__ int wave_tot[4];
  __shared__ int base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int side = 2 * kmax + 1;
  const int64_t ncand = int64_t(kmax + 1) * side * side;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < ncand; c0 += 256) {
    const int64_t cnd = c0 + threadIdx.x;
    const int l = int(cnd / (side * side)), m = int((cnd / side) % side) - kmax, n = int(cnd % side) - kmax;
    bool ok = cnd < ncand && l * l + m * m + n * n <= kmax * kmax;
    if (l == 0 && m < 0) ok = false;
    if (l == 0 && m == 0 && n < 1) ok = false;
    const unsigned long long b = __ballot(ok);
    const int before = __popcll(b & (lane ? (~0ull >> (64 - lane)) : 0ull));
    if (lane == 0) wave_tot[wv] = __popcll(b);
    __syncthreads();
    int off = base;
    for (int i = 0; i < wv; i++) off += wave_tot[i];
    if (ok) w[off + before] = Wave{{double(l), double(m), double(n)}, 0.0, 0.0};
    __syncthreads();
    if (threadIdx.x == 0) base += wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
    __syncthreads();
  }
}

void ewald_sum_noalloc(Ctx &c, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns, int ksize,
                       double alpha, double sigma, double cutoff, double cyc) {
  const int nw = ewald_wave_count(ksize);
  if (nw > kEwaldMaxWaves)
    fail(EXAFMM_ERR_CAPACITY, "ksize " + std::to_string(ksize) + " needs " + std::to_string(nw) +
                                  " waves; the allocation-free Ewald sum holds at most " +
                                  std::to_string(kEwaldMaxWaves));
  Wave *w = ewald_scratch(c.cfg.device);
  if (nw) {
    hipLaunchKernelGGL(k_ewald_waves, dim3(1), dim3(256), 0, c.s0, ksize, w);
    XF_LAUNCH();
  }
  ewald_launch(c, w, nw, t, nt, s, ns, alpha, sigma, cutoff, cyc);
}

}  // namespace xf
