// expansions.hip -- P2M, M2M (upward), L2L, L2P (downward).
//
// Reference operators: P2M 3d/kernel.h:187-202, M2M :204-231, L2L :262-291,
// L2P :293-322; passes 3d/traverse_lazy.h:7-24 (upward) and :65-80
// (downward).
//
// Storage: every cell's expansions are kept in radius-scaled real form,
//   M~[rf(n,m,.)] = (Re, Im) M_n^m / R^n,   L~[rf(j,k,.)] = (Re, Im) L_j^k R^(j+1),
// P^2 doubles per cell (the imaginary parts of the m = 0 coefficients are
// zero for real charges and are not stored).  In this form the child <-> parent
// shifts do not depend on the level: M2M and L2L are one fixed P^2 x P^2 FP64
// matrix per octant, generated once from the reference's shift formulas
// (m2m_coef / l2l_coef applied to unit vectors) and applied as batched
// matrix-vector products.  Harmonics come from the trig-free Cartesian
// recurrences of fmm_internal.h (same functions and normalisation as the
// reference's spherical evaluation).  P2M, M2M and L2L run in FP64; L2P
// evaluates in FP32 (FP64 in parity mode) and adds into FP64 body
// accumulators, with the force from Cartesian ladder identities of the
// solid harmonics (finite at r = 0 and on the z axis, where the reference's
// spherical form divides by zero, 3d/kernel.h:31-40).
#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

// 1/i for i <= 1024 (the harmonic recurrences need 1/(2m) and
// 1/((n+m)(n-m)), n, m < 16): correctly rounded at compile time, so a table
// read replaces an FP64 division in the kernels.
struct InvTable {
  double v[1025];
  constexpr InvTable() : v() {
    for (int i = 1; i <= 1024; i++) v[i] = 1.0 / i;
  }
};
static __constant__ InvTable kInv = InvTable();

// ---------------------------------------------------------------- P2M
// One wave per leaf: the leaf's bodies are staged in LDS 64 at a time; lane
// (g, m) = (lane / 16, lane % 16) sums column n = m..P-1 of conj(Y_n^m(u)),
// u = (x_b - X_c)/R, weighted by q over the bodies b = g mod 4
// (3d/kernel.h:193 evaluates Y at -beta, i.e. the conjugate); the four body
// groups are then added with cross-lane shuffles in a fixed order.
constexpr int kP2MWaves = 4;

template <int PMAX>
__device__ __attribute__((always_inline)) inline void p2m_body(int bx, int P, const int *__restrict__ leaves,
                                const int *__restrict__ nleaves,
                                const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                const double4 *__restrict__ cx,
                                const double4 *__restrict__ xs, double *__restrict__ M) {
  static_assert(PMAX <= 16, "16 orders per body group");
  __shared__ double4 sb[kP2MWaves][kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int li = bx * kP2MWaves + w;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const int beg = cbeg[c], cnt = ccnt[c];
  const double4 C = cx[c];
  const double invR = C.w > 0.0 ? 1.0 / C.w : 0.0;  // R = 0: every body at the centre
  const int m = lane & 15, g = lane >> 4;
  // per-lane constants (no divisions per body): 1/(2^m m!) and the
  // recurrence factors -1/((n+m)(n-m)), n = m+2..P-1
  double cm = 1.0, sn[PMAX];
  for (int k = 1; k <= m; k++) cm *= kInv.v[2 * k];
#pragma unroll
  for (int n = 0; n < PMAX; n++) sn[n] = (n >= m + 2 && n < P) ? -kInv.v[(n + m) * (n - m)] : 0.0;
  double ar[PMAX], ai[PMAX];
#pragma unroll
  for (int n = 0; n < PMAX; n++) ar[n] = ai[n] = 0;
  for (int b0 = 0; b0 < cnt; b0 += kWave) {
    const int nb = min(kWave, cnt - b0);
    if (lane < nb) sb[w][lane] = xs[beg + b0 + lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (m < P) {
      // two bodies (bi, bi + 4) per pass: independent recurrence chains
      // interleave; a missing second body has charge 0 and adds exact zeros
      for (int bi = g; bi < nb; bi += 8) {
        const bool two = bi + 4 < nb;
        const double4 B0 = sb[w][bi], B1 = sb[w][two ? bi + 4 : bi];
        const double x0 = (B0.x - C.x) * invR, y0 = -(B0.y - C.y) * invR, z0 = (B0.z - C.z) * invR;  // conjugate
        const double x1 = (B1.x - C.x) * invR, y1 = -(B1.y - C.y) * invR, z1 = (B1.z - C.z) * invR;
        const double r20 = x0 * x0 + y0 * y0 + z0 * z0, r21 = x1 * x1 + y1 * y1 + z1 * z1;
        double yr0 = 1.0, yi0 = 0, yr1 = 1.0, yi1 = 0;  // (x + iy)^m
        for (int k = 1; k <= m; k++) {
          const double tr0 = yr0 * x0 - yi0 * y0, ti0 = yr0 * y0 + yi0 * x0;
          const double tr1 = yr1 * x1 - yi1 * y1, ti1 = yr1 * y1 + yi1 * x1;
          yr0 = tr0;
          yi0 = ti0;
          yr1 = tr1;
          yi1 = ti1;
        }
        const double qc0 = B0.w * cm, qc1 = two ? B1.w * cm : 0.0;  // q * Y_m^m = q (x+iy)^m / (2^m m!)
        double pr0 = 0, pi0 = 0, cr0 = yr0 * qc0, ci0 = yi0 * qc0;
        double pr1 = 0, pi1 = 0, cr1 = yr1 * qc1, ci1 = yi1 * qc1;
#pragma unroll
        for (int n = 0; n < PMAX; n++) {
          if (n < m || n >= P) continue;
          if (n > m) {
            double nr0, ni0, nr1, ni1;
            if (n == m + 1) {
              nr0 = -z0 * cr0;
              ni0 = -z0 * ci0;
              nr1 = -z1 * cr1;
              ni1 = -z1 * ci1;
            } else {
              const double c2 = double(2 * n - 1);
              nr0 = (c2 * z0 * cr0 + r20 * pr0) * sn[n];
              ni0 = (c2 * z0 * ci0 + r20 * pi0) * sn[n];
              nr1 = (c2 * z1 * cr1 + r21 * pr1) * sn[n];
              ni1 = (c2 * z1 * ci1 + r21 * pi1) * sn[n];
            }
            pr0 = cr0;
            pi0 = ci0;
            cr0 = nr0;
            ci0 = ni0;
            pr1 = cr1;
            pi1 = ci1;
            cr1 = nr1;
            ci1 = ni1;
          }
          ar[n] += cr0;
          ai[n] += ci0;
          ar[n] += cr1;
          ai[n] += ci1;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int n = 0; n < PMAX; n++) {
    ar[n] += __shfl_xor(ar[n], 16);
    ai[n] += __shfl_xor(ai[n], 16);
    ar[n] += __shfl_xor(ar[n], 32);
    ai[n] += __shfl_xor(ai[n], 32);
  }
  if (g != 0 || m >= P) return;
  double *Mc = M + int64_t(c) * P * P;
#pragma unroll
  for (int n = 0; n < PMAX; n++) {
    if (n < m || n >= P) continue;
    Mc[rf(n, m, 0)] = ar[n];
    if (m > 0) Mc[rf(n, m, 1)] = ai[n];
  }
}

template <int PMAX>
__global__ void __launch_bounds__(kP2MWaves *kWave) k_p2m(int P, const int *__restrict__ leaves,
                                                          const int *__restrict__ nleaves,
                                                          const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                                          const double4 *__restrict__ cx,
                                                          const double4 *__restrict__ xs, double *__restrict__ M) {
  p2m_body<PMAX>(blockIdx.x, P, leaves, nleaves, cbeg, ccnt, cx, xs, M);
}

// All problems of a batch in one launch (block -> problem through start[]).
__global__ void __launch_bounds__(kP2MWaves *kWave) k_p2m_grouped(const GroupTree *__restrict__ g,
                                                                  const int *__restrict__ start, int nprob) {
  const int b = blockIdx.x, i = group_of(start, nprob, b);
  const GroupTree &t = g[i];
  p2m_body<kMaxP>(b - start[i], t.P, t.leaves, t.nleaves, t.cbeg, t.ccnt, t.cx, t.xs, t.M);
}


// ---------------------------------------------------------------- shift operators
// Column `col` of the scaled M2M (kind 0) or L2L (kind 1) operator of octant
// o, for a parent of radius 1 and a child of radius 1/2 centred at
// parent + (+/-1/2) per axis.  T[kind][o][col][row] (row fastest).
__global__ void k_shift_ops(int P, double *__restrict__ T) {
  const int PP = P * P, NT = P * (P + 1) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * 8 * PP) return;
  const int kind = t / (8 * PP), o = (t / PP) % 8, col = t % PP;
  cpx<double> Y[kMaxP * (kMaxP + 1) / 2];
  double2 in[kMaxP * (kMaxP + 1) / 2];
  for (int i = 0; i < NT; i++) in[i] = make_double2(0, 0);
  int n = 0;
  while ((n + 1) * (n + 1) <= col) n++;
  const int rem = col - n * n, m = (rem + 1) / 2;
  const bool im = rem > 0 && rem % 2 == 0;
  const double ox = (o & 1) ? 0.5 : -0.5, oy = (o & 2) ? 0.5 : -0.5, oz = (o & 4) ? 0.5 : -0.5;
  double *out = T + ((int64_t(kind) * 8 + o) * PP + col) * PP;
  if (kind == 0) {
    // child scaled unit vector -> unscaled child M (radius 1/2), shift by Xp - Xc
    const double v = ldexp(1.0, -n);
    in[nms(n, m)] = im ? make_double2(0, v) : make_double2(v, 0);
    regular_all<double>(P, -ox, -oy, -oz, Y);
    for (int jk = 0; jk < NT; jk++) {
      int j, k;
      jk_of(jk, j, k);
      const double2 r = m2m_coef(j, k, in, Y);  // parent radius 1: scaled == unscaled
      out[rf(j, k, 0)] = r.x;
      if (k > 0) out[rf(j, k, 1)] = r.y;
    }
  } else {
    // parent scaled unit vector (radius 1: L = L~) -> child L, scaled by (1/2)^(j+1)
    in[nms(n, m)] = im ? make_double2(0, 1) : make_double2(1, 0);
    regular_all<double>(P, ox, oy, oz, Y);
    for (int jk = 0; jk < NT; jk++) {
      int j, k;
      jk_of(jk, j, k);
      const double2 r = l2l_coef(P, j, k, in, Y);
      const double s = ldexp(1.0, -(j + 1));
      out[rf(j, k, 0)] = r.x * s;
      if (k > 0) out[rf(j, k, 1)] = r.y * s;
    }
  }
}

void shift_ops(Ctx &c, hipStream_t s) {
  if (c.shift_P == c.P) return;
  const int PP = c.P * c.P;
  c.shift_T.ensure(size_t(2) * 8 * PP * PP);
  hipLaunchKernelGGL(k_shift_ops, dim3(ceil_div(2 * 8 * PP, 64)), dim3(64), 0, s, c.P, c.shift_T.p);
  XF_LAUNCH();
  c.shift_P = c.P;
}

// ---------------------------------------------------------------- M2M
// M2M and L2L as FP64 MFMA GEMMs over the cells of one level
// (v_mfma_f64_16x16x4_f64: A lane l = row l%16, k l/16; B lane l = column
// l%16, k l/16; D register i of lane l = row l/16 + 4i, column l%16).  A wave owns one
// 16-row tile of the operator and 16 parents; its operands stream straight
// from L2 into registers (no LDS round trip: every loaded value feeds a
// 16x16x4 product), four k-steps per batch so the loads of a batch overlap.
// Degree structure: an M2M row of degree j reads columns of degree <= j, an
// L2L row of degree j reads columns of degree >= j, so a row tile stops
// (starts) at the column bound of its last (first) row.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kShiftWaves = 4;  // row tiles per workgroup
constexpr int kShiftBatch = 4;  // k-steps of 4 columns loaded together
// L2L runs alone on the device (after the GEMM), so it loads deeper: all
// 25 k-steps of order 10 in two rounds (M2M runs beside P2P, where more
// registers cost it more than the latency it saves)
constexpr int kL2LBatch = 13;

__device__ inline int degree_of(int row) {
  int j = 0;
  while ((j + 1) * (j + 1) <= row) j++;
  return j;
}

// Children of the parents p0 .. p0 + 15 of a level by octant: thread t < 128
// takes child t % 8 of parent p0 + t / 8 and files it under the octant its
// centre lies in relative to the parent's (strict >, 3d/build_tree.h:50-56);
// octants without a child keep -1.  Every load is issued at once (no walk
// over the children per octant).  `only` >= 0 keeps that octant alone, in
// row 0.  The caller fills chl with -1 and puts a barrier before and after.
__device__ inline void file_children(int tid, int p0, int lstart, int lcount, const int *__restrict__ cchild,
                                     const int *__restrict__ cnch, const double4 *__restrict__ cx, int only,
                                     int (*chl)[16]) {
  if (tid >= 128) return;
  const int q = tid >> 3, j = tid & 7;
  if (p0 + q >= lcount) return;
  const int p = lstart + p0 + q;
  const int nch = cnch[p], first = cchild[p];
  if (j >= nch) return;
  const double4 Cp = cx[p], Cc = cx[first + j];
  const int o = (Cc.x > Cp.x) + ((Cc.y > Cp.y) << 1) + ((Cc.z > Cp.z) << 2);
  if (only < 0)
    chl[o][q] = first + j;
  else if (o == only)
    chl[0][q] = first + j;
}

// M~_parent = sum over children of T_M2M[octant] M~_child (3d/kernel.h:206).
// grid: (parent tiles of 16, row tiles of 16).  The eight waves of a block
// take one octant each (K = that octant's columns), so a small level's
// launch waits for one octant's chain of loads, not eight; the eight
// partial tiles are then added in LDS in octant order.
constexpr int kM2MWaves = 8;
__device__ __attribute__((always_inline)) inline void m2m_body(int bx, int by, int P, int lstart, int lcount,
                                const int *__restrict__ cchild,
                                const int *__restrict__ cnch,
                                const double4 *__restrict__ cx,
                                const double *__restrict__ T, double *__restrict__ M) {
  __shared__ int chl[8][16];
  __shared__ v4d part[kM2MWaves][kWave];
  const int PP = P * P;
  const int lane = threadIdx.x % kWave, o = threadIdx.x / kWave;
  const int j16 = lane & 15, kq = lane >> 4;
  const int p0 = bx * 16;
  if (threadIdx.x < 128) chl[threadIdx.x / 16][threadIdx.x % 16] = -1;
  __syncthreads();
  file_children(threadIdx.x, p0, lstart, lcount, cchild, cnch, cx, -1, chl);
  __syncthreads();
  const int r0 = by * 16;  // the grid has ceil(PP / 16) row tiles: r0 < PP
  const int jmax = degree_of(min(r0 + 16, PP) - 1);
  const int ncol = (jmax + 1) * (jmax + 1);
  const int ksteps = (ncol + 3) / 4;
  const int row = min(r0 + j16, PP - 1);  // rows past PP compute garbage that is never stored
  const int ch = chl[o][j16];
  const double *To = T + int64_t(o) * PP * PP;
  const double *Mc = M + int64_t(ch < 0 ? 0 : ch) * PP;
  v4d acc = {0, 0, 0, 0};
  for (int k0 = 0; k0 < ksteps; k0 += kShiftBatch) {
    double a[kShiftBatch], b[kShiftBatch];
#pragma unroll
    for (int s = 0; s < kShiftBatch; s++) {
      const int col = (k0 + s) * 4 + kq;
      const bool in = col < ncol;
      a[s] = in ? To[int64_t(col) * PP + row] : 0.0;
      b[s] = (in && ch >= 0) ? Mc[col] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kShiftBatch; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
  }
  part[o][lane] = acc;
  __syncthreads();
  if (o != 0) return;
  v4d sum = part[0][lane];
#pragma unroll
  for (int oo = 1; oo < kM2MWaves; oo++) sum += part[oo][lane];
  const int pl = p0 + j16;
  if (pl >= lcount) return;
  const int p = lstart + pl;
  if (cnch[p] == 0) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = r0 + kq + 4 * i;
    if (r < PP) M[int64_t(p) * PP + r] = sum[i];
  }
}

__global__ void __launch_bounds__(kM2MWaves *kWave) k_m2m(int P, int lstart, int lcount,
                                                          const int *__restrict__ cchild,
                                                          const int *__restrict__ cnch,
                                                          const double4 *__restrict__ cx,
                                                          const double *__restrict__ T, double *__restrict__ M) {
  m2m_body(blockIdx.x, blockIdx.y, P, lstart, lcount, cchild, cnch, cx, T, M);
}

// Row-tile form of M2M and L2L (single tree): one workgroup per 16 parents
// of a level, wave w = row tile w of the P^2 rows (ceil(P^2/16) waves), each
// wave looping over the 8 octants.  All 8 x R waves that read a parent
// tile's operands (the children's multipoles, the parent's locals) run in
// one workgroup, so those rows come from HBM once and then from that CU's
// caches (the one-octant-per-workgroup layout fetched them once per XCD
// and row block).  Operator columns and the B operand are loaded
// kShiftBatch2 k-steps at a time.
constexpr int kShiftBatch2 = 8;
constexpr int kRowTileMaxWaves = 16;  // P <= 16
// Levels with fewer parents run the octant-parallel kernels (k_m2m: an
// octant per wave, k_l2l: an octant per workgroup): their chains of
// dependent loads are 8x shorter, and few parents leave the cache-sharing
// of the row-tile form little to share.
constexpr int kRowTileMinParents = 8192;

// waves of a row-tile workgroup: one per row tile, at least two (filing the
// children of 16 parents takes 128 threads)
static int row_tile_waves(int PP) { return std::max(2, ceil_div(PP, 16)); }

__device__ inline void file_children_all(int p0, int lstart, int lcount, const int *__restrict__ cchild,
                                         const int *__restrict__ cnch, const double4 *__restrict__ cx,
                                         int (*chl)[16]) {
  if (threadIdx.x < 128) chl[threadIdx.x / 16][threadIdx.x % 16] = -1;
  __syncthreads();
  file_children(threadIdx.x, p0, lstart, lcount, cchild, cnch, cx, -1, chl);
  __syncthreads();
}

// M~_parent = sum over octants of T_M2M[octant] M~_child (3d/kernel.h:206),
// octants added in order into one accumulator per row tile.
__global__ void __launch_bounds__(kRowTileMaxWaves *kWave) k_m2m_rows(int P, int lstart, int lcount,
                                                                      const int *__restrict__ cchild,
                                                                      const int *__restrict__ cnch,
                                                                      const double4 *__restrict__ cx,
                                                                      const double *__restrict__ T,
                                                                      double *__restrict__ M) {
  __shared__ int chl[8][16];
  const int PP = P * P;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int j16 = lane & 15, kq = lane >> 4;
  const int p0 = blockIdx.x * 16;
  file_children_all(p0, lstart, lcount, cchild, cnch, cx, chl);
  const int r0 = w * 16;  // blockDim.x = 64 max(2, ceil(PP / 16))
  if (r0 >= PP) return;
  const int jmax = degree_of(min(r0 + 16, PP) - 1);
  const int ncol = (jmax + 1) * (jmax + 1);
  const int ksteps = (ncol + 3) / 4;
  const int row = min(r0 + j16, PP - 1);
  v4d acc = {0, 0, 0, 0};
  for (int o = 0; o < 8; o++) {
    const int ch = chl[o][j16];
    if (!__any(ch >= 0)) continue;
    const double *To = T + int64_t(o) * PP * PP;
    const double *Mc = M + int64_t(ch < 0 ? 0 : ch) * PP;
    for (int k0 = 0; k0 < ksteps; k0 += kShiftBatch2) {
      double a[kShiftBatch2], b[kShiftBatch2];
#pragma unroll
      for (int s = 0; s < kShiftBatch2; s++) {
        const int col = (k0 + s) * 4 + kq;
        const bool in = col < ncol;
        a[s] = in ? To[int64_t(col) * PP + row] : 0.0;
        b[s] = (in && ch >= 0) ? Mc[col] : 0.0;
      }
#pragma unroll
      for (int s = 0; s < kShiftBatch2; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
    }
  }
  const int pl = p0 + j16;
  if (pl >= lcount) return;
  const int p = lstart + pl;
  if (cnch[p] == 0) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = r0 + kq + 4 * i;
    if (r < PP) M[int64_t(p) * PP + r] = acc[i];
  }
}

// L~_child += T_L2L[octant] L~_parent (3d/kernel.h:265-280) for every child
// of the 16 parents, octant by octant.
__global__ void __launch_bounds__(kRowTileMaxWaves *kWave) k_l2l_rows(int P, int lstart, int lcount,
                                                                      const int *__restrict__ cchild,
                                                                      const int *__restrict__ cnch,
                                                                      const double4 *__restrict__ cx,
                                                                      const double *__restrict__ T,
                                                                      double *__restrict__ L) {
  __shared__ int chl[8][16];
  const int PP = P * P;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int j16 = lane & 15, kq = lane >> 4;
  const int p0 = blockIdx.x * 16;
  file_children_all(p0, lstart, lcount, cchild, cnch, cx, chl);
  const int r0 = w * 16;
  if (r0 >= PP) return;
  const int jmin = degree_of(r0);
  const int c0 = jmin * jmin;
  const int ksteps = (PP - c0 + 3) / 4;
  const int row = min(r0 + j16, PP - 1);
  const int pl = min(p0 + j16, lcount - 1);
  const double *Lp = L + int64_t(lstart + pl) * PP;
  for (int o = 0; o < 8; o++) {
    const int ch = chl[o][j16];
    if (!__any(ch >= 0)) continue;
    const double *To = T + (int64_t(8) + o) * PP * PP;
    v4d acc = {0, 0, 0, 0};
    for (int k0 = 0; k0 < ksteps; k0 += kShiftBatch2) {
      double a[kShiftBatch2], b[kShiftBatch2];
#pragma unroll
      for (int s = 0; s < kShiftBatch2; s++) {
        const int col = c0 + (k0 + s) * 4 + kq;
        const bool in = col < PP;
        a[s] = in ? To[int64_t(col) * PP + row] : 0.0;
        b[s] = (in && ch >= 0) ? Lp[col] : 0.0;
      }
#pragma unroll
      for (int s = 0; s < kShiftBatch2; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
    }
    if (ch >= 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = r0 + kq + 4 * i;
        if (r < PP) L[int64_t(ch) * PP + r] += acc[i];
      }
    }
  }
}

// One level of every problem of a batch in one launch: lv[i] describes
// problem lv[i].prob's cells of this level as nbx x nby blocks.
__global__ void __launch_bounds__(kM2MWaves *kWave) k_m2m_grouped(const GroupTree *__restrict__ g,
                                                                    const GroupLevel *__restrict__ lv,
                                                                    const int *__restrict__ start, int n) {
  const int b = blockIdx.x, i = group_of(start, n, b), lb = b - start[i];
  const GroupLevel L = lv[i];
  const GroupTree &t = g[L.prob];
  m2m_body(lb % L.nbx, lb / L.nbx, t.P, L.lstart, L.lcount, t.cchild, t.cnch, t.cx, t.shift_T, t.M);
}


// L~_child += T_L2L[octant] L~_parent for the children of one level of
// parents (3d/kernel.h:265-280).  grid: (parent tiles of 16, row blocks of
// 64, octant).
__device__ __attribute__((always_inline)) inline void l2l_body(int bx, int by, int bz, int P, int lstart, int lcount,
                                const int *__restrict__ cchild,
                                const int *__restrict__ cnch,
                                const double4 *__restrict__ cx,
                                const double *__restrict__ T, double *__restrict__ L) {
  __shared__ int chl[1][16];
  const int PP = P * P;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int j16 = lane & 15, kq = lane >> 4;
  const int p0 = bx * 16, o = bz;
  if (threadIdx.x < 16) chl[0][threadIdx.x] = -1;
  __syncthreads();
  file_children(threadIdx.x, p0, lstart, lcount, cchild, cnch, cx, o, chl);
  __syncthreads();
  const int r0 = (by * kShiftWaves + w) * 16;
  if (r0 >= PP) return;
  const int jmin = degree_of(r0);
  const int c0 = jmin * jmin;
  const int ksteps = (PP - c0 + 3) / 4;
  const int row = min(r0 + j16, PP - 1);
  const int ch = chl[0][j16];
  const double *To = T + (int64_t(8) + o) * PP * PP;
  const int pl = min(p0 + j16, lcount - 1);
  const double *Lp = L + int64_t(lstart + pl) * PP;
  v4d acc = {0, 0, 0, 0};
  for (int k0 = 0; k0 < ksteps; k0 += kL2LBatch) {
    double a[kL2LBatch], b[kL2LBatch];
#pragma unroll
    for (int s = 0; s < kL2LBatch; s++) {
      const int col = c0 + (k0 + s) * 4 + kq;
      const bool in = col < PP;
      a[s] = in ? To[int64_t(col) * PP + row] : 0.0;
      b[s] = (in && ch >= 0) ? Lp[col] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kL2LBatch; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
  }
  if (ch < 0) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = r0 + kq + 4 * i;
    if (r < PP) L[int64_t(ch) * PP + r] += acc[i];
  }
}

__global__ void __launch_bounds__(kShiftWaves *kWave) k_l2l(int P, int lstart, int lcount,
                                                            const int *__restrict__ cchild,
                                                            const int *__restrict__ cnch,
                                                            const double4 *__restrict__ cx,
                                                            const double *__restrict__ T, double *__restrict__ L) {
  l2l_body(blockIdx.x, blockIdx.y, blockIdx.z, P, lstart, lcount, cchild, cnch, cx, T, L);
}

// One level of parents of every problem of a batch in one launch (blocks
// nbx x nby x 8 octants per problem).
__global__ void __launch_bounds__(kShiftWaves *kWave) k_l2l_grouped(const GroupTree *__restrict__ g,
                                                                    const GroupLevel *__restrict__ lv,
                                                                    const int *__restrict__ start, int n) {
  const int b = blockIdx.x, i = group_of(start, n, b), lb = b - start[i];
  const GroupLevel L = lv[i];
  const GroupTree &t = g[L.prob];
  const int bx = lb % L.nbx, r = lb / L.nbx;
  l2l_body(bx, r % L.nby, r / L.nby, t.P, L.lstart, L.lcount, t.cchild, t.cnch, t.cx, t.shift_T, t.L);
}


// ---------------------------------------------------------------- L2P
// One thread per body.  With u = (x - Xc)/R and
// Lh_n^m = L_n^m R^n = L~/R, phi = Re sum_{n, m>=0} w_m Lh_n^m Y_n^m(u)
// (w_0 = 1, w_m = 2, 3d/kernel.h:302-316).  The gradient uses
//   dz Y_n^m = -Y_{n-1}^m,  (dx + i dy) Y_n^m = -Y_{n-1}^{m+1},
//   (dx - i dy) Y_n^m = Y_{n-1}^{m-1} (m >= 1), = -conj(Y_{n-1}^1) (m = 0),
// collected per harmonic Y_n^m so each column m is generated once.
// Potential and field of one body at u = (x, y, z) (leaf units) from the
// leaf's folded coefficients Lh (LDS) and the recurrence factors rec.
// PC > 0: the order is the compile-time PC and the (n, m) loops unroll, so
// the LDS offsets and the branches below resolve at compile time.
template <class T, int PC = 0>
__device__ __attribute__((always_inline)) inline void l2p_eval(int Prt, T x, T y, T z, const cpx<T> *Lh,
                                                              const T *rec, T &pot, T &Fx2, T &Fy2, T &Dz) {
  const int P = PC > 0 ? PC : Prt;
  const T r2 = x * x + y * y + z * z;
  T Ar = 0, Ai = 0, Br = 0, Bi = 0;
  pot = 0;
  Dz = 0;
  T ymr = 1, ymi = 0;          // Y_m^m
  T pr = 0, pi = 0, cr = 0, ci = 0;  // Y_{n-1}^m, Y_n^m
  auto term = [&](int m, int n) {
    if (n == m) {
      if (m > 0) {
        const T inv = rec[nms(m, m)];
        const T tr = (ymr * x - ymi * y) * inv, ti = (ymr * y + ymi * x) * inv;
        ymr = tr;
        ymi = ti;
      }
      pr = 0;
      pi = 0;
      cr = ymr;
      ci = ymi;
    } else {
      T nr, ni;
      if (n == m + 1) {
        nr = -z * cr;
        ni = -z * ci;
      } else {
        const T s = rec[nms(n, m)];
        const T cz = T(2 * n - 1) * z;
        nr = (cz * cr + r2 * pr) * s;
        ni = (cz * ci + r2 * pi) * s;
      }
      pr = cr;
      pi = ci;
      cr = nr;
      ci = ni;
    }
    const cpx<T> l0 = Lh[nms(n, m)];
    pot += l0.re * cr - l0.im * ci;
    if (n + 1 < P) {
      const cpx<T> lz = Lh[nms(n + 1, m)];
      Dz -= lz.re * cr - lz.im * ci;
      if (m >= 1) {
        const cpx<T> la = Lh[nms(n + 1, m - 1)];
        Ar -= la.re * cr - la.im * ci;
        Ai -= la.re * ci + la.im * cr;
      }
      const cpx<T> lb = Lh[nms(n + 1, m + 1)];
      Br += lb.re * cr - lb.im * ci;
      Bi += lb.re * ci + lb.im * cr;
      if (m == 1) {  // the m = 0 terms' (dx - i dy): -Lh_{n+1}^0 conj(Y_n^1)
        const cpx<T> l00 = Lh[nms(n + 1, 0)];
        Br -= l00.re * cr + l00.im * ci;
        Bi -= l00.im * cr - l00.re * ci;
      }
    }
  };
  if constexpr (PC > 0) {
#pragma unroll
    for (int m = 0; m < PC; m++)
#pragma unroll
      for (int n = m; n < PC; n++) term(m, n);
  } else {
    for (int m = 0; m < P; m++)
      for (int n = m; n < P; n++) term(m, n);
  }
  Fx2 = Ar + Br;  // 2 dphi/dx
  Fy2 = Ai - Bi;  // 2 dphi/dy
}

// Lh (w_m and L~ -> L R^n folded in) and the recurrence factors of one leaf,
// entries e = lane, lane + stride, ...
template <class T>
__device__ inline void l2p_stage(int P, int e0, int stride, const double *__restrict__ Lc, double invR, cpx<T> *Lh,
                                 T *rec) {
  const int NT = P * (P + 1) / 2;
  for (int e = e0; e < NT; e += stride) {
    int n, m;
    jk_of(e, n, m);
    const double w = (m == 0 ? 1.0 : 2.0) * invR;
    Lh[e] = {T(Lc[rf(n, m, 0)] * w), T(m ? Lc[rf(n, m, 1)] * w : 0.0)};
    if (rec) rec[e] = n == m ? T(m ? kInv.v[2 * m] : 0.0) : (n >= m + 2 ? T(-kInv.v[(n + m) * (n - m)]) : T(0));
  }
}

template <int MAXP, class T>
__device__ __attribute__((always_inline)) inline void l2p_body(int bx, int P, const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                const double4 *__restrict__ cx, const double *__restrict__ L,
                                const double4 *__restrict__ xs, exafmm_body *__restrict__ bodies) {
  __shared__ cpx<T> Lh[MAXP * (MAXP + 1) / 2];
  __shared__ T rec[MAXP * (MAXP + 1) / 2];  // 1/(2m) at (m, m), -1/((n+m)(n-m)) at (n, m), n >= m + 2
  const int li = bx;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const double4 C = cx[c];
  const double invR = C.w > 0.0 ? 1.0 / C.w : 0.0;  // R = 0: every body at the centre
  l2p_stage<T>(P, threadIdx.x, 64, L + int64_t(c) * P * P, invR, Lh, rec);
  __syncthreads();
  const int beg = cbeg[c], cnt = ccnt[c];
  for (int b = beg + threadIdx.x; b < beg + cnt; b += 64) {
    const double4 B = xs[b];
    T pot, fx2, fy2, dz;
    l2p_eval<T>(P, T((B.x - C.x) * invR), T((B.y - C.y) * invR), T((B.z - C.z) * invR), Lh, rec, pot, fx2, fy2, dz);
    exafmm_body *o = bodies + b;
    o->p += double(pot);
    o->F[0] += 0.5 * double(fx2) * invR;
    o->F[1] += 0.5 * double(fy2) * invR;
    o->F[2] += double(dz) * invR;
  }
}

// FP32 L2P over the bodies in tree order, 64 per wave (full waves whatever
// the leaf sizes; one wave per leaf leaves half its lanes idle at 32 bodies
// a leaf).  The wave's bodies fall into runs of equal leaf; up to kL2PSlots
// runs at a time have their coefficients staged in LDS (one pass for
// uniform trees, a few when many small leaves share a wave), and each lane
// evaluates its own body against its run's slot.
constexpr int kL2PSlots = 8;
template <int MAXP, int PC>
__global__ void __launch_bounds__(64) k_l2p_flat(int P, int64_t n, const int *__restrict__ body_leaf,
                                                 const double4 *__restrict__ cx, const double *__restrict__ L,
                                                 const double4 *__restrict__ xs, exafmm_body *__restrict__ bodies) {
  constexpr int NTM = MAXP * (MAXP + 1) / 2;
  __shared__ cpx<float> Lh[kL2PSlots][NTM];
  __shared__ float rec[NTM];
  __shared__ int sleaf[kL2PSlots];
  const int lane = threadIdx.x;
  const int64_t b = int64_t(blockIdx.x) * kWave + lane;
  const bool valid = b < n;
  const int leaf = valid ? body_leaf[b] : -1;
  const int prev = __shfl_up(leaf, 1);
  const unsigned long long starts = __ballot(valid && (lane == 0 || prev != leaf));
  const int nruns = __popcll(starts);
  const int run = valid ? __popcll(starts & (~0ull >> (63 - lane))) - 1 : -1;
  const int NT = P * (P + 1) / 2;
  for (int e = lane; e < NT; e += kWave) {
    int nn, m;
    jk_of(e, nn, m);
    rec[e] = nn == m ? (m ? float(kInv.v[2 * m]) : 0.0f) : (nn >= m + 2 ? float(-kInv.v[(nn + m) * (nn - m)]) : 0.0f);
  }
  double4 C = make_double4(0, 0, 0, 1), B = make_double4(0, 0, 0, 0);
  if (valid) {
    C = cx[leaf];
    B = xs[b];
  }
  const double invR = C.w > 0.0 ? 1.0 / C.w : 0.0;  // R = 0: every body at the centre
  for (int g0 = 0; g0 < nruns; g0 += kL2PSlots) {
    const int ns = min(kL2PSlots, nruns - g0);
    if (valid && run >= g0 && run < g0 + ns && (lane == 0 || prev != leaf)) sleaf[run - g0] = leaf;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int e = lane; e < ns * NT; e += kWave) {
      const int sl = e / NT, k = e - sl * NT, c = sleaf[sl];
      const double Rc = cx[c].w;
      l2p_stage<float>(P, k, NT, L + int64_t(c) * P * P, Rc > 0.0 ? 1.0 / Rc : 0.0, Lh[sl], nullptr);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (valid && run >= g0 && run < g0 + ns) {
      float pot, fx2, fy2, dz;
      l2p_eval<float, PC>(P, float((B.x - C.x) * invR), float((B.y - C.y) * invR), float((B.z - C.z) * invR),
                          Lh[run - g0], rec, pot, fx2, fy2, dz);
      exafmm_body *o = bodies + b;
      o->p += double(pot);
      o->F[0] += 0.5 * double(fx2) * invR;
      o->F[1] += 0.5 * double(fy2) * invR;
      o->F[2] += double(dz) * invR;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <int MAXP, class T>
__global__ void __launch_bounds__(64) k_l2p(int P, const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                            const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                            const double4 *__restrict__ cx, const double *__restrict__ L,
                                            const double4 *__restrict__ xs, exafmm_body *__restrict__ bodies) {
  l2p_body<MAXP, T>(blockIdx.x, P, leaves, nleaves, cbeg, ccnt, cx, L, xs, bodies);
}

__global__ void __launch_bounds__(64) k_l2p_grouped(const GroupTree *__restrict__ g, const int *__restrict__ start,
                                                    int nprob) {
  const int b = blockIdx.x, i = group_of(start, nprob, b);
  const GroupTree &t = g[i];
  l2p_body<kMaxP, float>(b - start[i], t.P, t.leaves, t.nleaves, t.cbeg, t.ccnt, t.cx, t.L, t.xs, t.bodies);
}


// ---------------------------------------------------------------- grouped launches
int p2m_group_blocks(int64_t nleaves) { return ceil_div(nleaves, kP2MWaves); }
int shift_group_rows(int P) { return ceil_div(P * P, 16 * kShiftWaves); }
int m2m_group_rows(int P) { return ceil_div(P * P, 16); }

void p2m_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_p2m_grouped, dim3(nblocks), dim3(kP2MWaves * kWave), 0, s, g, start, nprob);
  XF_LAUNCH();
}
void m2m_grouped(const GroupTree *g, const GroupLevel *lv, const int *start, int n, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_m2m_grouped, dim3(nblocks), dim3(kM2MWaves * kWave), 0, s, g, lv, start, n);
  XF_LAUNCH();
}
void l2l_grouped(const GroupTree *g, const GroupLevel *lv, const int *start, int n, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_l2l_grouped, dim3(nblocks), dim3(kShiftWaves * kWave), 0, s, g, lv, start, n);
  XF_LAUNCH();
}
void l2p_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_l2p_grouped, dim3(nblocks), dim3(64), 0, s, g, start, nprob);
  XF_LAUNCH();
}

void upward(Ctx &c, hipStream_t s) {
  Tree &t = c.t;
  const int P = c.P, PP = P * P;
  shift_ops(c, s);
  c.M.ensure(size_t(t.ncells) * PP);
  const int64_t nl = t.nleaves;
  // the per-lane degree loop is unrolled to PMAX (masked past P): the
  // smallest even bound >= P keeps the masked iterations few
  const dim3 pg(ceil_div(nl, kP2MWaves)), pb(kP2MWaves * kWave);
  switch (std::max(4, (P + 1) & ~1)) {
#define XF_P2M_CASE(n)                                                                                      \
  case n:                                                                                                    \
    hipLaunchKernelGGL((k_p2m<n>), pg, pb, 0, s, P, t.leaves.p, t.misc.p + 17, t.cbeg.p, t.ccnt.p, t.cx.p, \
                       t.xs.p, c.M.p);                                                                       \
    break;
    XF_P2M_CASE(4)
    XF_P2M_CASE(6)
    XF_P2M_CASE(8)
    XF_P2M_CASE(10)
    XF_P2M_CASE(12)
    XF_P2M_CASE(14)
    XF_P2M_CASE(16)
#undef XF_P2M_CASE
    default: fail(EXAFMM_ERR_UNSUPPORTED, "expansion order");
  }
  XF_LAUNCH();
  for (int l = t.depth - 1; l >= t.lroot; l--) {  // (a forest stops at its problems' roots)
    const int cnt = t.lstart[l + 1] - t.lstart[l];
    if (cnt >= kRowTileMinParents)
      hipLaunchKernelGGL(k_m2m_rows, dim3(ceil_div(cnt, 16)), dim3(row_tile_waves(PP) * kWave), 0, s, c.P,
                         t.lstart[l], cnt, t.cchild.p, t.cnch.p, t.cx.p, c.shift_T.p, c.M.p);
    else
      hipLaunchKernelGGL(k_m2m, dim3(ceil_div(cnt, 16), m2m_group_rows(P)), dim3(kM2MWaves * kWave), 0, s, c.P,
                         t.lstart[l], cnt, t.cchild.p, t.cnch.p, t.cx.p, c.shift_T.p, c.M.p);
    XF_LAUNCH();
  }
  c.have_M = true;
}

void l2l_pass(Ctx &c) {
  Tree &t = c.t;
  hipStream_t s = c.s0;
  const int PP = c.P * c.P;
  shift_ops(c, s);
  for (int l = t.lroot; l < t.depth; l++) {  // parents at level l
    const int cnt = t.lstart[l + 1] - t.lstart[l];
    if (cnt >= kRowTileMinParents)
      hipLaunchKernelGGL(k_l2l_rows, dim3(ceil_div(cnt, 16)), dim3(row_tile_waves(PP) * kWave), 0, s, c.P,
                         t.lstart[l], cnt, t.cchild.p, t.cnch.p, t.cx.p, c.shift_T.p, c.L.p);
    else
      hipLaunchKernelGGL(k_l2l, dim3(ceil_div(cnt, 16), ceil_div(PP, 16 * kShiftWaves), 8),
                         dim3(kShiftWaves * kWave), 0, s, c.P, t.lstart[l], cnt, t.cchild.p, t.cnch.p, t.cx.p,
                         c.shift_T.p, c.L.p);
    XF_LAUNCH();
  }
}

void l2p_pass(Ctx &c) {
  Tree &t = c.t;
  hipStream_t s = c.s0;
  const int P = c.P;
  // FP64 evaluation in parity mode, FP32 otherwise
  if (c.cfg.m2l_precision == EXAFMM_PREC_FP64)
    hipLaunchKernelGGL((k_l2p<kMaxP, double>), dim3(t.nleaves), dim3(64), 0, s, P, t.leaves.p, t.misc.p + 17,
                       t.cbeg.p, t.ccnt.p, t.cx.p, c.L.p, t.xs.p, c.bodies);
  else {
    // the common orders unrolled at compile time, any other order generic
    auto flat = k_l2p_flat<kMaxP, 0>;
    if (P == 6) flat = k_l2p_flat<kMaxP, 6>;
    if (P == 8) flat = k_l2p_flat<kMaxP, 8>;
    if (P == 10) flat = k_l2p_flat<kMaxP, 10>;
    hipLaunchKernelGGL(flat, dim3(ceil_div(t.n, kWave)), dim3(kWave), 0, s, P, t.n, t.body_leaf.p, t.cx.p, c.L.p,
                       t.xs.p, c.bodies);
  }
  XF_LAUNCH();
}

}  // namespace xf
