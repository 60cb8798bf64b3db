// p2p.hip -- near-field body-body interactions and the FP64 direct sum.
//
// Reference: P2P 3d/kernel.h:122-150 over each target's P2P list
// (3d/traverse_lazy.h:56), direct 3d/traverse_eager.h:193-202.
// Both add p += q/r and F -= q*dX/r^3 (dX = x_target - x_source) and skip
// pairs at zero distance.
//
// FP32 path: one wavefront per target leaf.  The leaf's targets sit in lanes;
// when a leaf holds n_i <= 32 bodies the wave splits into S = 64/n_i lane
// groups that take every S-th source, so small leaves still fill the wave.
// Sources of consecutive list entries are staged 256 at a time in LDS,
// already shifted into the target leaf's frame (coordinates are stored
// relative to their own leaf centre, so every difference is formed from
// small FP32 numbers).  1/r is the hardware reciprocal square root refined
// by one Newton step.  Each staged round is summed in FP32 and folded into
// FP64 accumulators (compensated accumulation: the FP64 carry keeps the
// low-order bits a long FP32 chain would drop), and lane groups are
// reduced through LDS at the end.
#include "fmm_internal.h"

namespace xf {

constexpr int kP2PWaves = 4;

constexpr int kP2PStage = 256;  // sources staged per round, per wave

template <bool NEWTON>
__device__ inline void interact(const float4 xi, const float4 sj, float &pot, float &ax, float &ay, float &az) {
  const float dx = xi.x - sj.x, dy = xi.y - sj.y, dz = xi.z - sj.z;
  const float r2 = dx * dx + dy * dy + dz * dz;
  float inv = __builtin_amdgcn_rsqf(r2);
  if (NEWTON) inv = inv * (1.5f - 0.5f * r2 * inv * inv);
  inv = r2 > 0.0f ? inv : 0.0f;
  const float qi = sj.w * inv;
  const float q3 = qi * inv * inv;
  pot += qi;
  ax += dx * q3;
  ay += dy * q3;
  az += dz * q3;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <bool NEWTON>
__global__ void __launch_bounds__(kP2PWaves *kWave) k_p2p32(
    const int *__restrict__ leaves, const int *__restrict__ nleaves, const int *__restrict__ cbeg,
    const int *__restrict__ ccnt, const double4 *__restrict__ cx, const float4 *__restrict__ xq,
    const int *__restrict__ lbeg, const int *__restrict__ lcnt, const int *__restrict__ list,
    const int *__restrict__ scbeg, const int *__restrict__ sccnt, const double4 *__restrict__ scx,
    const float4 *__restrict__ sxq, double cycle, exafmm_body *__restrict__ bodies) {
  __shared__ float4 src[kP2PWaves][kP2PStage];
  __shared__ double red[kP2PWaves][4][kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int li = blockIdx.x * kP2PWaves + w;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const int tb = cbeg[c], tn = ccnt[c];
  const double4 Cc = cx[c];
  const int b0 = lbeg[c], nl = lcnt[c];
  float4 *stage = src[w];
  for (int t0 = 0; t0 < tn; t0 += kWave) {
    const int ni = min(kWave, tn - t0);
    const int S = kWave / ni;
    const int i = lane % ni, g = lane / ni;
    const bool active = g < S;
    const float4 xi = xq[tb + t0 + i];
    double P64 = 0, X64 = 0, Y64 = 0, Z64 = 0;
    int s = 0, j_in = 0;  // next source leaf of the list, offset inside it
    while (s < nl) {
      // stage up to kP2PStage sources from consecutive list entries, shifted
      // into the target leaf's frame (periodic image included)
      int nfill = 0;
      while (nfill < kP2PStage && s < nl) {
        const int e = list[b0 + s];
        const int sc = e >> 5, img = e & 31;
        const int sb = scbeg[sc], sn = sccnt[sc];
        const double4 Cs = scx[sc];
        const float ox = float(Cs.x + ((img % 3) - 1) * cycle - Cc.x);
        const float oy = float(Cs.y + (((img / 3) % 3) - 1) * cycle - Cc.y);
        const float oz = float(Cs.z + ((img / 9) - 1) * cycle - Cc.z);
        const int take = min(sn - j_in, kP2PStage - nfill);
        for (int k = lane; k < take; k += kWave) {
          const float4 v = sxq[sb + j_in + k];
          stage[nfill + k] = make_float4(v.x + ox, v.y + oy, v.z + oz, v.w);
        }
        nfill += take;
        j_in += take;
        if (j_in == sn) {
          s++;
          j_in = 0;
        }
      }
      wave_sync();
      float pot = 0, ax = 0, ay = 0, az = 0;
      if (active) {
        int j = g;
        for (; j + 3 * S < nfill; j += 4 * S) {
          const float4 s0 = stage[j], s1 = stage[j + S], s2 = stage[j + 2 * S], s3 = stage[j + 3 * S];
          interact<NEWTON>(xi, s0, pot, ax, ay, az);
          interact<NEWTON>(xi, s1, pot, ax, ay, az);
          interact<NEWTON>(xi, s2, pot, ax, ay, az);
          interact<NEWTON>(xi, s3, pot, ax, ay, az);
        }
        for (; j < nfill; j += S) interact<NEWTON>(xi, stage[j], pot, ax, ay, az);
      }
      // fold the round's FP32 partials into the FP64 accumulators
      P64 += pot;
      X64 += ax;
      Y64 += ay;
      Z64 += az;
      wave_sync();
    }
    red[w][0][lane] = P64;
    red[w][1][lane] = X64;
    red[w][2][lane] = Y64;
    red[w][3][lane] = Z64;
    wave_sync();
    if (lane < ni) {
      double p = 0, fx = 0, fy = 0, fz = 0;
      for (int gg = 0; gg < S; gg++) {
        p += red[w][0][gg * ni + lane];
        fx += red[w][1][gg * ni + lane];
        fy += red[w][2][gg * ni + lane];
        fz += red[w][3][gg * ni + lane];
      }
      exafmm_body *o = bodies + tb + t0 + lane;
      o->p += p;
      o->F[0] -= fx;
      o->F[1] -= fy;
      o->F[2] -= fz;
    }
    wave_sync();
  }
}

// FP64 path: the reference's arithmetic per interaction (1/R2, q*sqrt, dX
// scaled by invR2*invR) on absolute FP64 coordinates; one thread per target
// body, sources in list order.
__global__ void __launch_bounds__(64) k_p2p64(const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                              const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                              const double4 *__restrict__ xs, const int *__restrict__ lbeg,
                                              const int *__restrict__ lcnt, const int *__restrict__ list,
                                              const int *__restrict__ scbeg, const int *__restrict__ sccnt,
                                              const double4 *__restrict__ sxs, double cycle,
                                              exafmm_body *__restrict__ bodies) {
#pragma clang fp contract(off)
  const int li = blockIdx.x;
  if (li >= *nleaves) return;
  const int c = leaves[li];
  const int tb = cbeg[c], tn = ccnt[c];
  const int b0 = lbeg[c], nl = lcnt[c];
  for (int i = threadIdx.x; i < tn; i += 64) {
    const double4 xi = xs[tb + i];
    double pot = 0, ax = 0, ay = 0, az = 0;
    for (int s = 0; s < nl; s++) {
      const int e = list[b0 + s];
      const int sc = e >> 5, img = e & 31;
      const double sx = ((img % 3) - 1) * cycle, sy = (((img / 3) % 3) - 1) * cycle, sz = ((img / 9) - 1) * cycle;
      const int sb = scbeg[sc], sn = sccnt[sc];
      for (int j = sb; j < sb + sn; j++) {
        const double4 xj = sxs[j];
        // 3dp/kernel.h:133: dX = Xi - Xj - iX * cycle
        double dx = xi.x - xj.x - sx, dy = xi.y - xj.y - sy, dz = xi.z - xj.z - sz;
        const double R2 = dx * dx + dy * dy + dz * dz;
        if (R2 != 0) {
          const double invR2 = 1.0 / R2;
          const double invR = xj.w * sqrt(invR2);
          const double f = invR2 * invR;
          pot += invR;
          ax += dx * f;
          ay += dy * f;
          az += dz * f;
        }
      }
    }
    exafmm_body *o = bodies + tb + i;
    o->p += pot;
    o->F[0] -= ax;
    o->F[1] -= ay;
    o->F[2] -= az;
  }
}

void p2p_apply(Ctx &tc, const Ctx &sc, hipStream_t s) {
  const Tree &T = tc.t, &S = sc.t;
  const Lists &L = tc.lists;
  const double cycle = tc.cfg.images > 0 ? tc.cfg.cycle : 0.0;
  if (tc.cfg.p2p_precision == EXAFMM_PREC_FP64) {
    hipLaunchKernelGGL(k_p2p64, dim3(T.nleaves), dim3(64), 0, s, T.leaves.p, T.misc.p + 17, T.cbeg.p, T.ccnt.p,
                       T.xs.p, L.p2p_beg.p, L.p2p_cnt.p, L.p2p.p, S.cbeg.p, S.ccnt.p, S.xs.p, cycle, tc.bodies);
  } else {
    hipLaunchKernelGGL(k_p2p32<true>, dim3(ceil_div(T.nleaves, kP2PWaves)), dim3(kP2PWaves * kWave), 0, s,
                       T.leaves.p, T.misc.p + 17, T.cbeg.p, T.ccnt.p, T.cx.p, T.xq.p, L.p2p_beg.p, L.p2p_cnt.p,
                       L.p2p.p, S.cbeg.p, S.ccnt.p, S.cx.p, S.xq.p, cycle, tc.bodies);
  }
  XF_LAUNCH();
}

// ---------------------------------------------------------------- direct sum
// FP64 oracle on the device: one workgroup per target, threads stride the
// sources, fixed-order tree reduction (deterministic).
__global__ void __launch_bounds__(256) k_direct(exafmm_body *__restrict__ t, int64_t nt,
                                               const exafmm_body *__restrict__ s, int64_t ns) {
#pragma clang fp contract(off)
  __shared__ double red[4][256];
  const int64_t ti = blockIdx.x;
  if (ti >= nt) return;
  const double xi = t[ti].X[0], yi = t[ti].X[1], zi = t[ti].X[2];
  double pot = 0, ax = 0, ay = 0, az = 0;
  for (int64_t j = threadIdx.x; j < ns; j += 256) {
    const double dx = xi - s[j].X[0], dy = yi - s[j].X[1], dz = zi - s[j].X[2];
    const double R2 = dx * dx + dy * dy + dz * dz;
    if (R2 != 0) {
      const double invR2 = 1.0 / R2;
      const double invR = s[j].q * sqrt(invR2);
      const double f = invR2 * invR;
      pot += invR;
      ax += dx * f;
      ay += dy * f;
      az += dz * f;
    }
  }
  red[0][threadIdx.x] = pot;
  red[1][threadIdx.x] = ax;
  red[2][threadIdx.x] = ay;
  red[3][threadIdx.x] = az;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    t[ti].p += red[0][0];
    t[ti].F[0] -= red[1][0];
    t[ti].F[1] -= red[2][0];
    t[ti].F[2] -= red[3][0];
  }
}

void direct_sum(int, hipStream_t st, exafmm_body *t, int64_t nt, const exafmm_body *s, int64_t ns) {
  if (nt <= 0 || ns <= 0) return;
  hipLaunchKernelGGL(k_direct, dim3(nt), dim3(256), 0, st, t, nt, s, ns);
  XF_LAUNCH();
}

}  // namespace xf
