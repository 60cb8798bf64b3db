// p2p.hip -- near-field body-body interactions and the FP64 direct sum.
//
// Reference: P2P 3d/kernel.h:122-150 over each target's P2P list
// (3d/traverse_lazy.h:56), direct 3d/traverse_eager.h:193-202.
// Both add p += q/r and F -= q*dX/r^3 (dX = x_target - x_source) and skip
// pairs at zero distance.
//
// FP32 path: one wavefront per target leaf.  The leaf's list entries
// (source leaves, with their periodic image) are staged in LDS with a prefix
// of their body counts; the sources themselves are then staged 256 at a
// time as float(X_s + shift - C_t) -- computed in FP64 from the tree-order
// coordinates and rounded once -- so a body that is both target and source
// gets the same FP32 coordinates either way and its zero distance is exact
// (also between two different trees).  Targets are float(X_t - C_t), the
// tree's leaf-relative copy.
//
// The leaf's targets go in balanced blocks of at most kP2PTC: a block's
// targets are wave-uniform (scalar registers), and lane l takes staged
// source 64k + l in round k, so each round is one LDS read per lane and TC
// interactions per lane, with no padding of the target block.  Per lane and
// target the four sums (p and the three force components) run over the
// lane's sources; at the end the 64 lane partials are added by recursive
// halving across lanes, in a fixed order.
//
// EXAFMM_PREC_FP32_NEWTON refines the hardware 1/sqrt by one Newton step
// y (3 - r^2 y^2) / 2 and compensates the sums: every kP2PFlush rounds the
// lane partials (FP32 sums of at most kP2PFlush terms) are reduced across
// the lanes and added into Kahan-compensated running sums, one per lane
// (a lane holds one target's p or force component after the reduction).
#include <type_traits>

#include "fmm_internal.h"

namespace xf {

constexpr int kP2PWaves = 4;      // target leaves per workgroup, one wave each
constexpr int kP2PSeg = 128;      // list entries staged per segment
constexpr int kP2PChunk = 256;    // sources staged per chunk (4 rounds of 64)
constexpr int kP2PTC = 16;        // targets per block (plain)
#ifndef P2P_FLUSH
#define P2P_FLUSH 64
#endif
constexpr int kP2PFlush = P2P_FLUSH;  // rounds per compensated block (Newton + compensated)
#ifndef P2P_TC_STEP
#define P2P_TC_STEP 1
#endif
constexpr int kP2PTCStep = P2P_TC_STEP;  // block sizes are multiples of this (instantiated variants)

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Sum over the 64 lanes of TC per-lane values by recursive halving: at each
// step a lane keeps one half of its values and sends the other half to its
// partner (TC - 1 shuffles instead of 6 TC).  Afterwards lane l holds the
// partial sum of value l >> (6 - log2 TC) over the lanes that share its low
// bits; the caller finishes with a butterfly over those bits.
template <int TC, class T>
__device__ inline T halving(T (&v)[TC], int lane) {
  int c = TC;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    if (c > 1) {
      const bool up = lane & m;
#pragma unroll
      for (int i = 0; i < TC / 2; i++) {
        if (i < c / 2) {
          const T lo = v[i], hi = v[i + c / 2];
          v[i] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, m);
        }
      }
      c >>= 1;
    }
  }
  return v[0];
}

template <int N>
constexpr int pow2ceil() {
  return N <= 1 ? 1 : 2 * pow2ceil<(N + 1) / 2>();
}

// Kahan step: s + v with the rounding error carried in c (s - c is the sum).
__device__ inline void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

struct P2PLds {
  int pre[kP2PSeg + 1];    // prefix of the staged entries' body counts
  int beg[kP2PSeg];        // first source body of each entry
  int cnt[kP2PSeg];        // body count of each entry
  double dlx[kP2PSeg], dly[kP2PSeg], dlz[kP2PSeg];  // shift - C_t of each entry
  float4 src[kP2PChunk];   // staged sources: x, y, z in the target frame, q
};

// Stages list entries [seg, seg + ne) of the target leaf: the leaf's own
// entry (image 13) moves to position 0; returns the number of sources and
// sets *nself to the body count of the leaf's own entry (0 if absent).
__device__ inline int p2p_stage_entries(P2PLds &L, const int *__restrict__ list, int b0, int seg, int ne,
                                        int self_key, const int *__restrict__ scbeg,
                                        const int *__restrict__ sccnt, double4 Cc, double cycle, int lane,
                                        int *nself) {
  constexpr int kPer = kP2PSeg / kWave;
  wave_sync();
  int self = -1;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int e = lane * kPer + k;
    if (e < ne && list[b0 + seg + e] == self_key) self = e;
  }
  const unsigned long long hit = __ballot(self >= 0);
  const int sidx = hit ? __shfl(self, __ffsll((long long)hit) - 1) : -1;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int e = lane * kPer + k;
    if (e < ne) {
      const int pos = e == sidx ? 0 : (e == 0 && sidx > 0 ? sidx : e);
      const int en = list[b0 + seg + e];
      const int sc = en >> 5, img = en & 31;
      L.cnt[pos] = sccnt[sc];
      L.beg[pos] = scbeg[sc];
      L.dlx[pos] = ((img % 3) - 1) * cycle - Cc.x;
      L.dly[pos] = (((img / 3) % 3) - 1) * cycle - Cc.y;
      L.dlz[pos] = ((img / 9) - 1) * cycle - Cc.z;
    }
  }
  wave_sync();
  int cnt[kPer], run = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int e = lane * kPer + k;
    cnt[k] = e < ne ? L.cnt[e] : 0;
    run += cnt[k];
  }
  int incl = run;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  int acc = incl - run;
  if (lane == 0) L.pre[0] = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int e = lane * kPer + k;
    acc += cnt[k];
    if (e < ne) L.pre[e + 1] = acc;
  }
  wave_sync();
  *nself = sidx >= 0 ? L.pre[1] : 0;
  return L.pre[ne];
}

// Stages sources [c0, c0 + nsrc) of the staged entries, padded to whole
// rounds of 64 with far sources of zero charge (r^2 ~ 3e36: finite, so the
// Newton step stays finite, and the pair adds exact zeros).  e is the lane's
// entry cursor (positions only move forward).
__device__ inline void p2p_stage_sources(P2PLds &L, const double4 *__restrict__ sxs, int c0, int nsrc, int lane,
                                         int &e) {
  const int nround = (nsrc + kWave - 1) / kWave;
  const float4 far_src = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
  wave_sync();  // the previous chunk has been read
#pragma unroll
  for (int k = 0; k < kP2PChunk / kWave; k++) {
    const int p = k * kWave + lane;
    if (p < nsrc) {
      const int g = c0 + p;
      while (L.pre[e + 1] <= g) e++;
      const double4 X = sxs[L.beg[e] + g - L.pre[e]];
      L.src[p] = make_float4(float(X.x + L.dlx[e]), float(X.y + L.dly[e]), float(X.z + L.dlz[e]), float(X.w));
    } else if (k < nround) {
      L.src[p] = far_src;
    }
  }
  wave_sync();
}

// One round: the lane's staged source against the TC block targets.
// CHECK skips pairs at zero distance (3d/kernel.h:139): only rounds that can
// hold such pairs use it.
template <int TC, bool CHECK, bool REFINE>
__device__ inline void p2p_round(float4 s, const float (&xt)[TC], const float (&yt)[TC], const float (&zt)[TC],
                                 float (&ap)[TC], float (&ax)[TC], float (&ay)[TC], float (&az)[TC]) {
#pragma unroll
  for (int t = 0; t < TC; t++) {
    const float dx = xt[t] - s.x, dy = yt[t] - s.y, dz = zt[t] - s.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float y = __builtin_amdgcn_rsqf(r2);
    if (REFINE) {
      const float h = (r2 * y) * y;
      y = y * fmaf(h, -0.5f, 1.5f);
    }
    if (CHECK) y = r2 > 0.0f ? y : 0.0f;
    const float qi = s.w * y;
    const float q3 = qi * (y * y);
    ap[t] += qi;
    ax[t] = fmaf(dx, q3, ax[t]);
    ay[t] = fmaf(dy, q3, ay[t]);
    az[t] = fmaf(dz, q3, az[t]);
  }
}

// Cross-lane totals of the block's four sums: halving over the targets,
// then a butterfly over the lanes that share a target.  Lane l receives
// component l % kLow (p, F_x, F_y, F_z for 0..3) of target l / kLow, with
// kLow = 64 / T2 (T2 = TC rounded up to a power of two); other lanes 0.
template <int TC>
__device__ inline float p2p_lane_total(const float (&ap)[TC], const float (&ax)[TC], const float (&ay)[TC],
                                       const float (&az)[TC], int lane) {
  constexpr int T2 = pow2ceil<TC>();
  constexpr int kLow = kWave / T2;
  float vp[T2], vx[T2], vy[T2], vz[T2];
#pragma unroll
  for (int i = 0; i < T2; i++) {
    vp[i] = i < TC ? ap[i] : 0.0f;
    vx[i] = i < TC ? ax[i] : 0.0f;
    vy[i] = i < TC ? ay[i] : 0.0f;
    vz[i] = i < TC ? az[i] : 0.0f;
  }
  float rp = halving<T2>(vp, lane), rx = halving<T2>(vx, lane), ry = halving<T2>(vy, lane),
        rz = halving<T2>(vz, lane);
#pragma unroll
  for (int m = 1; m < kLow; m <<= 1) {
    rp += __shfl_xor(rp, m);
    rx += __shfl_xor(rx, m);
    ry += __shfl_xor(ry, m);
    rz += __shfl_xor(rz, m);
  }
  const int comp = lane % kLow;
  return comp == 0 ? rp : comp == 1 ? rx : comp == 2 ? ry : comp == 3 ? rz : 0.0f;
}

// Targets t0 .. t0 + tc - 1 (tc <= TC) of a leaf (a wave) against its whole list.
// same_tree: sources and targets are one tree, so only the leaf's own entry
// can hold zero distances; otherwise every round checks.  REFINE: every
// kP2PFlush rounds the lane partials are reduced across lanes and added
// into one Kahan-compensated running sum per lane (target, component).
template <int TC, bool REFINE>
__device__ __attribute__((always_inline)) inline void p2p_block(int t0, int tc, int tb, double4 Cc, int b0, int nl,
                                                                const float4 *__restrict__ xq,
                                                                const int *__restrict__ list,
                                                                const int *__restrict__ scbeg,
                                                                const int *__restrict__ sccnt,
                                                                const double4 *__restrict__ sxs, double cycle,
                                                                int self_key, int same_tree,
                                                                exafmm_body *__restrict__ bodies, P2PLds &L) {
  const int lane = threadIdx.x & (kWave - 1);
  float xt[TC], yt[TC], zt[TC];
#pragma unroll
  for (int t = 0; t < TC; t++) {
    const float4 v = xq[tb + t0 + min(t, tc - 1)];  // wave-uniform address: scalar loads
    xt[t] = v.x;
    yt[t] = v.y;
    zt[t] = v.z;
  }
  float ap[TC], ax[TC], ay[TC], az[TC];
#pragma unroll
  for (int t = 0; t < TC; t++) ap[t] = ax[t] = ay[t] = az[t] = 0.0f;
  float ks = 0.0f, kc = 0.0f;  // compensated sum of this lane's (target, component)
  int since = 0;               // rounds since the last flush
  auto flush = [&]() {
    kahan_add(ks, kc, p2p_lane_total<TC>(ap, ax, ay, az, lane));
#pragma unroll
    for (int t = 0; t < TC; t++) ap[t] = ax[t] = ay[t] = az[t] = 0.0f;
    since = 0;
  };
  for (int seg = 0; seg < nl; seg += kP2PSeg) {
    const int ne = min(kP2PSeg, nl - seg);
    int nself = 0;
    const int total = p2p_stage_entries(L, list, b0, seg, ne, self_key, scbeg, sccnt, Cc, cycle, lane, &nself);
    // sources that can sit at zero distance: the leaf's own (staged first),
    // or all of them when the trees differ
    const int nchecked = same_tree ? nself : total;
    int e = 0;
    for (int c0 = 0; c0 < total; c0 += kP2PChunk) {
      const int nsrc = min(kP2PChunk, total - c0);
      const int nround = (nsrc + kWave - 1) / kWave;
      p2p_stage_sources(L, sxs, c0, nsrc, lane, e);
      const int nck = min(nround, max(0, (nchecked - c0 + kWave - 1) / kWave));
      int k = 0;
      for (; k < nck; k++) p2p_round<TC, true, REFINE>(L.src[k * kWave + lane], xt, yt, zt, ap, ax, ay, az);
      for (; k < nround; k++) p2p_round<TC, false, REFINE>(L.src[k * kWave + lane], xt, yt, zt, ap, ax, ay, az);
      if (REFINE) {
        since += nround;
        if (since >= kP2PFlush) flush();
      }
    }
  }
  constexpr int kLow = kWave / pow2ceil<TC>();
  const int t = lane / kLow, comp = lane % kLow;
  double v;
  if (REFINE) {
    flush();
    v = double(ks) - double(kc);
  } else {
    v = double(p2p_lane_total<TC>(ap, ax, ay, az, lane));
  }
  if (comp < 4 && t < tc) {  // slots tc .. TC - 1 repeat the last target and are not written
    exafmm_body *ob = bodies + tb + t0 + t;
    if (comp == 0)
      ob->p += v;
    else
      ob->F[comp - 1] -= v;
  }
}

// One target leaf (a wave): its targets in nb = ceil(tn / kP2PTC) blocks of
// balanced size, each run by the variant of the next multiple of
// kP2PTCStep (the extra slots repeat the block's last target).
template <bool REFINE>
__device__ __attribute__((always_inline)) inline void p2p_leaf(
    int li, const int *__restrict__ leaves, const int *__restrict__ cbeg, const int *__restrict__ ccnt,
    const double4 *__restrict__ cx, const float4 *__restrict__ xq, const int *__restrict__ lbeg,
    const int *__restrict__ lcnt, const int *__restrict__ list, const int *__restrict__ scbeg,
    const int *__restrict__ sccnt, const double4 *__restrict__ sxs, double cycle, int same_tree,
    exafmm_body *__restrict__ bodies, P2PLds &L) {
  const int c = leaves[li];
  const int tb = cbeg[c], tn = ccnt[c];
  const double4 Cc = cx[c];
  const int b0 = lbeg[c], nl = lcnt[c];
  const int self_key = same_tree ? ((c << 5) | 13) : -1;
  const int nb = (tn + kP2PTC - 1) / kP2PTC;
  const int per = (tn + nb - 1) / nb;
  for (int t0 = 0; t0 < tn; t0 += per) {
    const int tc = min(per, tn - t0);
    const int tv = (tc + kP2PTCStep - 1) / kP2PTCStep * kP2PTCStep;
#define XF_P2P_BLOCK(N)                                                                                        \
  case N:                                                                                                      \
    if (N % kP2PTCStep == 0)                                                                                   \
      p2p_block<N, REFINE>(t0, tc, tb, Cc, b0, nl, xq, list, scbeg, sccnt, sxs, cycle, self_key, same_tree, bodies, \
                           L);                                                                                 \
    break;
    switch (tv) {
      XF_P2P_BLOCK(1)
      XF_P2P_BLOCK(2)
      XF_P2P_BLOCK(3)
      XF_P2P_BLOCK(4)
      XF_P2P_BLOCK(5)
      XF_P2P_BLOCK(6)
      XF_P2P_BLOCK(7)
      XF_P2P_BLOCK(8)
      XF_P2P_BLOCK(9)
      XF_P2P_BLOCK(10)
      XF_P2P_BLOCK(11)
      XF_P2P_BLOCK(12)
      XF_P2P_BLOCK(13)
      XF_P2P_BLOCK(14)
      XF_P2P_BLOCK(15)
      XF_P2P_BLOCK(16)
      default:
        break;
    }
#undef XF_P2P_BLOCK
  }
}

// A P2P launched before the host has read the traversal's flags: it does
// nothing if that pass is to be repeated (list capacity or key table full)
// or failed (work queue overflow, offset out of key range), as its lists
// are then incomplete.
__device__ inline bool p2p_skip(const int *list_flags) {
  return list_flags && (__atomic_load_n(list_flags, __ATOMIC_RELAXED) & (kListRetry | 4 | 16));
}

template <bool REFINE>
__global__ void __launch_bounds__(kP2PWaves *kWave) __attribute__((amdgpu_waves_per_eu(REFINE ? 3 : 4))) k_p2p32(
    const int *__restrict__ leaves, const int *__restrict__ nleaves, const int *__restrict__ cbeg,
    const int *__restrict__ ccnt, const double4 *__restrict__ cx, const float4 *__restrict__ xq,
    const int *__restrict__ lbeg, const int *__restrict__ lcnt, const int *__restrict__ list,
    const int *__restrict__ scbeg, const int *__restrict__ sccnt, const double4 *__restrict__ sxs, double cycle,
    int same_tree, exafmm_body *__restrict__ bodies, const int *__restrict__ list_flags) {
  __shared__ P2PLds lds[kP2PWaves];
  // wave-uniform (scalar) leaf index: the per-leaf quantities, the block
  // targets and the loop bounds stay in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int li = blockIdx.x * kP2PWaves + w;
  if (li >= *nleaves || p2p_skip(list_flags)) return;  // whole wave; the kernel uses wave-level syncs only
  p2p_leaf<REFINE>(li, leaves, cbeg, ccnt, cx, xq, lbeg, lcnt, list, scbeg, sccnt, sxs, cycle, same_tree, bodies,
                   lds[w]);
}

// All problems of a batch in one launch: block b belongs to the problem i
// with start[i] <= b < start[i + 1] (each problem's own tree is source and
// target).
__global__ void __launch_bounds__(kP2PWaves *kWave) k_p2p32_grouped(const GroupTree *__restrict__ g,
                                                                    const int *__restrict__ start, int nprob) {
  __shared__ P2PLds lds[kP2PWaves];
  const int b = blockIdx.x, i = group_of(start, nprob, b);
  const GroupTree &t = g[i];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int li = (b - start[i]) * kP2PWaves + w;
  if (li >= *t.nleaves) return;
  p2p_leaf<false>(li, t.leaves, t.cbeg, t.ccnt, t.cx, t.xq, t.lbeg, t.lcnt, t.list, t.cbeg, t.ccnt, t.xs, 0.0, 1,
                  t.bodies, lds[w]);
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
                                              exafmm_body *__restrict__ bodies,
                                              const int *__restrict__ list_flags) {
#pragma clang fp contract(off)
  const int li = blockIdx.x;
  if (li >= *nleaves || p2p_skip(list_flags)) return;
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

int p2p_group_blocks(int64_t nleaves) { return ceil_div(nleaves, kP2PWaves); }

void p2p_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_p2p32_grouped, dim3(nblocks), dim3(kP2PWaves * kWave), 0, s, g, start, nprob);
  XF_LAUNCH();
}

void p2p_apply(Ctx &tc, const Ctx &sc, hipStream_t s, const int *list_flags) {
  const Tree &T = tc.t, &S = sc.t;
  const Lists &L = tc.lists;
  const double cycle = tc.cfg.images > 0 ? tc.cfg.cycle : 0.0;
  if (tc.cfg.p2p_precision == EXAFMM_PREC_FP64) {
    hipLaunchKernelGGL(k_p2p64, dim3(T.nleaves), dim3(64), 0, s, T.leaves.p, T.misc.p + 17, T.cbeg.p, T.ccnt.p,
                       T.xs.p, L.p2p_beg.p, L.p2p_cnt.p, L.p2p.p, S.cbeg.p, S.ccnt.p, S.xs.p, cycle, tc.bodies,
                       list_flags);
  } else {
    auto *k = tc.cfg.p2p_precision == EXAFMM_PREC_FP32_NEWTON ? k_p2p32<true> : k_p2p32<false>;
    hipLaunchKernelGGL(k, dim3(ceil_div(T.nleaves, kP2PWaves)), dim3(kP2PWaves * kWave), 0, s, T.leaves.p,
                       T.misc.p + 17, T.cbeg.p, T.ccnt.p, T.cx.p, T.xq.p, L.p2p_beg.p, L.p2p_cnt.p, L.p2p.p,
                       S.cbeg.p, S.ccnt.p, S.xs.p, cycle, int(&tc == &sc), tc.bodies, list_flags);
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
