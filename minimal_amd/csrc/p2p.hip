// p2p.hip -- near-field body-body interactions and the FP64 direct sum.
//
// Reference: P2P 3d/kernel.h:122-150 over each target's P2P list
// (3d/traverse_lazy.h:56), direct 3d/traverse_eager.h:193-202.
// Both add p += q/r and F -= q*dX/r^3 (dX = x_target - x_source) and skip
// pairs at zero distance.
//
// FP32 path: one wavefront per target leaf, sources in lanes.  The leaf's
// list entries are staged in LDS (source range and the shift of the source
// leaf into the target leaf's frame, periodic image included) with a prefix
// of their body counts; lane l takes source 64k + l of the concatenated
// list in round k.  Targets are wave-uniform, 16 at a time (8 or 4 for the
// last ones of a leaf), so each interaction reads nothing from LDS: per lane and target it
// accumulates p and the three force components over its sources, and the
// 64 lane partials are added with a fixed butterfly at the end.
// Coordinates are stored relative to their own leaf centre, so every
// difference is formed from small FP32 numbers.  1/r is the hardware
// reciprocal square root; EXAFMM_PREC_FP32_NEWTON refines it by one Newton
// step and keeps Kahan-compensated partial sums (a second kernel, so the
// default one keeps its register budget).
#include <type_traits>

#include "fmm_internal.h"

namespace xf {

constexpr int kP2PWaves = 4;   // target leaves per workgroup, one wave each
constexpr int kP2PSeg = 256;   // list entries staged per segment
constexpr int kP2PTC = 16;     // targets per register block

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline float wave_sum(float v) {
  v += __shfl_xor(v, 32);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

// Sum over the 64 lanes of TC per-lane values by recursive halving: at each
// step a lane keeps one half of its values and sends the other half to its
// partner (TC - 1 shuffles instead of 6 TC).  Afterwards lane l holds the
// partial sum of value l >> (6 - log2 TC) over the lanes that share its low
// bits; the caller finishes with a butterfly over those bits.
template <int TC>
__device__ inline float halving(float (&v)[TC], int lane) {
  int c = TC;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    if (c > 1) {
      const bool up = lane & m;
#pragma unroll
      for (int i = 0; i < TC / 2; i++) {
        if (i < c / 2) {
          const float lo = v[i], hi = v[i + c / 2];
          v[i] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, m);
        }
      }
      c >>= 1;
    }
  }
  return v[0];
}

// Kahan step: s + v with the rounding error carried in c (s - c is the sum).
__device__ inline void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// The interactions of one staged source per lane with the TC targets (the
// TC chains are independent, so they interleave).  CHECK skips pairs at zero distance (3d/kernel.h:139); only
// rounds that can hold such pairs (the target leaf's own bodies, or any
// round when sources and targets come from different trees) use it.
// REFINE (EXAFMM_PREC_FP32_NEWTON) refines the hardware 1/sqrt by one
// Newton step, y (3 - r^2 y^2) / 2, and adds each term into Kahan-
// compensated partial sums (c*).
template <bool CHECK, int TC, bool REFINE>
__device__ inline void interact_block(const float3 (&xt)[TC], float4 sj, float (&ap)[TC], float (&ax)[TC],
                                      float (&ay)[TC], float (&az)[TC], float (&cp)[TC], float (&cx)[TC],
                                      float (&cy)[TC], float (&cz)[TC]) {
#pragma unroll
  for (int t = 0; t < TC; t++) {
    const float dx = xt[t].x - sj.x, dy = xt[t].y - sj.y, dz = xt[t].z - sj.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    float v = __builtin_amdgcn_rsqf(r2);
    if (REFINE) v = v * fmaf(-0.5f * r2 * v, v, 1.5f);
    const float inv = CHECK ? (r2 > 0.0f ? v : 0.0f) : v;
    const float qi = sj.w * inv;
    const float q3 = qi * inv * inv;
    if (REFINE) {
      kahan_add(ap[t], cp[t], qi);
      kahan_add(ax[t], cx[t], dx * q3);
      kahan_add(ay[t], cy[t], dy * q3);
      kahan_add(az[t], cz[t], dz * q3);
    } else {
      ap[t] += qi;
      ax[t] += dx * q3;
      ay[t] += dy * q3;
      az[t] += dz * q3;
    }
  }
}

// One block of TC wave-uniform targets against the whole list of the leaf.
// Every target slot is computed (slots past tc repeat the last target and
// are not written).  The entry of the target leaf itself is staged first,
// so only the first rounds need the zero-distance check.  Unused source
// lanes sit 1e18 away with q = 0: the pair adds exact zeros.
template <int TC, bool REFINE>
__device__ inline void p2p_block(int tc, const float4 *__restrict__ xt_src, int b0, int nl, double4 Cc,
                                 int self_cell, const int *__restrict__ list, const int *__restrict__ scbeg,
                                 const int *__restrict__ sccnt, const double4 *__restrict__ scx,
                                 const float4 *__restrict__ sxq, double cycle, int *pre, int *sbg, float4 *off,
                                 int *cnt_s, int lane, exafmm_body *__restrict__ out) {
  float3 xt[TC];
#pragma unroll
  for (int t = 0; t < TC; t++) {
    const float4 v = xt_src[min(t, tc - 1)];
    xt[t] = make_float3(v.x, v.y, v.z);
  }
  float ap[TC], ax[TC], ay[TC], az[TC], cp[TC], cx[TC], cy[TC], cz[TC];
#pragma unroll
  for (int t = 0; t < TC; t++) ap[t] = ax[t] = ay[t] = az[t] = cp[t] = cx[t] = cy[t] = cz[t] = 0.0f;
  constexpr int kPer = kP2PSeg / kWave;
  for (int seg = 0; seg < nl; seg += kP2PSeg) {
    const int ne = min(kP2PSeg, nl - seg);
    wave_sync();
    // where is the target leaf's own entry (image 13 = no shift)?
    int self = -1;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int e = lane * kPer + k;
      if (e < ne && list[b0 + seg + e] == ((self_cell << 5) | 13)) self = e;
    }
    const unsigned long long hit = __ballot(self >= 0);
    const int sidx = hit ? __shfl(self, __ffsll((long long)hit) - 1) : -1;
    // stage entry e at position 0 if it is the leaf itself (entry 0 takes its place)
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int e = lane * kPer + k;
      if (e < ne) {
        const int pos = e == sidx ? 0 : (e == 0 && sidx > 0 ? sidx : e);
        const int en = list[b0 + seg + e];
        const int sc = en >> 5, img = en & 31;
        cnt_s[pos] = sccnt[sc];
        sbg[pos] = scbeg[sc];
        const double4 Cs = scx[sc];
        off[pos] = make_float4(float(Cs.x + ((img % 3) - 1) * cycle - Cc.x),
                               float(Cs.y + (((img / 3) % 3) - 1) * cycle - Cc.y),
                               float(Cs.z + ((img / 9) - 1) * cycle - Cc.z), 1.0f);
      }
    }
    wave_sync();
    // prefix of the staged counts into pre[]
    int cnt[kPer], run = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int e = lane * kPer + k;
      cnt[k] = e < ne ? cnt_s[e] : 0;
      run += cnt[k];
    }
    int incl = run;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    int acc = incl - run;
    if (lane == 0) pre[0] = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int e = lane * kPer + k;
      acc += cnt[k];
      if (e < ne) pre[e + 1] = acc;
    }
    wave_sync();
    const int total = pre[ne];
    const int nchecked = self_cell == -2 ? total : (sidx >= 0 ? pre[1] : 0);
    // Round k: lane takes source k*64 + lane; its entry cursor only moves
    // forward.  Sources are fetched two rounds ahead into two register
    // buffers used alternately (no copies, so a fetch stays in flight while
    // the round before it computes).  A lane past the end fetches a valid
    // body and gets the offset (1e18, 1e18, 1e18) and charge factor 0: r^2
    // is about 3e36 (finite, so a Newton step on its rsqrt stays finite),
    // and the pair adds exact zeros (body 0 of the source tree always
    // exists).
    const float4 far_off = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
    int e = 0;
    struct Src {
      float4 c, o;
    };
    auto fetch = [&](int g) {
      const bool in = g < total;
      if (in)
        while (pre[e + 1] <= g) e++;
      const int idx = in ? sbg[e] + g - pre[e] : 0;
      const float4 t = off[e];  // e < ne always
      const float4 ov = make_float4(in ? t.x : far_off.x, in ? t.y : far_off.y, in ? t.z : far_off.z,
                                    in ? t.w : far_off.w);
      return Src{sxq[idx], ov};
    };
    auto step = [&](int k0, Src b, auto check) {
      const float4 sj = make_float4(b.c.x + b.o.x, b.c.y + b.o.y, b.c.z + b.o.z, b.c.w * b.o.w);
      const Src nb = fetch(k0 + 2 * kWave + lane);
      interact_block<decltype(check)::value, TC, REFINE>(xt, sj, ap, ax, ay, az, cp, cx, cy, cz);
      return nb;
    };
    Src b0 = fetch(lane), b1 = fetch(kWave + lane);
    // Rounds go in pairs (an odd count ends with one round of far lanes, so
    // every loop iteration issues the same loads and waits for the older
    // buffer only).  Rounds that may hold zero-distance pairs first (rounded
    // up to a pair), then the rest: two loops, one interaction variant each.
    const int nck = ((nchecked + 2 * kWave - 1) / (2 * kWave)) * (2 * kWave);
    int k0 = 0;
    for (; k0 < total && k0 < nck; k0 += 2 * kWave) {
      b0 = step(k0, b0, std::true_type());
      b1 = step(k0 + kWave, b1, std::true_type());
    }
    for (; k0 < total; k0 += 2 * kWave) {
      b0 = step(k0, b0, std::false_type());
      b1 = step(k0 + kWave, b1, std::false_type());
    }
  }
  // 64-lane sums: halving over the targets, then a butterfly over the 64 / TC
  // lanes that share a target; lane l writes component l % (64 / TC) of
  // target l / (64 / TC) when that remainder is below 4
  constexpr int kLow = kWave / TC;  // lanes sharing a target after the halving
  if (REFINE) {
#pragma unroll
    for (int t = 0; t < TC; t++) {
      ap[t] -= cp[t];
      ax[t] -= cx[t];
      ay[t] -= cy[t];
      az[t] -= cz[t];
    }
  }
  float sp = halving<TC>(ap, lane), sx = halving<TC>(ax, lane), sy = halving<TC>(ay, lane),
        sz = halving<TC>(az, lane);
#pragma unroll
  for (int m = 1; m < kLow; m <<= 1) {
    sp += __shfl_xor(sp, m);
    sx += __shfl_xor(sx, m);
    sy += __shfl_xor(sy, m);
    sz += __shfl_xor(sz, m);
  }
  const int t = lane / kLow, comp = lane % kLow;
  if (comp < 4 && t < tc) {
    exafmm_body *ob = out + t;
    if (comp == 0)
      ob->p += sp;
    else
      ob->F[comp - 1] -= comp == 1 ? sx : (comp == 2 ? sy : sz);
  }
}

// One workgroup's share (block `bid`): kP2PWaves target leaves, one wave each.
template <bool REFINE>
__device__ __attribute__((always_inline)) inline void p2p32_body(int bid, const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                  const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                  const double4 *__restrict__ cx, const float4 *__restrict__ xq,
                                  const int *__restrict__ lbeg, const int *__restrict__ lcnt,
                                  const int *__restrict__ list, const int *__restrict__ scbeg,
                                  const int *__restrict__ sccnt, const double4 *__restrict__ scx,
                                  const float4 *__restrict__ sxq, double cycle, int same_tree,
                                  exafmm_body *__restrict__ bodies) {
  __shared__ int s_pre[kP2PWaves][kP2PSeg + 1];  // prefix of the entries' body counts
  __shared__ int s_beg[kP2PWaves][kP2PSeg];      // first source body of each entry
  __shared__ int s_cnt[kP2PWaves][kP2PSeg];      // body count of each entry
  __shared__ float4 s_off[kP2PWaves][kP2PSeg];   // source leaf frame -> target leaf frame
  // (targets stay in vector registers: a block of them in scalar registers
  // would spill)
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int li = bid * kP2PWaves + w;
  if (li >= *nleaves) return;  // whole wave; the kernel uses wave-level syncs only
  const int c = leaves[li];
  const int tb = cbeg[c], tn = ccnt[c];
  const double4 Cc = cx[c];
  const int b0 = lbeg[c], nl = lcnt[c];
  const int self_cell = same_tree ? c : -2;  // -2: every round checks zero distances
  // REFINE carries a compensation per accumulator: blocks of at most 8
  constexpr int kTC = REFINE ? kP2PTC / 2 : kP2PTC;
  for (int t0 = 0; t0 < tn; t0 += kTC) {
    const int tc = min(kTC, tn - t0);
    if (tc > kTC / 2)
      p2p_block<kTC, REFINE>(tc, xq + tb + t0, b0, nl, Cc, self_cell, list, scbeg, sccnt, scx, sxq, cycle, s_pre[w],
                             s_beg[w], s_off[w], s_cnt[w], lane, bodies + tb + t0);
    else if (!REFINE && tc > kTC / 4)
      p2p_block<kTC / 2, REFINE>(tc, xq + tb + t0, b0, nl, Cc, self_cell, list, scbeg, sccnt, scx, sxq, cycle,
                                 s_pre[w], s_beg[w], s_off[w], s_cnt[w], lane, bodies + tb + t0);
    else  // the last 1..4 targets of a leaf (a quarter of the leaves at 32 bodies each)
      p2p_block<kP2PTC / 4, REFINE>(tc, xq + tb + t0, b0, nl, Cc, self_cell, list, scbeg, sccnt, scx, sxq, cycle,
                                    s_pre[w], s_beg[w], s_off[w], s_cnt[w], lane, bodies + tb + t0);
  }
}

__global__ void __launch_bounds__(kP2PWaves *kWave, 3) k_p2p32(
                                                            const int *__restrict__ leaves, const int *__restrict__ nleaves, const int *__restrict__ cbeg,
                                                            const int *__restrict__ ccnt, const double4 *__restrict__ cx, const float4 *__restrict__ xq,
                                                            const int *__restrict__ lbeg, const int *__restrict__ lcnt, const int *__restrict__ list,
                                                            const int *__restrict__ scbeg, const int *__restrict__ sccnt, const double4 *__restrict__ scx,
                                                            const float4 *__restrict__ sxq, double cycle, int same_tree, exafmm_body *__restrict__ bodies) {
  p2p32_body<false>(blockIdx.x, leaves, nleaves, cbeg, ccnt, cx, xq, lbeg, lcnt, list, scbeg, sccnt, scx, sxq,
                    cycle, same_tree, bodies);
}

// EXAFMM_PREC_FP32_NEWTON: Newton-refined 1/sqrt and compensated sums.
__global__ void __launch_bounds__(kP2PWaves *kWave) k_p2p32_refined(
    const int *__restrict__ leaves, const int *__restrict__ nleaves, const int *__restrict__ cbeg,
    const int *__restrict__ ccnt, const double4 *__restrict__ cx, const float4 *__restrict__ xq,
    const int *__restrict__ lbeg, const int *__restrict__ lcnt, const int *__restrict__ list,
    const int *__restrict__ scbeg, const int *__restrict__ sccnt, const double4 *__restrict__ scx,
    const float4 *__restrict__ sxq, double cycle, int same_tree, exafmm_body *__restrict__ bodies) {
  p2p32_body<true>(blockIdx.x, leaves, nleaves, cbeg, ccnt, cx, xq, lbeg, lcnt, list, scbeg, sccnt, scx, sxq,
                   cycle, same_tree, bodies);
}

// All problems of a batch in one launch: block b belongs to the problem i
// with start[i] <= b < start[i + 1] (each problem's own tree is source and
// target).
__global__ void __launch_bounds__(kP2PWaves *kWave, 3) k_p2p32_grouped(const GroupTree *__restrict__ g,
                                                                    const int *__restrict__ start, int nprob) {
  const int b = blockIdx.x, i = group_of(start, nprob, b);
  const GroupTree &t = g[i];
  p2p32_body<false>(b - start[i], t.leaves, t.nleaves, t.cbeg, t.ccnt, t.cx, t.xq, t.lbeg, t.lcnt, t.list, t.cbeg, t.ccnt,
             t.cx, t.xq, 0.0, 1, t.bodies);
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

int p2p_group_blocks(int64_t nleaves) { return ceil_div(nleaves, kP2PWaves); }

void p2p_grouped(const GroupTree *g, const int *start, int nprob, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_p2p32_grouped, dim3(nblocks), dim3(kP2PWaves * kWave), 0, s, g, start, nprob);
  XF_LAUNCH();
}

void p2p_apply(Ctx &tc, const Ctx &sc, hipStream_t s) {
  const Tree &T = tc.t, &S = sc.t;
  const Lists &L = tc.lists;
  const double cycle = tc.cfg.images > 0 ? tc.cfg.cycle : 0.0;
  if (tc.cfg.p2p_precision == EXAFMM_PREC_FP64) {
    hipLaunchKernelGGL(k_p2p64, dim3(T.nleaves), dim3(64), 0, s, T.leaves.p, T.misc.p + 17, T.cbeg.p, T.ccnt.p,
                       T.xs.p, L.p2p_beg.p, L.p2p_cnt.p, L.p2p.p, S.cbeg.p, S.ccnt.p, S.xs.p, cycle, tc.bodies);
  } else {
    auto *k = tc.cfg.p2p_precision == EXAFMM_PREC_FP32_NEWTON ? k_p2p32_refined : k_p2p32;
    hipLaunchKernelGGL(k, dim3(ceil_div(T.nleaves, kP2PWaves)), dim3(kP2PWaves * kWave), 0, s, T.leaves.p,
                       T.misc.p + 17, T.cbeg.p, T.ccnt.p, T.cx.p, T.xq.p, L.p2p_beg.p, L.p2p_cnt.p, L.p2p.p,
                       S.cbeg.p, S.ccnt.p, S.cx.p, S.xq.p, cycle, int(&tc == &sc), tc.bodies);
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
