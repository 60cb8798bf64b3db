// traverse.hip -- dual-tree traversal emitting flat M2L / P2P lists.
//
// Reference: the dual-tree rule of 3d/traverse_eager.h:147-165 and its lazy
// list form 3d/traverse_lazy.h:27-43:
//   MAC |Xi-Xj|^2 theta^2 > (Ri+Rj)^2          -> M2L(Ci, Cj)
//   both leaves                                -> P2P(Ci, Cj)
//   Cj leaf or (Ri >= Rj and Ci not a leaf)    -> split Ci
//   otherwise                                  -> split Cj
// Target-centric, level-synchronous form: a pair (c, b) is visited iff b
// is handed down from c's parent (the parent's pair split on the target
// side) or b is a child of a pair (c, parent(b)) split on the source side.
// One wavefront per target cell works through its pairs breadth-first in
// LDS; every target's lists are therefore built by one wave, in a fixed
// order, with no atomics.  Each level runs twice: a count pass, an
// exclusive scan, and a fill pass writing at the scanned offsets.
// The MAC is evaluated in FP64 with contraction disabled, so each pair is
// classified exactly as the reference classifies it.
//
// Periodic boundary (3dp/traverse_lazy.h:39-57, 124-135): the source tree
// is visited in its 27 near images; every list entry carries its image
// index (packed as cell * 32 + image, image = ix+1 + 3(iy+1) + 9(iz+1)) and
// the MAC uses dX = Xc - Xb - iX * cycle.  Free space uses image 13 (0,0,0).
#include "fmm_internal.h"

namespace xf {

constexpr int kTravWaves = 4;
constexpr int kQCap = 2048;
constexpr unsigned long long kEmpty = ~0ull;

struct TravArgs {
  // target tree
  const double4 *cx;
  const int *cnch, *clevel, *cparent;
  // source tree
  const double4 *scx;
  const int *scnch, *scchild, *sclevel;
  double theta;
  double cycle;
  int periodic;
  int lstart, lcount;
  // hand-down lists
  const int *in_beg, *in_cnt, *in;
  int *out_cnt, *out_beg, *out;
  // results
  int *m2l_cnt, *m2l_beg;
  int2 *m2l;
  int *p2p_cnt, *p2p_beg, *p2p;
  int *flags;
  int64_t cap_m2l, cap_p2p, cap_out;
  // translation keys
  unsigned long long *hslots;
  int hmask;
  int want_keys;
  int self_start;  // first level: each cell starts from (itself, no shift) -- a forest's problem roots
};

__device__ inline unsigned long long mix64(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Key of the M2L translation (c <- b): integer offset (Xc - Xb) / Rf in units
// of the finer radius Rf, level difference, and -- periodic only -- the
// image and the level of Rf (the image shift is not a whole number of
// cell widths).  Under radius scaling of M and L the operator depends on
// nothing else.  Layout: dx,dy,dz 16 bits each, image 5, level 5, diff 6.
__device__ inline bool m2l_key(double4 Xc, int lc, double4 Xb, int lb, int img, int periodic,
                               unsigned long long &key) {
  // The offset is a whole number of finer radii (up to the rounding of the
  // centres), so one reciprocal serves the three components: its relative
  // error moves an offset of at most 32767 by far less than the 0.5 rint
  // tolerates.
  const double inv = 1.0 / fmin(Xc.w, Xb.w);
  const double dx = rint((Xc.x - Xb.x) * inv), dy = rint((Xc.y - Xb.y) * inv), dz = rint((Xc.z - Xb.z) * inv);
  const double lim = 32767.0;
  const int ld = lc - lb;
  if (fabs(dx) > lim || fabs(dy) > lim || fabs(dz) > lim || ld < -31 || ld > 31) return false;
  const unsigned long long ix = (unsigned long long)((long long)dx + 32768);
  const unsigned long long iy = (unsigned long long)((long long)dy + 32768);
  const unsigned long long iz = (unsigned long long)((long long)dz + 32768);
  const unsigned long long im = periodic ? unsigned(img) : 13u;
  const unsigned long long lf = periodic ? unsigned(max(lc, lb)) : 0u;
  key = (ix << 48) | (iy << 32) | (iz << 16) | (im << 11) | (lf << 6) | (unsigned long long)(ld + 32);
  return true;
}

// Slots only go from empty to a key during a pass, so the probe reads at
// workgroup scope (served from the CU's L1 once the few hundred live keys
// are there): a stale read shows either the empty value, and the CAS then
// returns the slot's real content, or another key, and the probe moves on;
// it never shows a key the slot does not hold.
__device__ inline int key_insert(unsigned long long *slots, int mask, unsigned long long key, int *flags) {
  int h = int(mix64(key) & (unsigned long long)mask);
  for (int probe = 0; probe <= mask; probe++) {
    unsigned long long v = __hip_atomic_load(&slots[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (v == key) return h;
    if (v == kEmpty) {
      const unsigned long long prev = atomicCAS(&slots[h], kEmpty, key);
      if (prev == kEmpty || prev == key) return h;
    }
    h = (h + 1) & mask;
  }
  atomicOr(flags, 8);
  return 0;
}

__device__ inline int wave_excl_scan(int v, int lane, int &total) {
  int x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  total = __shfl(x, kWave - 1, kWave);
  return x - v;
}

template <bool FILL>
__global__ void __launch_bounds__(kTravWaves *kWave) k_traverse(TravArgs a) {
  __shared__ int queue[kTravWaves][kQCap];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int t = blockIdx.x * kTravWaves + w;
  if (t >= a.lcount) return;
  const int c = a.lstart + t;
  int *q = queue[w];
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;

  // initial pairs: handed down from the parent, or (root, root)
  int tail;
  if (a.in) {
    const int par = a.cparent[c];
    const int ib = a.in_beg[par];
    tail = a.in_cnt[par];
    if (tail > kQCap || int64_t(ib) + tail > a.cap_out) {
      if (lane == 0) atomicOr(a.flags, 4);
      return;
    }
    for (int i = lane; i < tail; i += kWave) q[i] = a.in[ib + i];
  } else if (a.self_start) {
    tail = 1;
    if (lane == 0) q[0] = (c << 5) | 13;
  } else if (a.periodic) {
    tail = 27;
    if (lane < 27) q[lane] = lane;  // (root, image) for the 27 near images
  } else {
    tail = 1;
    if (lane == 0) q[0] = 13;  // (root, no shift)
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

  const double4 Xc = a.cx[c];
  const int leafc = a.cnch[c] == 0;
  const int lc = a.clevel[c];
  const double th2 = a.theta;
  int head = 0;
  int nm = 0, np = 0, no = 0;
  int bm = 0, bp = 0, bo = 0;
  if (FILL) {
    bm = a.m2l_beg[c];
    bp = a.p2p_beg[c];
    bo = a.out_beg[c];
  }
  while (head < tail) {
    const int k = min(kWave, tail - head);
    int b = -1, kind = -1, nchb = 0, lb = 0;
    double4 Xb;
    int img = 13;
    if (lane < k) {
      const int e = q[(head + lane) % kQCap];
      b = e >> 5;
      img = e & 31;
      Xb = a.scx[b];
      nchb = a.scnch[b];
      if (FILL && a.want_keys) lb = a.sclevel[b];  // loaded beside the centre, ahead of the MAC
      const double sx = ((img % 3) - 1) * a.cycle, sy = (((img / 3) % 3) - 1) * a.cycle,
                   sz = ((img / 9) - 1) * a.cycle;
      bool mac;
      {
#pragma clang fp contract(off)
        const double dx = Xc.x - Xb.x - sx, dy = Xc.y - Xb.y - sy, dz = Xc.z - Xb.z - sz;
        const double R2 = (dx * dx + dy * dy + dz * dz) * th2 * th2;
        const double Rs = Xc.w + Xb.w;
        mac = R2 > Rs * Rs;
      }
      if (mac)
        kind = 0;
      else if (leafc && nchb == 0)
        kind = 1;
      else if (nchb == 0 || (Xc.w >= Xb.w && !leafc))
        kind = 2;
      else
        kind = 3;
    }
    const unsigned long long mm = __ballot(kind == 0), mp = __ballot(kind == 1), mo = __ballot(kind == 2);
    if (FILL) {
      if (kind == 0) {
        const int pos = bm + nm + __popcll(mm & lt);
        int slot = 0;
        if (a.want_keys) {
          unsigned long long key;
          if (m2l_key(Xc, lc, Xb, lb, img, a.periodic, key))
            slot = key_insert(a.hslots, a.hmask, key, a.flags);
          else
            atomicOr(a.flags, 16);
        }
        if (pos < a.cap_m2l) a.m2l[pos] = make_int2((b << 5) | img, slot);
      } else if (kind == 1) {
        const int pos = bp + np + __popcll(mp & lt);
        if (pos < a.cap_p2p) a.p2p[pos] = (b << 5) | img;
      } else if (kind == 2) {
        const int pos = bo + no + __popcll(mo & lt);
        if (pos < a.cap_out) a.out[pos] = (b << 5) | img;
      }
    }
    nm += __popcll(mm);
    np += __popcll(mp);
    no += __popcll(mo);
    // split the source: enqueue its children
    const int nadd = kind == 3 ? nchb : 0;
    int total;
    const int off = wave_excl_scan(nadd, lane, total);
    if (tail - head - k + total > kQCap) {
      if (lane == 0) atomicOr(a.flags, 4);
      break;
    }
    if (nadd) {
      const int first = a.scchild[b];
      for (int i = 0; i < nadd; i++) q[(tail + off + i) % kQCap] = ((first + i) << 5) | img;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    head += k;
    tail += total;
  }
  if (!FILL && lane == 0) {
    a.m2l_cnt[c] = nm;
    a.p2p_cnt[c] = np;
    a.out_cnt[c] = no;
  }
}

// per-level offsets: exclusive scans of the three counts + running totals
__global__ void k_hash_clear(unsigned long long *slots, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) slots[i] = kEmpty;
}

__global__ void k_hash_collect(const unsigned long long *__restrict__ slots, int n,
                               unsigned long long *__restrict__ keys, int *__restrict__ sl, int *__restrict__ count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && slots[i] != kEmpty) {
    const int p = atomicAdd(count, 1);
    keys[p] = slots[i];
    sl[p] = i;
  }
}

// M2L entries carry their hash slot out of the traversal; after the keys are
// sorted the slot becomes the key's rank (the id the operators are stored by).
// The count is the device total, capped at the list capacity; flags (if
// given) skip a pass that is repeated or ranked by the radix sort instead.
__global__ void k_entry_ids(int2 *__restrict__ m2l, const int *__restrict__ total, int64_t cap,
                            const int *__restrict__ slot2id, const int *__restrict__ flags) {
  if (flags && (*flags & (1 | 8 | 32))) return;
  const int64_t n = min(int64_t(*total), cap);
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    m2l[i].y = slot2id[m2l[i].y];
}

// Key sort for small key sets (the usual case: hundreds to a few thousand
// distinct translations): one workgroup, keys in LDS, each key's rank is the
// number of smaller keys (keys are distinct).  Writes the sorted keys in
// place and the slot -> rank map.  A larger set raises bit 32 of *flags
// and is left to the radix sort.
constexpr int kRankMax = 6144;
__global__ void __launch_bounds__(1024) k_key_rank(unsigned long long *__restrict__ keys, const int *__restrict__ sl,
                                                   const int *__restrict__ count, int *__restrict__ slot2id,
                                                   int *__restrict__ flags) {
  __shared__ unsigned long long K[kRankMax];
  const int n = *count;
  if (n > kRankMax) {
    if (threadIdx.x == 0) atomicOr(flags, 32);
    return;
  }
  for (int i = threadIdx.x; i < n; i += 1024) K[i] = keys[i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024) {
    const unsigned long long k = K[i];
    int r = 0;
    for (int j = 0; j < n; j++) r += K[j] < k;
    keys[r] = k;
    slot2id[sl[i]] = r;
  }
}

__global__ void k_slot_ids(const int *__restrict__ sl, int n, int *__restrict__ slot2id) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) slot2id[sl[i]] = i;
}

void traverse(Ctx &tc, const Ctx &sc, const std::function<void()> &after_kernels) {
  Tree &T = tc.t;
  const Tree &S = sc.t;
  Lists &Lr = tc.lists;
  hipStream_t s = tc.s0;
  const int64_t nc = T.ncells;
  const bool keys = (&tc == &sc);
  if (S.ncells >= (int64_t(1) << 26)) fail(EXAFMM_ERR_CAPACITY, "more than 2^26 cells");
  Lr.m2l_beg.ensure(nc);
  Lr.m2l_cnt.ensure(nc);
  Lr.p2p_beg.ensure(nc);
  Lr.p2p_cnt.ensure(nc);
  Lr.out_beg.ensure(nc);
  Lr.out_cnt.ensure(nc);
  Lr.tot.ensure(8);
  if (Lr.cap_m2l < 200 * nc) Lr.cap_m2l = 200 * nc;
  if (Lr.cap_p2p < 96 * nc) Lr.cap_p2p = 96 * nc;
  if (Lr.cap_out < 96 * nc) Lr.cap_out = 96 * nc;
  if (keys && Lr.hcap == 0) Lr.hcap = 1 << 16;

  int tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool done = false;
  for (int attempt = 0; attempt < 4; attempt++) {
    Lr.m2l.ensure(Lr.cap_m2l);
    Lr.p2p.ensure(Lr.cap_p2p);
    Lr.out.ensure(Lr.cap_out);
    XF_HIP(hipMemsetAsync(Lr.tot.p, 0, 8 * sizeof(int), s));
    // capacities on the device, uploaded when they change (from pinned
    // memory: an asynchronous copy)
    const int64_t caps[3] = {Lr.cap_m2l, Lr.cap_p2p, Lr.cap_out};
    if (!Lr.capd.p || !std::equal(caps, caps + 3, Lr.caps_up)) {
      Lr.capd.ensure(3);
      Lr.caps_pin.ensure(3);
      std::copy(caps, caps + 3, Lr.caps_pin.p);
      XF_HIP(hipMemcpyAsync(Lr.capd.p, Lr.caps_pin.p, sizeof(caps), hipMemcpyHostToDevice, s));
      std::copy(caps, caps + 3, Lr.caps_up);
    }
    if (keys) {
      Lr.hslots.ensure(Lr.hcap);
      hipLaunchKernelGGL(k_hash_clear, dim3(ceil_div(Lr.hcap, 256)), dim3(256), 0, s, Lr.hslots.p, Lr.hcap);
      XF_LAUNCH();
    }
    TravArgs a{};
    a.cx = T.cx.p;
    a.cnch = T.cnch.p;
    a.clevel = T.clevel.p;
    a.cparent = T.cparent.p;
    a.scx = S.cx.p;
    a.scnch = S.cnch.p;
    a.scchild = S.cchild.p;
    a.sclevel = S.clevel.p;
    a.theta = tc.cfg.theta;
    a.periodic = tc.cfg.images > 0;
    a.cycle = tc.cfg.images > 0 ? tc.cfg.cycle : 0.0;
    a.out_cnt = Lr.out_cnt.p;
    a.out_beg = Lr.out_beg.p;
    a.out = Lr.out.p;
    a.m2l_cnt = Lr.m2l_cnt.p;
    a.m2l_beg = Lr.m2l_beg.p;
    a.m2l = Lr.m2l.p;
    a.p2p_cnt = Lr.p2p_cnt.p;
    a.p2p_beg = Lr.p2p_beg.p;
    a.p2p = Lr.p2p.p;
    a.flags = Lr.tot.p + 3;
    a.cap_m2l = Lr.cap_m2l;
    a.cap_p2p = Lr.cap_p2p;
    a.cap_out = Lr.cap_out;
    a.hslots = Lr.hslots.p;
    a.hmask = Lr.hcap - 1;
    a.want_keys = keys;
    // a forest starts at its problems' roots (level lroot): the cells above
    // are never paired, their lists stay empty
    const int l0 = T.lroot;
    if (l0 > 0 && &tc != &sc) fail(EXAFMM_ERR_UNSUPPORTED, "a forest's horizontal pass needs source == target");
    a.self_start = l0 > 0;
    if (l0 > 0) {
      const size_t nv = size_t(T.lstart[l0]);
      for (int *z : {Lr.m2l_cnt.p, Lr.m2l_beg.p, Lr.p2p_cnt.p, Lr.p2p_beg.p, Lr.out_cnt.p, Lr.out_beg.p})
        XF_HIP(hipMemsetAsync(z, 0, sizeof(int) * nv, s));
    }
    for (int l = l0; l <= T.depth; l++) {
      a.lstart = T.lstart[l];
      a.lcount = T.lstart[l + 1] - T.lstart[l];
      a.in = l > l0 ? Lr.out.p : nullptr;
      a.in_beg = Lr.out_beg.p;
      a.in_cnt = Lr.out_cnt.p;
      const int g = ceil_div(a.lcount, kTravWaves);
      hipLaunchKernelGGL(k_traverse<false>, dim3(g), dim3(kTravWaves * kWave), 0, s, a);
      XF_LAUNCH();
      scan_exclusive3(Lr.m2l_cnt.p + a.lstart, Lr.p2p_cnt.p + a.lstart, Lr.out_cnt.p + a.lstart,
                      Lr.m2l_beg.p + a.lstart, Lr.p2p_beg.p + a.lstart, Lr.out_beg.p + a.lstart, a.lcount, Lr.tot.p,
                      Lr.capd.p, Lr.tot.p + 3, Lr.scan_tmp, s);
      hipLaunchKernelGGL(k_traverse<true>, dim3(g), dim3(kTravWaves * kWave), 0, s, a);
      XF_LAUNCH();
    }
    if (after_kernels) after_kernels();
    if (keys) {
      // distinct translation keys out of the hash table (discarded if the
      // pass has to be repeated), counted into tot[6]
      Lr.keys.ensure(size_t(Lr.hcap) * 2);
      Lr.kscratch.ensure(size_t(Lr.hcap) * 2 + 8);
      Lr.slot2id.ensure(Lr.hcap);
      hipLaunchKernelGGL(k_hash_collect, dim3(ceil_div(Lr.hcap, 256)), dim3(256), 0, s, Lr.hslots.p, Lr.hcap,
                         Lr.keys.p, Lr.kscratch.p, Lr.tot.p + 6);
      XF_LAUNCH();
      // The usual small key set is ranked, and the entries renumbered,
      // before the totals are read, so the read overlaps them (counts from
      // the device; a repeated pass redoes it all, a set above kRankMax is
      // flagged and sorted below).
      hipLaunchKernelGGL(k_key_rank, dim3(1), dim3(1024), 0, s, Lr.keys.p, Lr.kscratch.p, Lr.tot.p + 6,
                         Lr.slot2id.p, Lr.tot.p + 3);
      XF_LAUNCH();
      hipLaunchKernelGGL(k_entry_ids, dim3(int(std::min<int64_t>(ceil_div(Lr.cap_m2l, 256), 4096))), dim3(256), 0, s,
                         Lr.m2l.p, Lr.tot.p, Lr.cap_m2l, Lr.slot2id.p, Lr.tot.p + 3);
      XF_LAUNCH();
    }
    tc.pin_i.ensure(kPinInts);
    XF_HIP(hipMemcpyAsync(tc.pin_i.p + kPinTot, Lr.tot.p, sizeof(tot), hipMemcpyDeviceToHost, s));
    stream_wait(s);
    std::copy(tc.pin_i.p + kPinTot, tc.pin_i.p + kPinTot + 8, tot);
    const int flags = tot[3];
    // work started by after_kernels ends before a repeated pass resets the flags
    if (after_kernels && (flags & kListRetry)) stream_wait(tc.s1);
    if (flags & 4) fail(EXAFMM_ERR_CAPACITY, "traversal work queue overflow (pairs per target cell > 2048)");
    if (flags & 16) fail(EXAFMM_ERR_CAPACITY, "M2L offset out of key range");
    if (flags & 1) {  // list capacity: retry with the exact totals
      Lr.cap_m2l = std::max<int64_t>(Lr.cap_m2l * 2, tot[0]);
      Lr.cap_p2p = std::max<int64_t>(Lr.cap_p2p * 2, tot[1]);
      Lr.cap_out = std::max<int64_t>(Lr.cap_out * 2, tot[2]);
      continue;
    }
    if (flags & 8) {  // key table full
      Lr.hcap *= 4;
      continue;
    }
    Lr.nm2l = tot[0];
    Lr.np2p = tot[1];
    done = true;
    break;
  }
  if (!done) fail(EXAFMM_ERR_CAPACITY, "interaction lists or key table still full after 4 passes");
  Lr.valid = true;
  Lr.nkeys = 0;
  if (keys) {
    unsigned long long *k1 = Lr.keys.p, *k2 = Lr.keys.p + Lr.hcap;
    int *v1 = Lr.kscratch.p, *v2 = Lr.kscratch.p + Lr.hcap;
    const int nk = tot[6];
    if (tot[3] & 32) {  // more keys than k_key_rank holds
      radix_sort_pairs(k1, v1, k2, v2, nk, 0, 64, T.rs_hist, s);
      hipLaunchKernelGGL(k_slot_ids, dim3(ceil_div(std::max(nk, 1), 256)), dim3(256), 0, s, v1, nk, Lr.slot2id.p);
      XF_LAUNCH();
      if (k1 != Lr.keys.p)
        XF_HIP(hipMemcpyAsync(Lr.keys.p, k1, sizeof(unsigned long long) * nk, hipMemcpyDeviceToDevice, s));
      if (Lr.nm2l > 0) {
        hipLaunchKernelGGL(k_entry_ids, dim3(std::min(ceil_div(Lr.nm2l, 256), 4096)), dim3(256), 0, s, Lr.m2l.p,
                           Lr.tot.p, Lr.cap_m2l, Lr.slot2id.p, (const int *)nullptr);
        XF_LAUNCH();
      }
    }
    Lr.nkeys = nk;
  }
}

}  // namespace xf
