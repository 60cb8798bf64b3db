// tree.hip -- octree construction on the GPU, bit-exact with the reference.
//
// Reference: 3d/build_tree.h:7-96.  The reference builds the tree top-down
// by recursive stable octant partitions against FP64 cell centres
// (child centre = parent centre +/- R0/2^(l+1), strict '>'), stopping at
// NBODY <= ncrit.  Here each body's full octant path is computed once from
// the same FP64 centre sequence (63-bit key, 21 levels), the keys are
// radix-sorted (stable, so bodies with equal prefixes keep input order),
// levels are split by binary search on the sorted keys, and each leaf's
// bodies are put back into input order -- which is the order the
// reference's stable partitions leave inside a leaf.  Cells are stored
// level by level (BFS); the reference's numbering (children blocks in DFS
// preorder of their parents) is computed on export.
#include <algorithm>

#include "fmm_internal.h"

namespace xf {

constexpr int kHistLevels = 7;
__host__ __device__ constexpr int hist_off(int l) {  // offset of level l (1..7) in the histogram array
  return l <= 1 ? 0 : hist_off(l - 1) + (1 << (3 * (l - 1)));
}
constexpr int kHistTotal = hist_off(kHistLevels + 1);

// ---------------------------------------------------------------- bounds
__global__ void __launch_bounds__(256) k_bounds_part(const exafmm_body *__restrict__ b, int64_t n,
                                                     double *__restrict__ part, int *__restrict__ flags) {
  __shared__ double s[6][256];
  double lo[3] = {b[0].X[0], b[0].X[1], b[0].X[2]};
  double hi[3] = {lo[0], lo[1], lo[2]};
  int bad = 0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    for (int d = 0; d < 3; d++) {
      const double x = b[i].X[d];
      if (!isfinite(x)) bad = 1;
      lo[d] = fmin(x, lo[d]);
      hi[d] = fmax(x, hi[d]);
    }
    if (!isfinite(b[i].q)) bad = 1;
  }
  if (bad) atomicOr(flags, 1);
  for (int d = 0; d < 3; d++) {
    s[d][threadIdx.x] = lo[d];
    s[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + o]);
        s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// X0 = (max+min)/2, R0 = max half-extent * 1.00001 (3d/build_tree.h:7-21).
__global__ void __launch_bounds__(256) k_bounds_final(const double *__restrict__ part, int nparts,
                                                      double *__restrict__ root) {
  __shared__ double s[6][256];
  const int t = threadIdx.x;
  double v[6];
  for (int d = 0; d < 6; d++) v[d] = part[d];
  for (int p = t; p < nparts; p += 256)
    for (int d = 0; d < 3; d++) {
      v[d] = fmin(part[p * 6 + d], v[d]);
      v[3 + d] = fmax(part[p * 6 + 3 + d], v[3 + d]);
    }
  for (int d = 0; d < 6; d++) s[d][t] = v[d];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int d = 0; d < 3; d++) {
        s[d][t] = fmin(s[d][t], s[d][t + o]);
        s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + o]);
      }
    __syncthreads();
  }
  if (t != 0) return;
  double lo[3], hi[3], X0[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = s[d][0];
    hi[d] = s[3 + d][0];
    X0[d] = (hi[d] + lo[d]) / 2;
  }
  double R0 = 0;
  for (int d = 0; d < 3; d++) {
    R0 = fmax(X0[d] - lo[d], R0);
    R0 = fmax(hi[d] - X0[d], R0);
  }
  R0 *= 1.00001;
  root[0] = X0[0];
  root[1] = X0[1];
  root[2] = X0[2];
  root[3] = R0;
}

// Forest: one workgroup per problem (bodies off[p] .. off[p+1] - 1), the
// root of k_bounds_final per problem (min / max are exact, so any reduction
// order gives the same root).  Empty problems get a unit root.
constexpr int kFBThreads = 1024;  // one workgroup per problem: up to ~10^5 bodies each
__global__ void __launch_bounds__(kFBThreads) k_forest_bounds(const exafmm_body *__restrict__ b,
                                                       const int64_t *__restrict__ off,
                                                       double4 *__restrict__ roots, int *__restrict__ flags) {
  __shared__ double s[6][kFBThreads];
  const int p = blockIdx.x, t = threadIdx.x;
  const int64_t i0 = off[p], i1 = off[p + 1];
  if (i1 <= i0) {
    if (t == 0) roots[p] = make_double4(0.0, 0.0, 0.0, 1.0);
    return;
  }
  double lo[3] = {b[i0].X[0], b[i0].X[1], b[i0].X[2]};
  double hi[3] = {lo[0], lo[1], lo[2]};
  int bad = 0;
  for (int64_t i = i0 + t; i < i1; i += kFBThreads) {
    for (int d = 0; d < 3; d++) {
      const double x = b[i].X[d];
      if (!isfinite(x)) bad = 1;
      lo[d] = fmin(x, lo[d]);
      hi[d] = fmax(x, hi[d]);
    }
    if (!isfinite(b[i].q)) bad = 1;
  }
  if (bad) atomicOr(flags, 1);
  for (int d = 0; d < 3; d++) {
    s[d][t] = lo[d];
    s[3 + d][t] = hi[d];
  }
  __syncthreads();
  for (int o = kFBThreads / 2; o > 0; o >>= 1) {
    if (t < o)
      for (int d = 0; d < 3; d++) {
        s[d][t] = fmin(s[d][t], s[d][t + o]);
        s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + o]);
      }
    __syncthreads();
  }
  if (t != 0) return;
  double X0[3], R0 = 0;
  for (int d = 0; d < 3; d++) X0[d] = (s[3 + d][0] + s[d][0]) / 2;
  for (int d = 0; d < 3; d++) {
    R0 = fmax(X0[d] - s[d][0], R0);
    R0 = fmax(s[3 + d][0] - X0[d], R0);
  }
  R0 *= 1.00001;
  roots[p] = make_double4(X0[0], X0[1], X0[2], R0);
}

// ---------------------------------------------------------------- keys
// Octant path of each body through the reference's centre sequence, and
// the population of its level-7 prefix (dense, 8^7 bins).  With AGG the
// workgroup first counts its 256 bodies' prefixes in an LDS table and adds
// one atomic per distinct prefix: clustered inputs (a Plummer core puts
// most bodies in a few hundred prefixes) would otherwise serialise on the
// same global counters; spread-out inputs take the direct atomics.
constexpr int kKeysThreads = 256;
constexpr int kKeysSlots = 2 * kKeysThreads;  // load factor <= 1/2: every probe ends
template <bool AGG>
__global__ void __launch_bounds__(kKeysThreads) k_keys(const exafmm_body *__restrict__ b, int64_t n,
                                                       const double *__restrict__ root,
                                                       unsigned long long *__restrict__ key, int *__restrict__ val,
                                                       int *__restrict__ hist7, const double4 *__restrict__ froots,
                                                       const int64_t *__restrict__ off, int nprob) {
  __shared__ unsigned tbin[AGG ? kKeysSlots : 1];  // prefix + 1 (0: empty)
  __shared__ int tcnt[AGG ? kKeysSlots : 1];
  if (AGG) {
    for (int i = threadIdx.x; i < kKeysSlots; i += kKeysThreads) {
      tbin[i] = 0u;
      tcnt[i] = 0;
    }
    __syncthreads();
  }
  const int64_t i = int64_t(blockIdx.x) * kKeysThreads + threadIdx.x;
  if (i < n) {
    const double x = b[i].X[0], y = b[i].X[1], z = b[i].X[2];
    double R0 = root[3], cx = root[0], cy = root[1], cz = root[2];
    unsigned long long k = 0;
    int l0 = 0;
    if (froots) {
      // forest: the problem's index fills the digits of levels 0 .. kForestLevel - 1,
      // its own root's octant path the rest
      int lo = 0, hi = nprob - 1;  // last problem with off[p] <= i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i)
          lo = mid;
        else
          hi = mid - 1;
      }
      const double4 Rt = froots[lo];
      cx = Rt.x;
      cy = Rt.y;
      cz = Rt.z;
      R0 = Rt.w;
      k = (unsigned long long)lo;
      l0 = kForestLevel;
    }
    for (int l = l0; l < kMaxLevel; l++) {
      const int bx = x > cx, by = y > cy, bz = z > cz;
      const int oct = bx + (by << 1) + (bz << 2);
      k = (k << 3) | unsigned(oct);
      const double rr = R0 / (1 << (l - l0 + 1));
      cx += bx ? rr : -rr;
      cy += by ? rr : -rr;
      cz += bz ? rr : -rr;
    }
    key[i] = k;
    val[i] = int(i);
    const unsigned bin = unsigned(k >> (3 * (kMaxLevel - kHistLevels)));
    if (!AGG) {
      atomicAdd(&hist7[bin], 1);
    } else {
      unsigned h = ((bin + 1u) * 2654435761u) & unsigned(kKeysSlots - 1);
      for (;;) {
        const unsigned prev = atomicCAS(&tbin[h], 0u, bin + 1u);
        if (prev == 0u || prev == bin + 1u) {
          atomicAdd(&tcnt[h], 1);
          break;
        }
        h = (h + 1u) & unsigned(kKeysSlots - 1);
      }
    }
  }
  if (AGG) {
    __syncthreads();
    for (int j = threadIdx.x; j < kKeysSlots; j += kKeysThreads)
      if (tcnt[j]) atomicAdd(&hist7[tbin[j] - 1u], tcnt[j]);
  }
}

// Levels 6..3 of the prefix histogram from level 7 (prefixes are in Morton
// order, so prefix p of level l has children 8p..8p+7 at level l+1).  Block
// b sums the 4096 level-7 bins under level-3 prefix b: thread t loads bins
// 4t..4t+3 in one 16-byte load, lane shuffles give levels 6 and 5, LDS
// levels 4 and 3.  The block's largest population per level goes to
// bmax[8b + l] (l = 3..7).
static_assert(kHistLevels == 7, "k_hist_pyramid is shaped for 8^7 level-7 bins");
__global__ void __launch_bounds__(1024) k_hist_pyramid(int *__restrict__ hist, int *__restrict__ bmax) {
  __shared__ int w4[16], wm[16][3], s4[8];
  const int t = threadIdx.x, b = blockIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const int4 v = reinterpret_cast<const int4 *>(hist + hist_off(7))[b * 1024 + t];
  int m7 = max(max(v.x, v.y), max(v.z, v.w));
  int s6 = v.x + v.y + v.z + v.w;
  s6 += __shfl_xor(s6, 1);  // 8 level-7 bins: 2 threads
  int s5 = s6 + __shfl_xor(s6, 2);
  s5 += __shfl_xor(s5, 4);
  s5 += __shfl_xor(s5, 8);  // 64 bins: 16 threads
  int h = s5 + __shfl_xor(s5, 16);
  h += __shfl_xor(h, 32);  // the wave's half of a level-4 prefix
  if ((t & 1) == 0) hist[hist_off(6) + b * 512 + (t >> 1)] = s6;
  if ((t & 15) == 0) hist[hist_off(5) + b * 64 + (t >> 4)] = s5;
  int m6 = s6, m5 = s5;
  for (int o = 1; o < kWave; o <<= 1) {
    m7 = max(m7, __shfl_xor(m7, o));
    m6 = max(m6, __shfl_xor(m6, o));
    m5 = max(m5, __shfl_xor(m5, o));
  }
  if (lane == 0) {
    w4[w] = h;
    wm[w][0] = m7;
    wm[w][1] = m6;
    wm[w][2] = m5;
  }
  __syncthreads();
  if (t < 8) {
    s4[t] = w4[2 * t] + w4[2 * t + 1];
    hist[hist_off(4) + b * 8 + t] = s4[t];
  }
  __syncthreads();
  if (t == 0) {
    int a7 = 0, a6 = 0, a5 = 0, a4 = 0, s3 = 0;
    for (int i = 0; i < 16; i++) {
      a7 = max(a7, wm[i][0]);
      a6 = max(a6, wm[i][1]);
      a5 = max(a5, wm[i][2]);
    }
    for (int i = 0; i < 8; i++) {
      a4 = max(a4, s4[i]);
      s3 += s4[i];
    }
    hist[hist_off(3) + b] = s3;
    bmax[8 * b + 3] = s3;
    bmax[8 * b + 4] = a4;
    bmax[8 * b + 5] = a5;
    bmax[8 * b + 6] = a6;
    bmax[8 * b + 7] = a7;
  }
}

// Levels 2 and 1 from level 3, and the largest population of every level
// (lmax[l], l = 1..7) from the blocks' maxima: one workgroup of 512.
__global__ void __launch_bounds__(512) k_hist_top(int *__restrict__ hist, const int *__restrict__ bmax,
                                                  int *__restrict__ lmax) {
  __shared__ int red[512], s2[64];
  const int t = threadIdx.x;
  for (int l = 3; l <= 7; l++) {
    red[t] = bmax[8 * t + l];
    __syncthreads();
    for (int o = 256; o > 0; o >>= 1) {
      if (t < o) red[t] = max(red[t], red[t + o]);
      __syncthreads();
    }
    if (t == 0) lmax[l] = red[0];
    __syncthreads();
  }
  if (t < 64) {
    int s = 0;
    for (int c = 0; c < 8; c++) s += hist[hist_off(3) + t * 8 + c];
    s2[t] = s;
    hist[hist_off(2) + t] = s;
  }
  __syncthreads();
  if (t == 0) {
    int a2 = 0, a1 = 0;
    for (int i = 0; i < 64; i++) a2 = max(a2, s2[i]);
    for (int o = 0; o < 8; o++) {
      int s = 0;
      for (int c = 0; c < 8; c++) s += s2[o * 8 + c];
      hist[hist_off(1) + o] = s;
      a1 = max(a1, s);
    }
    lmax[2] = a2;
    lmax[1] = a1;
  }
}

// ---------------------------------------------------------------- levels
// Node t of level l: if it holds more than ncrit bodies, find the body
// sub-ranges of its 8 octants (keys are sorted by the level-(l+1) digit
// inside the node's range) by binary search.
// Child boundaries of a level-l cell holding bodies [beg, beg + cnt) of the
// sorted keys: b[o] .. b[o + 1] are octant o's bodies.  Returns the number of
// non-empty children (0: the cell is a leaf, b untouched).
__device__ inline int split_bounds(int l, int ncrit, int beg, int cnt, const unsigned long long *__restrict__ key,
                                   const int *__restrict__ hist, int b[9], int lforce) {
  // levels above lforce (a forest's problem-index levels) always split
  if ((cnt <= ncrit && l >= lforce) || cnt == 0 || l >= kMaxLevel) return 0;
  const int sh = 3 * (kMaxLevel - 1 - l);
  b[0] = beg;
  b[8] = beg + cnt;
  if (l + 1 <= kHistLevels) {
    // the children's populations are the level-(l+1) prefix histogram
    // entries of this node's prefix (all its bodies share it)
    const unsigned long long pre = key[beg] >> (sh + 3);
    const int *h = hist + hist_off(l + 1) + int(pre) * 8;
    for (int o = 1; o < 8; o++) b[o] = b[o - 1] + h[o - 1];
  } else {
    for (int o = 1; o < 8; o++) {
      int lo = b[o - 1], hi = beg + cnt;  // first index with digit >= o
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (int((key[mid] >> sh) & 7ull) < o)
          lo = mid + 1;
        else
          hi = mid;
      }
      b[o] = lo;
    }
  }
  int k = 0;
  for (int o = 0; o < 8; o++) k += b[o + 1] > b[o];
  return k;
}

__global__ void __launch_bounds__(256) k_split(int l, int ncrit, int lforce, const int *__restrict__ lcount,
                                               const int *__restrict__ lstart, const int *__restrict__ cbeg,
                                               const int *__restrict__ ccnt,
                                               const unsigned long long *__restrict__ key,
                                               const int *__restrict__ hist, int *__restrict__ nch,
                                               int *__restrict__ bnd) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= lcount[l]) return;
  const int c = lstart[l] + t;
  int b[9];
  const int k = split_bounds(l, ncrit, cbeg[c], ccnt[c], key, hist, b, lforce);
  nch[t] = k;
  if (k == 0) return;
  for (int o = 0; o < 9; o++) bnd[t * 9 + o] = b[o];
}

// Writes the non-empty children of level-l cell c (boundaries b) to cells
// first, first + 1, ... in octant order.
// r = R_parent / 2 is R0 / 2^(l+1) exactly.  Forest: the cells of level
// kForestLevel are the problems' roots (froots, by the problem index in the
// top key digits).
__device__ inline void emit_children(int l, int c, int first, const int b[9], const double4 *__restrict__ froots,
                                     const unsigned long long *__restrict__ key, double4 *__restrict__ cx,
                                     int *__restrict__ cbeg, int *__restrict__ ccnt, int *__restrict__ cchild,
                                     int *__restrict__ cparent, int *__restrict__ clevel) {
  cchild[c] = first;
  const double4 P = cx[c];
  const double r = P.w * 0.5;
  int j = 0;
  for (int o = 0; o < 8; o++) {
    const int b0 = b[o], b1 = b[o + 1];
    if (b1 <= b0) continue;
    const int ch = first + j++;
    double4 X;
    X.x = P.x + ((o & 1) ? r : -r);
    X.y = P.y + ((o & 2) ? r : -r);
    X.z = P.z + ((o & 4) ? r : -r);
    X.w = r;
    if (froots && l + 1 == kForestLevel) X = froots[int(key[b0] >> (3 * (kMaxLevel - kForestLevel)))];
    cx[ch] = X;
    cbeg[ch] = b0;
    ccnt[ch] = b1 - b0;
    cparent[ch] = c;
    clevel[ch] = l + 1;
  }
}

__global__ void k_level_next(int l, int *__restrict__ lcount, int *__restrict__ lstart, const int *__restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    lcount[l + 1] = *total;
    lstart[l + 1] = lstart[l] + lcount[l];
  }
}

// Children of level-l node t, in octant order (3d/build_tree.h:115-132):
// centre = parent centre + r*(+/-1), r = R0/2^(l+1); R = R0/2^(l+1).
__global__ void __launch_bounds__(256) k_emit(int l, const int *__restrict__ lcount, const int *__restrict__ lstart,
                                              const int *__restrict__ nch, const int *__restrict__ off,
                                              const int *__restrict__ bnd, const double4 *__restrict__ froots,
                                              const unsigned long long *__restrict__ key,
                                              double4 *__restrict__ cx, int *__restrict__ cbeg,
                                              int *__restrict__ ccnt, int *__restrict__ cchild,
                                              int *__restrict__ cnch, int *__restrict__ cparent,
                                              int *__restrict__ clevel, int *__restrict__ nsplit) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const bool in = t < lcount[l];
  const int k = in ? nch[t] : 0;
  // cells that split, counted per wave (leaves = cells - split cells)
  const unsigned long long sp = __ballot(k > 0);
  if ((threadIdx.x & (kWave - 1)) == 0 && sp) atomicAdd(nsplit, int(__popcll(sp)));
  if (!in) return;
  const int c = lstart[l] + t;
  cnch[c] = k;
  if (k == 0) {
    cchild[c] = -1;
    return;
  }
  int b[9];
  for (int o = 0; o < 9; o++) b[o] = bnd[t * 9 + o];
  emit_children(l, c, lstart[l + 1] + off[t], b, froots, key, cx, cbeg, ccnt, cchild, cparent, clevel);
}

// The top levels, 0 .. lend - 1, split by one workgroup: per level the
// k_split / exclusive scan / k_level_next / k_emit sequence above, with a
// workgroup scan and barriers in place of launches.  Level l holds at most
// 8^l cells, so levels up to kTopLevels - 1 fit kTopPer cells per thread;
// thread t owns cells kTopPer t .. kTopPer t + kTopPer - 1 of its level.
// Level 4 (4096 cells) stays on the multi-workgroup kernels: one workgroup
// walking four cells per thread through their dependent loads (counts, key,
// histogram, centre) took longer than the four launches it saves.
constexpr int kTopThreads = 1024, kTopPer = 1, kTopLevels = 4;
static_assert(1 << (3 * (kTopLevels - 1)) <= kTopThreads * kTopPer, "top levels exceed the workgroup");
__global__ void __launch_bounds__(kTopThreads)
    k_top_levels(int lend, int ncrit, int lforce, int *__restrict__ lcount, int *__restrict__ lstart,
                 const unsigned long long *__restrict__ key, const int *__restrict__ hist,
                 const double4 *__restrict__ froots, double4 *__restrict__ cx, int *__restrict__ cbeg,
                 int *__restrict__ ccnt, int *__restrict__ cchild, int *__restrict__ cnch,
                 int *__restrict__ cparent, int *__restrict__ clevel, int *__restrict__ nsplit) {
  __shared__ int wsum[kTopThreads / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  int cnt = lcount[0], start = lstart[0];
  for (int l = 0; l < lend; l++) {
    int b[kTopPer][9], k[kTopPer];
    int mine = 0, nspl = 0;
#pragma unroll
    for (int j = 0; j < kTopPer; j++) {
      const int t = tid * kTopPer + j;
      k[j] = 0;
      if (t < cnt) {
        const int c = start + t;
        k[j] = split_bounds(l, ncrit, cbeg[c], ccnt[c], key, hist, b[j], lforce);
      }
      mine += k[j];
      nspl += k[j] > 0;
    }
    // exclusive scan of the children counts over the workgroup
    int incl = mine;
    for (int d = 1; d < kWave; d <<= 1) {
      const int v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    for (int d = kWave / 2; d > 0; d >>= 1) nspl += __shfl_xor(nspl, d);
    if (lane == 0 && nspl) atomicAdd(nsplit, nspl);  // cells that split
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int off = incl - mine, total = 0;
    for (int i = 0; i < kTopThreads / kWave; i++) {
      off += i < w ? wsum[i] : 0;
      total += wsum[i];
    }
    const int next = start + cnt;
#pragma unroll
    for (int j = 0; j < kTopPer; j++) {
      const int t = tid * kTopPer + j;
      if (t < cnt) {
        const int c = start + t;
        cnch[c] = k[j];
        if (k[j] == 0)
          cchild[c] = -1;
        else
          emit_children(l, c, next + off, b[j], froots, key, cx, cbeg, ccnt, cchild, cparent, clevel);
        off += k[j];
      }
    }
    if (tid == 0) {
      lcount[l + 1] = total;
      lstart[l + 1] = next;
    }
    __syncthreads();  // children written before the next level reads them
    start = next;
    cnt = total;
  }
}

__global__ void k_root(int64_t n, const double *__restrict__ root, double4 *__restrict__ cx, int *__restrict__ cbeg,
                       int *__restrict__ ccnt, int *__restrict__ cparent, int *__restrict__ clevel,
                       int *__restrict__ lcount, int *__restrict__ lstart) {
  cx[0] = make_double4(root[0], root[1], root[2], root[3]);
  cbeg[0] = 0;
  ccnt[0] = int(n);
  cparent[0] = -1;
  clevel[0] = 0;
  lcount[0] = 1;
  lstart[0] = 0;
}

// Cells of the deepest sorted level: leaves; more than ncrit bodies there
// means the tree needs more than 21 levels.
__global__ void __launch_bounds__(256) k_last_level(int l, int ncrit, const int *__restrict__ lcount,
                                                    const int *__restrict__ lstart, const int *__restrict__ ccnt,
                                                    int *__restrict__ cchild, int *__restrict__ cnch,
                                                    int *__restrict__ flags) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t == 0) flags[19] = lstart[l] + lcount[l];  // cells in the tree
  if (t >= lcount[l]) return;
  const int c = lstart[l] + t;
  cnch[c] = 0;
  cchild[c] = -1;
  if (ccnt[c] > ncrit) atomicOr(flags, 2);
}

// Leaf flags and the compacted leaf list over the *ncells cells (a device
// count: grid-stride loops over a grid sized from a host bound).
__global__ void __launch_bounds__(256) k_leaf_flag(const int *__restrict__ ncells, const int *__restrict__ cnch,
                                                   int *__restrict__ f) {
  const int nc = *ncells;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < nc; c += gridDim.x * 256) f[c] = cnch[c] == 0;
}
__global__ void __launch_bounds__(256) k_leaf_compact(const int *__restrict__ ncells, const int *__restrict__ cnch,
                                                      const int *__restrict__ off, int *__restrict__ leaves) {
  const int nc = *ncells;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < nc; c += gridDim.x * 256)
    if (cnch[c] == 0) leaves[off[c]] = c;
}

// One workgroup per leaf: restore input order inside the leaf (rank of
// each original index among the leaf's), then gather X, q (FP64) and the
// leaf-relative FP32 copy used by P2P.
__global__ void __launch_bounds__(64) k_leaf_gather(const int *__restrict__ leaves, const int *__restrict__ nleaves,
                                                    const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                                    const double4 *__restrict__ cx, const int *__restrict__ val,
                                                    const exafmm_body *__restrict__ b, int *__restrict__ perm,
                                                    double4 *__restrict__ xs, float4 *__restrict__ xq,
                                                    int *__restrict__ body_leaf, int ncrit) {
  extern __shared__ int sidx[];  // ncrit ints
  const int nl = *nleaves;
  for (int li = blockIdx.x; li < nl; li += gridDim.x) {
    const int c = leaves[li];
    const int beg = cbeg[c], cnt = ccnt[c];
    // a leaf above ncrit: a tree the host rejects or builds again (flag 2)
    if (cnt > ncrit) continue;
    for (int i = threadIdx.x; i < cnt; i += 64) sidx[i] = val[beg + i];
    __syncthreads();
    const double4 C = cx[c];
    for (int i = threadIdx.x; i < cnt; i += 64) {
      const int me = sidx[i];
      int r = 0;
      for (int j = 0; j < cnt; j++) r += sidx[j] < me;
      const int pos = beg + r;
      perm[pos] = me;
      const double x = b[me].X[0], y = b[me].X[1], z = b[me].X[2], q = b[me].q;
      xs[pos] = make_double4(x, y, z, q);
      xq[pos] = make_float4(float(x - C.x), float(y - C.y), float(z - C.z), float(q));
      body_leaf[pos] = c;
    }
    __syncthreads();
  }
}

// Reference permutes X and q of the caller's array in place (p, F stay).
// Bodies back to the caller in tree order -- unless the build flagged a
// non-finite input (bit 1) or a leaf above ncrit (bit 2): the caller's
// array is then left as it was.
__global__ void __launch_bounds__(256) k_writeback(int64_t n, const double4 *__restrict__ xs,
                                                   exafmm_body *__restrict__ b, const int *__restrict__ flags) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n || (*flags & 3)) return;
  const double4 v = xs[i];
  b[i].X[0] = v.x;
  b[i].X[1] = v.y;
  b[i].X[2] = v.z;
  b[i].q = v.w;
}

// Sort depth from the largest prefix population of levels 1..7 (lmax[l]).
static int sort_depth(const int *lmax, int64_t n, int ncrit, bool forest, int lforce) {
  if (n <= ncrit && !forest) return 0;
  for (int l = std::max(1, lforce); l <= kHistLevels; l++)
    if (lmax[l] <= ncrit) return l;
  return kMaxLevel;
}

void writeback_bodies(Ctx &c, hipStream_t s) {
  const int64_t n = c.t.n;
  hipLaunchKernelGGL(k_writeback, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, c.t.xs.p, c.bodies, c.t.misc.p);
  XF_LAUNCH();
  c.writeback_pending = false;
}

void tree_build(Ctx &c) {
  Tree &t = c.t;
  const int64_t n = t.n;
  hipStream_t s = c.s0;
  const int ncrit = c.cfg.ncrit;
  if (n > (int64_t(1) << 30)) fail(EXAFMM_ERR_ARG, "N above 2^30 is not supported");

  // misc layout (ints): [0] flags, [1..8] level max, [16] level total,
  // [17] leaves, [18] cells that split
  t.misc.ensure(64);
  t.bounds_part.ensure(256 * 6 + 8);
  double *root = t.bounds_part.p + 256 * 6;
  XF_HIP(hipMemsetAsync(t.misc.p, 0, 64 * sizeof(int), s));

  // forest: per-problem roots; the root of levels 0 .. kForestLevel - 1
  // (problem-index cells, never paired by the traversal) is a unit cube
  const bool forest = c.forest_n > 0;
  const double4 *froots = forest ? c.forest_roots.p : nullptr;
  const int lforce = forest ? kForestLevel : 0;
  t.lroot = lforce;
  if (forest) {
    hipLaunchKernelGGL(k_forest_bounds, dim3(c.forest_n), dim3(kFBThreads), 0, s, c.bodies, c.forest_off.p,
                       c.forest_roots.p, t.misc.p);
    XF_LAUNCH();
    const double unit[4] = {0.0, 0.0, 0.0, 1.0};
    XF_HIP(hipMemcpyAsync(root, unit, sizeof(unit), hipMemcpyHostToDevice, s));
  } else {
    hipLaunchKernelGGL(k_bounds_part, dim3(256), dim3(256), 0, s, c.bodies, n, t.bounds_part.p, t.misc.p);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, t.bounds_part.p, 256, root);
    XF_LAUNCH();
  }

  t.key.ensure(n);
  t.key2.ensure(n);
  t.val.ensure(n);
  t.val2.ensure(n);
  t.hist.ensure(kHistTotal);
  XF_HIP(hipMemsetAsync(t.hist.p + hist_off(kHistLevels), 0, sizeof(int) * (1 << (3 * kHistLevels)), s));
  // crowded level-7 prefixes in the last tree: count them per workgroup first
  hipLaunchKernelGGL(t.hist7_max > 64 ? k_keys<true> : k_keys<false>, dim3(ceil_div(n, kKeysThreads)),
                     dim3(kKeysThreads), 0, s, c.bodies, n, root, t.key.p, t.val.p,
                     t.hist.p + hist_off(kHistLevels), froots, forest ? c.forest_off.p : nullptr, c.forest_n);
  XF_LAUNCH();
  // one block per level-3 prefix (4096 level-7 bins each)
  t.hist_bmax.ensure(8 << (3 * 3));
  hipLaunchKernelGGL(k_hist_pyramid, dim3(1 << (3 * 3)), dim3(1024), 0, s, t.hist.p, t.hist_bmax.p);
  XF_LAUNCH();
  hipLaunchKernelGGL(k_hist_top, dim3(1), dim3(1 << (3 * 3)), 0, s, t.hist.p, t.hist_bmax.p, t.misc.p);
  XF_LAUNCH();

  // Sort depth Ds: the first level whose most populous prefix holds at
  // most ncrit bodies (kMaxLevel if none of levels 1..7 does).  It needs the
  // histogram on the host; instead of waiting for it here the build reuses
  // the previous call's depth and checks it at the level-size read below.
  // A depth at or above the true one builds the same tree (a cell splits
  // only above ncrit bodies, and each leaf keeps its bodies in input order),
  // so only a smaller one makes the build start over.
  const bool speculate = t.ds_hint >= 0;
  int Ds = t.ds_hint;
  c.pin_i.ensure(kPinInts);
  c.pin_d.ensure(4);
  int *hm = c.pin_i.p + kPinLevelMax;
  if (!speculate) {
    XF_HIP(hipMemcpyAsync(hm, t.misc.p, sizeof(int) * 9, hipMemcpyDeviceToHost, s));
    stream_wait(s);
    if (hm[0] & 1) fail(EXAFMM_ERR_ARG, "non-finite body coordinate or charge");
    Ds = sort_depth(hm, n, ncrit, forest, lforce);
  }

  // stable sort by the top Ds octant digits
  unsigned long long *k1 = t.key.p, *k2 = t.key2.p;
  int *v1 = t.val.p, *v2 = t.val2.p;
  radix_sort_pairs(k1, v1, k2, v2, int(n), 3 * (kMaxLevel - Ds), 3 * kMaxLevel, t.rs_hist, s);

  // level-by-level split
  std::vector<int64_t> ub(Ds + 2);
  int64_t ubtot = 0;
  // nodes of level l+1 are children of level-l nodes holding > ncrit bodies
  const int64_t ub_split = 8 * (n / (ncrit + 1)) + 8;
  for (int l = 0; l <= Ds; l++) {
    ub[l] = std::min<int64_t>(l >= 10 ? n : (int64_t(1) << (3 * l)), n);
    if (l > lforce) ub[l] = std::min(ub[l], ub_split);  // (a forest's top levels split unconditionally)
    ubtot += ub[l];
  }
  t.cx.ensure(ubtot);
  t.cbeg.ensure(ubtot);
  t.ccnt.ensure(ubtot);
  t.cchild.ensure(ubtot);
  t.cnch.ensure(ubtot);
  t.cparent.ensure(ubtot);
  t.clevel.ensure(ubtot);
  t.lcount.ensure(kMaxLevel + 2);
  t.lstart_d.ensure(kMaxLevel + 2);
  XF_HIP(hipMemsetAsync(t.lcount.p, 0, sizeof(int) * (kMaxLevel + 2), s));
  int64_t ubmax = 1;
  for (int l = 0; l <= Ds; l++) ubmax = std::max(ubmax, ub[l]);
  // (the leaf flags and offsets below span all cells: ubtot)
  t.tmp_nch.ensure(ubtot);
  t.tmp_off.ensure(ubtot);
  t.tmp_bnd.ensure(ubmax * 9);
  t.leaves.ensure(ubtot);
  t.perm.ensure(n);
  t.xs.ensure(n);
  t.xq.ensure(n);
  t.body_leaf.ensure(n);
  hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, s, n, root, t.cx.p, t.cbeg.p, t.ccnt.p, t.cparent.p, t.clevel.p,
                     t.lcount.p, t.lstart_d.p);
  XF_LAUNCH();
  const int lfuse = std::min(Ds, kTopLevels);
  if (lfuse > 0) {
    hipLaunchKernelGGL(k_top_levels, dim3(1), dim3(kTopThreads), 0, s, lfuse, ncrit, lforce, t.lcount.p,
                       t.lstart_d.p, k1, t.hist.p, froots, t.cx.p, t.cbeg.p, t.ccnt.p, t.cchild.p, t.cnch.p,
                       t.cparent.p, t.clevel.p, t.misc.p + 18);
    XF_LAUNCH();
  }
  for (int l = lfuse; l < Ds; l++) {
    const int g = ceil_div(ub[l], 256);
    hipLaunchKernelGGL(k_split, dim3(g), dim3(256), 0, s, l, ncrit, lforce, t.lcount.p, t.lstart_d.p, t.cbeg.p,
                       t.ccnt.p, k1, t.hist.p, t.tmp_nch.p, t.tmp_bnd.p);
    XF_LAUNCH();
    scan_exclusive_dev(t.tmp_nch.p, t.tmp_off.p, t.lcount.p + l, int(ub[l]), t.misc.p + 16, t.scan_tmp, s);
    hipLaunchKernelGGL(k_level_next, dim3(1), dim3(1), 0, s, l, t.lcount.p, t.lstart_d.p, t.misc.p + 16);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_emit, dim3(g), dim3(256), 0, s, l, t.lcount.p, t.lstart_d.p, t.tmp_nch.p, t.tmp_off.p,
                       t.tmp_bnd.p, froots, k1, t.cx.p, t.cbeg.p, t.ccnt.p, t.cchild.p, t.cnch.p, t.cparent.p,
                       t.clevel.p, t.misc.p + 18);
    XF_LAUNCH();
  }
  hipLaunchKernelGGL(k_last_level, dim3(ceil_div(ub[Ds], 256)), dim3(256), 0, s, Ds, ncrit, t.lcount.p,
                     t.lstart_d.p, t.ccnt.p, t.cchild.p, t.cnch.p, t.misc.p);
  XF_LAUNCH();

  // Leaves, the gather into tree order and the write-back run before the
  // level sizes are read, so that read overlaps them: the cell count comes
  // from the device (misc[19]), the leaf count too (misc[17]), and grids are
  // sized from the bound ubtot.  A tree the checks below reject or rebuild
  // (flag bit 1 or 2) leaves the caller's bodies untouched.
  const int gcells = int(std::min<int64_t>(ceil_div(ubtot, 256), 2048));
  hipLaunchKernelGGL(k_leaf_flag, dim3(gcells), dim3(256), 0, s, t.misc.p + 19, t.cnch.p, t.tmp_nch.p);
  XF_LAUNCH();
  scan_exclusive_dev(t.tmp_nch.p, t.tmp_off.p, t.misc.p + 19, int(std::min<int64_t>(ubtot, INT32_MAX)), t.misc.p + 17,
                     t.scan_tmp, s);
  hipLaunchKernelGGL(k_leaf_compact, dim3(gcells), dim3(256), 0, s, t.misc.p + 19, t.cnch.p, t.tmp_off.p,
                     t.leaves.p);
  XF_LAUNCH();
  hipLaunchKernelGGL(k_leaf_gather, dim3(int(std::min<int64_t>(ubtot, 16384))), dim3(64), sizeof(int) * ncrit, s,
                     t.leaves.p, t.misc.p + 17, t.cbeg.p, t.ccnt.p, t.cx.p, v1, c.bodies, t.perm.p, t.xs.p, t.xq.p,
                     t.body_leaf.p, ncrit);
  XF_LAUNCH();
  if (c.defer_writeback)
    c.writeback_pending = true;
  else
    writeback_bodies(c, s);

  // level sizes, flags, level maxima, root
  const int *lc = c.pin_i.p + kPinLevels, *fl = c.pin_i.p + kPinMisc;
  const double *rootv = c.pin_d.p;
  XF_HIP(hipMemcpyAsync(c.pin_i.p + kPinLevels, t.lcount.p, sizeof(int) * (kMaxLevel + 2), hipMemcpyDeviceToHost,
                        s));
  XF_HIP(hipMemcpyAsync(c.pin_i.p + kPinMisc, t.misc.p, sizeof(int) * 19, hipMemcpyDeviceToHost, s));
  XF_HIP(hipMemcpyAsync(c.pin_d.p, root, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
  stream_wait(s);
  if (fl[0] & 1) fail(EXAFMM_ERR_ARG, "non-finite body coordinate or charge");
  // The histogram fixes depths up to kHistLevels; a deeper tree was sorted
  // at the last tree's depth + 1, which is enough unless a cell of that
  // level still holds more than ncrit bodies (flag 2, below kMaxLevel).
  const int need = sort_depth(fl, n, ncrit, forest, lforce);
  const bool too_shallow = need <= kHistLevels ? need > Ds : (Ds < kMaxLevel && (fl[0] & 2));
  if (too_shallow) {  // build again at the depth the histogram gives
    t.ds_hint = need;
    tree_build(c);
    return;
  }
  for (int d = 0; d < 3; d++) t.X0[d] = rootv[d];
  t.R0 = rootv[3];
  if (fl[0] & 2) fail(EXAFMM_ERR_DEPTH, "more than ncrit bodies share a 21-level octant path");
  t.lstart.assign(1, 0);
  int depth = 0;
  for (int l = 0; l <= Ds; l++) {
    if (lc[l] == 0) break;
    depth = l;
    t.lstart.push_back(t.lstart.back() + lc[l]);
  }
  t.depth = depth;
  t.ds_hint = need <= kHistLevels ? need : std::min(kMaxLevel, depth + 1);
  t.hist7_max = fl[kHistLevels];
  t.ncells = t.lstart.back();
  const int nc = int(t.ncells);
  // every cell either split or is a leaf (the scan above wrote the same
  // count to misc[17] for the kernels)
  t.nleaves = nc - fl[18];
  t.built = true;
}

// Reference numbering: the reference appends a parent's children block
// when it visits the parent, in DFS preorder (3d/build_tree.h:115-132).
void tree_ref_order(Ctx &c, std::vector<int64_t> &bfs2ref, std::vector<int64_t> &ref2bfs) {
  Tree &t = c.t;
  const int64_t nc = t.ncells;
  std::vector<int> cchild(nc), cnch(nc);
  XF_HIP(hipMemcpy(cchild.data(), t.cchild.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(cnch.data(), t.cnch.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  bfs2ref.assign(nc, -1);
  ref2bfs.assign(nc, -1);
  bfs2ref[0] = 0;
  int64_t next = 1;
  std::vector<int> stack{0};
  while (!stack.empty()) {
    const int u = stack.back();
    stack.pop_back();
    if (cnch[u] == 0) continue;
    const int64_t first = next;
    next += cnch[u];
    for (int k = 0; k < cnch[u]; k++) bfs2ref[cchild[u] + k] = first + k;
    for (int k = cnch[u] - 1; k >= 0; k--) stack.push_back(cchild[u] + k);
  }
  for (int64_t i = 0; i < nc; i++) ref2bfs[bfs2ref[i]] = i;
}

void tree_export_cells(Ctx &c, exafmm_cell *out) {
  Tree &t = c.t;
  const int64_t nc = t.ncells;
  std::vector<int64_t> b2r, r2b;
  tree_ref_order(c, b2r, r2b);
  std::vector<double4> cx(nc);
  std::vector<int> cbeg(nc), ccnt(nc), cchild(nc), cnch(nc);
  XF_HIP(hipMemcpy(cx.data(), t.cx.p, sizeof(double4) * nc, hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(cbeg.data(), t.cbeg.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(ccnt.data(), t.ccnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(cchild.data(), t.cchild.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  XF_HIP(hipMemcpy(cnch.data(), t.cnch.p, sizeof(int) * nc, hipMemcpyDeviceToHost));
  for (int64_t b = 0; b < nc; b++) {
    exafmm_cell &o = out[b2r[b]];
    o.nchild = cnch[b];
    o.nbody = ccnt[b];
    o.child = cnch[b] ? b2r[cchild[b]] : -1;
    o.body = cbeg[b];
    o.X[0] = cx[b].x;
    o.X[1] = cx[b].y;
    o.X[2] = cx[b].z;
    o.R = cx[b].w;
  }
}

}  // namespace xf
