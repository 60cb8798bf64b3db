// scan.hip -- device-wide exclusive scan and stable LSD radix sort.
#include <algorithm>

#include "fmm_internal.h"

namespace xf {

// Device-wide exclusive scan of NA int arrays of the same length in two
// launches: each of G workgroups reduces a contiguous range (coalesced
// 1024-element tiles), then each rescans its range starting from the sum of
// the ranges before it.  The length may live on the device (n_dev); the
// grid is sized from a host upper bound and idle workgroups exit.  base
// (optional) is added to every output and receives the totals; cap/flag
// (optional) raise bit 1 of *flag when a total exceeds cap[k].
constexpr int kScanThreads = 1024;
constexpr int kScanMaxBlocks = 64;

template <int NA>
struct ScanArrays {
  const int *in[NA];
  int *out[NA];
};

__device__ inline int block_sum(int v, int *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  if (lane == 0) red[w] = v;
  __syncthreads();
  int s = 0;
  for (int k = 0; k < kScanThreads / 64; k++) s += red[k];
  __syncthreads();
  return s;
}

constexpr int kScanChunk = 8 * kScanThreads;  // elements per workgroup range granule (one tile)
__device__ inline void scan_range(int n, int G, int &b0, int &e) {
  int chunk = (n + G - 1) / G;
  chunk = (chunk + kScanChunk - 1) / kScanChunk * kScanChunk;
  b0 = blockIdx.x * chunk;
  e = min(b0 + chunk, n);
}

// Items per thread in one pass over a tile of kScanThreads * kScanItems
// elements: the tile is loaded with every load in flight (coalesced, item j
// of thread t at j * kScanThreads + t), transposed through LDS so that each
// thread owns kScanItems consecutive elements, scanned in registers, and
// written back the same way.  Padding one int per 32 keeps the transposed
// reads free of bank conflicts.
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
static_assert(kScanTile == kScanChunk, "a workgroup range is a whole number of tiles");
__device__ inline int scan_pad(int i) { return i + (i >> 5); }

template <int NA>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(ScanArrays<NA> a, int n_host, const int *n_dev,
                                                              int *__restrict__ sums, const int *base) {
  __shared__ int red[kScanThreads / 64];
  const int n = n_dev ? *n_dev : n_host;
  const int G = gridDim.x;
  int b0, e;
  scan_range(n, G, b0, e);
#pragma unroll
  for (int k = 0; k < NA; k++) {
    int v = 0;
    for (int t0 = b0; t0 < e; t0 += kScanTile) {
      int x[kScanItems];
#pragma unroll
      for (int j = 0; j < kScanItems; j++) {
        const int i = t0 + j * kScanThreads + threadIdx.x;
        x[j] = i < e ? a.in[k][i] : 0;
      }
#pragma unroll
      for (int j = 0; j < kScanItems; j++) v += x[j];
    }
    v = block_sum(v, red);
    if (threadIdx.x == 0) sums[blockIdx.x * NA + k] = v;
  }
  // keep the incoming running totals where the apply pass reads them
  if (blockIdx.x == 0 && threadIdx.x < NA) sums[kScanMaxBlocks * NA + threadIdx.x] = base ? base[threadIdx.x] : 0;
}

template <int NA>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(ScanArrays<NA> a, int n_host, const int *n_dev,
                                                             const int *__restrict__ sums, int *total, int *base,
                                                             const int64_t *cap, int *flag) {
  __shared__ int red[kScanThreads / 64];
  __shared__ int wtot[kScanThreads / 64];
  __shared__ int sm[kScanTile + kScanTile / 32];
  const int n = n_dev ? *n_dev : n_host;
  const int G = gridDim.x;
  int b0, e;
  scan_range(n, G, b0, e);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int k = 0; k < NA; k++) {
    // one workgroup (sums == nullptr): no reduce pass, the running total
    // comes straight from base
    int carry = sums ? sums[kScanMaxBlocks * NA + k] : (base ? base[k] : 0);
    if (sums) {
      int v = 0;
      for (int j = tid; j < blockIdx.x; j += kScanThreads) v += sums[j * NA + k];
      carry += block_sum(v, red);
    }
    for (int t0 = b0; t0 < e; t0 += kScanTile) {
      int x[kScanItems];
#pragma unroll
      for (int j = 0; j < kScanItems; j++) {
        const int i = t0 + j * kScanThreads + tid;
        x[j] = i < e ? a.in[k][i] : 0;
      }
#pragma unroll
      for (int j = 0; j < kScanItems; j++) sm[scan_pad(j * kScanThreads + tid)] = x[j];
      __syncthreads();
      // this thread's kScanItems consecutive elements
      int run = 0;
#pragma unroll
      for (int j = 0; j < kScanItems; j++) {
        x[j] = sm[scan_pad(tid * kScanItems + j)];
        run += x[j];
      }
      int incl = run;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(incl, d);
        if (lane >= d) incl += u;
      }
      if (lane == 63) wtot[w] = incl;
      __syncthreads();
      int before = carry, tile = 0;
      for (int q = 0; q < kScanThreads / 64; q++) {
        if (q < w) before += wtot[q];
        tile += wtot[q];
      }
      before += incl - run;
#pragma unroll
      for (int j = 0; j < kScanItems; j++) {
        sm[scan_pad(tid * kScanItems + j)] = before;
        before += x[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kScanItems; j++) {
        const int i = t0 + j * kScanThreads + tid;
        if (i < e) a.out[k][i] = sm[scan_pad(j * kScanThreads + tid)];
      }
      carry += tile;
      __syncthreads();
    }
    __syncthreads();  // every thread has read base[k] before it is overwritten
    if (blockIdx.x == G - 1 && tid == 0) {
      if (total) total[k] = carry;
      if (base) base[k] = carry;
      if (cap && flag && carry > cap[k]) atomicOr(flag, 1);
    }
  }
}

template <int NA>
static void scan_launch(const ScanArrays<NA> &a, int n_host, const int *n_dev, int n_bound, int *total, int *base,
                        const int64_t *cap, int *flag, Buf<int> &scratch, hipStream_t s) {
  const int G = std::max(1, std::min(kScanMaxBlocks, ceil_div(std::max(n_bound, 1), 8 * kScanThreads)));
  scratch.ensure(size_t(kScanMaxBlocks + 1) * NA);
  if (G > 1) {
    hipLaunchKernelGGL(k_scan_reduce<NA>, dim3(G), dim3(kScanThreads), 0, s, a, n_host, n_dev, scratch.p,
                       (const int *)base);
    XF_LAUNCH();
  }
  hipLaunchKernelGGL(k_scan_apply<NA>, dim3(G), dim3(kScanThreads), 0, s, a, n_host, n_dev,
                     G > 1 ? (const int *)scratch.p : nullptr, total, base, cap, flag);
  XF_LAUNCH();
}

void scan_exclusive(const int *in, int *out, int count, int *total, Buf<int> &scratch, hipStream_t s) {
  ScanArrays<1> a{{in}, {out}};
  scan_launch<1>(a, count, nullptr, count, total, nullptr, nullptr, nullptr, scratch, s);
}

void scan_exclusive_dev(const int *in, int *out, const int *count_dev, int count_bound, int *total,
                        Buf<int> &scratch, hipStream_t s) {
  ScanArrays<1> a{{in}, {out}};
  scan_launch<1>(a, 0, count_dev, count_bound, total, nullptr, nullptr, nullptr, scratch, s);
}

void scan_exclusive3(const int *in0, const int *in1, const int *in2, int *out0, int *out1, int *out2, int count,
                     int *running, const int64_t *cap, int *flag, Buf<int> &scratch, hipStream_t s) {
  ScanArrays<3> a{{in0, in1, in2}, {out0, out1, out2}};
  scan_launch<3>(a, count, nullptr, count, nullptr, running, cap, flag, scratch, s);
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 4;
constexpr int kRsTile = kRsThreads * kRsItems;

__global__ void __launch_bounds__(kRsThreads) k_rs_hist(const unsigned long long *__restrict__ key, int n, int shift,
                                                        int nbits, int *__restrict__ hist, int nblocks) {
  __shared__ int h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const unsigned mask = (1u << nbits) - 1u;
  const int base = blockIdx.x * kRsTile;
  for (int r = 0; r < kRsItems; r++) {
    const int i = base + r * kRsThreads + t;
    if (i < n) atomicAdd(&h[unsigned(key[i] >> shift) & mask], 1);
  }
  __syncthreads();
  hist[t * nblocks + blockIdx.x] = h[t];
}

// Stable scatter: rounds of 256 consecutive items; inside a round the rank
// of an item among equal digits comes from 8 wave ballots (lanes with the
// same digit), counts of earlier waves, and a running per-digit counter
// over earlier rounds, so the output keeps input order for equal digits.
__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const unsigned long long *__restrict__ key,
                                                           const int *__restrict__ val, int n, int shift, int nbits,
                                                           const int *__restrict__ off, int nblocks,
                                                           const int *__restrict__ dtot,
                                                           unsigned long long *__restrict__ key_out,
                                                           int *__restrict__ val_out) {
  static_assert(kRsThreads == 256, "one thread per digit");
  __shared__ int run[256];
  __shared__ int wcnt[kRsThreads / kWave][256];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1), w = t / kWave;
  // digit bases: exclusive scan of the digit totals
  run[t] = dtot[t];
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int v = t >= o ? run[t - o] : 0;
    __syncthreads();
    run[t] += v;
    __syncthreads();
  }
  const int dbase = run[t] - dtot[t];
  __syncthreads();
  run[t] = dbase + off[t * nblocks + blockIdx.x];
  const unsigned mask = (1u << nbits) - 1u;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int base = blockIdx.x * kRsTile;
  for (int r = 0; r < kRsItems; r++) {
    for (int k = t; k < (kRsThreads / kWave) * 256; k += kRsThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const int i = base + r * kRsThreads + t;
    const bool valid = i < n;
    unsigned long long k = 0;
    int v = 0;
    unsigned d = 0;
    if (valid) {
      k = key[i];
      v = val[i];
      d = unsigned(k >> shift) & mask;
    }
    unsigned long long eq = __ballot(valid);
    for (int b = 0; b < nbits; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      eq &= bit ? bb : ~bb;
    }
    const int rank_w = __popcll(eq & lt);
    if (valid && rank_w == 0) wcnt[w][d] = __popcll(eq);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rank_w;
      for (int ww = 0; ww < w; ww++) pos += wcnt[ww][d];
      key_out[pos] = k;
      val_out[pos] = v;
    }
    __syncthreads();
    {
      int s = 0;
      for (int ww = 0; ww < kRsThreads / kWave; ww++) s += wcnt[ww][t];
      run[t] += s;
    }
    __syncthreads();
  }
}

// Offsets for the scatter: one workgroup per digit scans that digit's
// per-block counts (exclusive) and records the digit total; each scatter
// workgroup adds the digit bases (scan of the totals) itself.
__global__ void __launch_bounds__(256) k_rs_scan_digit(const int *__restrict__ h, int *__restrict__ o, int nblocks,
                                                       int *__restrict__ dtot) {
  __shared__ int part[256];
  const int d = blockIdx.x, t = threadIdx.x;
  const int chunk = (nblocks + 255) / 256;
  const int b = t * chunk, e = min(b + chunk, nblocks);
  const int *hd = h + d * nblocks;
  int s = 0;
  for (int i = b; i < e; i++) s += hd[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  for (int i = b; i < e; i++) {
    o[d * nblocks + i] = run;
    run += hd[i];
  }
  if (t == 255) dtot[d] = part[255];
}

// Sorts (k, v) by key bits [lo_bit, hi_bit); the result ends in (k, v)
// (the pointers are swapped as the passes ping-pong).
void radix_sort_pairs(unsigned long long *&k, int *&v, unsigned long long *&k2, int *&v2, int n, int lo_bit,
                      int hi_bit, Buf<int> &hist, hipStream_t s) {
  if (n <= 1 || hi_bit <= lo_bit) return;
  const int nb = ceil_div(n, kRsTile);
  hist.ensure(size_t(256) * nb * 2 + 256);
  int *h = hist.p, *o = hist.p + 256 * nb, *dt = hist.p + 512 * nb;
  for (int sh = lo_bit; sh < hi_bit; sh += 8) {
    const int bits = min(8, hi_bit - sh);
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsThreads), 0, s, k, n, sh, bits, h, nb);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_rs_scan_digit, dim3(256), dim3(256), 0, s, h, o, nb, dt);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsThreads), 0, s, k, v, n, sh, bits, o, nb, dt, k2, v2);
    XF_LAUNCH();
    std::swap(k, k2);
    std::swap(v, v2);
  }
}

}  // namespace xf
