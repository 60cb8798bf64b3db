// scan.hip -- device-wide exclusive scan and stable LSD radix sort.
#include "fmm_internal.h"

namespace xf {

// Single-workgroup exclusive scan: each of the 1024 threads scans a
// contiguous chunk, chunk sums are scanned in LDS.  Used for the small
// per-level arrays (<= a few 10^5 entries) where one launch beats a
// multi-pass scan.  count comes from count_dev when non-null.
__global__ void __launch_bounds__(1024) k_scan1(const int *__restrict__ in, int *__restrict__ out, int count_host,
                                                const int *__restrict__ count_dev, int *__restrict__ total) {
  __shared__ int part[1024];
  const int count = count_dev ? *count_dev : count_host;
  const int t = threadIdx.x;
  const int chunk = (count + 1023) / 1024;
  const int b = t * chunk, e = min(b + chunk, count);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  for (int i = b; i < e; i++) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
  if (t == 1023 && total) *total = part[1023];
}

void scan_exclusive(const int *in, int *out, int count, int *total, hipStream_t s) {
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, in, out, count, nullptr, total);
  XF_LAUNCH();
}

void scan_exclusive_dev(const int *in, int *out, const int *count_dev, int *total, hipStream_t s) {
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, in, out, 0, count_dev, total);
  XF_LAUNCH();
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
