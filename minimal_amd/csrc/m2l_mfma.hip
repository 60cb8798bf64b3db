// m2l_mfma.hip -- M2L grouped by translation key, applied on MFMA.
//
// Reference: M2L 3d/kernel.h:233-260 (singular harmonics 3d/kernel.h:83-116).
//
// Radius scaling makes the translation operator of a pair depend only on
// its key (integer offset d in units of the finer radius, and the level
// difference): with M~_n^m = M_n^m / Rs^n and L~_j^k = L_j^k Rc^{j+1},
//   L~ = T(d) M~,  T = reference coefficients x I_{j+n}(d) x 2^{a(j+1) + b n},
// where Rc = Rf 2^a, Rs = Rf 2^b.  In real form (P^2 real degrees of freedom
// per expansion: Re M_n^0, Re/Im M_n^m for m > 0) T is a dense P^2 x P^2
// matrix, generated once per distinct key in FP64.
//
// Targets are grouped in tiles of 256 cells of one level and one octant
// position (cells with the same position relative to their parent have the
// same key set away from the domain boundary).  For each key in a tile's
// key union the 256 source columns are gathered and multiplied by the key's
// operator; all keys of a tile accumulate into the same MFMA accumulators,
// so each target's L is written once, without atomics, in a fixed order.
//
// Precision: FP32 uses v_mfma_f32_16x16x4_f32.  BF16X3 splits operator and
// multipoles into bf16 hi + lo parts and forms hi*hi + hi*lo + lo*hi with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulate), ~16 significant bits per
// product.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

// GEMM workgroup: 8 waves = column groups of 64 targets x KPAR k-parities
// (a wave multiplies the k-steps kk = kp, kp + KPAR, ... of its columns):
// tiles of 512 targets with one parity, 256 with two.
constexpr int kGemmThreads = 512;
constexpr int kGemmWaves = kGemmThreads / 64;
template <int KPAR>
constexpr int tile_width() {
  return 64 * (kGemmWaves / KPAR);
}
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct Geom {
  int P, PP, R;  // R = row tiles of 16 (rows padded to 16R)
  int K4;        // FP32 k-steps of 4
  int K32;       // BF16 k-steps of 32
};
// BF16 multipole rows: P^2 padded to a multiple of 8, plus a tail of zeros
// after the zero row that covers the last k-step's k-groups.
constexpr int kBf16Tail = 32;
static int bf16_row(const Geom &g) { return (g.PP + 7) / 8 * 8; }
// elements of each of the two BF16 cast arrays (hi, lo) of a tree of nc cells
static size_t bf16_cast_elems(const Geom &g, int64_t nc) { return size_t(nc + 1) * bf16_row(g) + kBf16Tail; }

static Geom geom(int P) {
  Geom g;
  g.P = P;
  g.PP = P * P;
  g.R = (g.PP + 15) / 16;
  g.K4 = (g.PP + 3) / 4;
  g.K32 = (g.PP + 31) / 32;
  return g;
}

// ---------------------------------------------------------------- tiles
constexpr int kTileUnroll = 8;  // list entries in flight per thread (list walks below)
// MurmurHash3's 32-bit finaliser: spreads key ids before taking the minimum.
__device__ inline unsigned mix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// Sort key of a target cell: group (level, octant) in bits 16..23 (255 = no
// M2L), and with `minhash` the top 16 bits of the minimum hashed key id of
// its list below: cells with similar key sets share that minimum with a
// probability equal to the sets' Jaccard similarity, so sorting by it packs
// cells of like key sets into the same tile (fewer empty GEMM slots).
// One wave per cell: the lanes take the list entries in turn and the
// minimum hash is a wave reduction.
__global__ void __launch_bounds__(256) k_group_keys(int ncells, const double4 *__restrict__ cx,
                                                    const int *__restrict__ cparent, const int *__restrict__ clevel,
                                                    const int *__restrict__ m2l_beg, const int *__restrict__ m2l_cnt,
                                                    const int2 *__restrict__ list, int minhash, int lmin,
                                                    unsigned long long *__restrict__ key, int *__restrict__ val) {
  const int lane = threadIdx.x & (kWave - 1);
  const int c = blockIdx.x * (256 / kWave) + int(threadIdx.x / kWave);
  if (c >= ncells) return;  // whole wave
  unsigned g = 255, mh = 0;
  // cells above level lmin take their M2L elsewhere (FP64, per pair)
  const int n = clevel[c] < lmin ? 0 : m2l_cnt[c];
  if (n > 0) {
    int o = 0;
    const int p = cparent[c];
    if (p >= 0) {
      const double4 a = cx[c], b = cx[p];
      o = (a.x > b.x) + ((a.y > b.y) << 1) + ((a.z > b.z) << 2);
    }
    g = unsigned(min(clevel[c], 30)) * 8 + o;
    if (minhash) {
      unsigned m = 0xffffffffu;
      const int b0 = m2l_beg[c];
      for (int i = lane; i < n; i += kWave) m = min(m, mix32(unsigned(list[b0 + i].y)));
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) m = min(m, unsigned(__shfl_xor(int(m), d)));
      mh = m >> 16;  // 16 bits: three 8-bit sort passes over the 24-bit key
    }
  }
  if (lane == 0) {
    key[c] = (unsigned long long)(g << 16 | mh);
    val[c] = c;
  }
}

__global__ void __launch_bounds__(256) k_group_hist(int ncells, const unsigned long long *__restrict__ key,
                                                    int *__restrict__ cnt) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (int c = blockIdx.x * 256 + threadIdx.x; c < ncells; c += gridDim.x * 256) atomicAdd(&h[int(key[c] >> 16)], 1);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

// Tile kernels: one thread per target column walks that column's M2L list
// (entries (source << 5 | image, key id)), kTileUnroll entries in flight.

// Distinct keys of one tile: bitmap over key ids in LDS.
// Four threads per column (blockDim.x / 4 columns at a time), entries i = sub mod 4.
__device__ inline void tile_bitmap(const int *cells, int ncol, const int *beg, const int *cnt, const int2 *list,
                                   unsigned *bm, int nwords) {
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) bm[i] = 0;
  __syncthreads();
  const int sub = threadIdx.x & 3;
  for (int t = threadIdx.x >> 2; t < ncol; t += blockDim.x >> 2) {
    const int c = cells[t];
    const int b = beg[c], n = cnt[c];
    for (int i0 = sub; i0 < n; i0 += 4 * kTileUnroll) {
      int id[kTileUnroll];
#pragma unroll
      for (int k = 0; k < kTileUnroll; k++) id[k] = i0 + 4 * k < n ? list[b + i0 + 4 * k].y : -1;
#pragma unroll
      for (int k = 0; k < kTileUnroll; k++)
        if (id[k] >= 0) atomicOr(&bm[id[k] >> 5], 1u << (id[k] & 31));
    }
  }
  __syncthreads();
}

// Tiles from the group populations gc[0..254] (group 255 = cells without
// M2L, not tiled): each group's cells, in sorted order, are cut into tiles
// of at most tn.  Writes each tile's first cell and count, and plan[0] =
// tiles, plan[1] = cells in tiles.
__global__ void __launch_bounds__(256) k_tile_plan(const int *__restrict__ gc, int tn, int *__restrict__ tfirst,
                                                   int *__restrict__ tcount, int *__restrict__ plan) {
  __shared__ int tb[256], cb[256];
  const int gi = threadIdx.x;
  const int n = gi < 255 ? gc[gi] : 0;
  tb[gi] = (n + tn - 1) / tn;
  cb[gi] = n;
  __syncthreads();
  if (gi == 0) {  // exclusive prefixes of tiles and cells over the groups
    int t = 0, cc = 0;
    for (int i = 0; i < 256; i++) {
      const int a = tb[i], b = cb[i];
      tb[i] = t;
      cb[i] = cc;
      t += a;
      cc += b;
    }
    plan[0] = t;
    plan[1] = cc;
  }
  __syncthreads();
  for (int j = 0; j * tn < n; j++) {
    tfirst[tb[gi] + j] = cb[gi] + j * tn;
    tcount[tb[gi] + j] = min(tn, n - j * tn);
  }
}

constexpr int kTileCountThreads = 1024;
__global__ void __launch_bounds__(kTileCountThreads) k_tile_count(const int *__restrict__ tfirst, const int *__restrict__ tcount,
                                                    const int *__restrict__ sorted, const int *__restrict__ beg,
                                                    const int *__restrict__ cnt, const int2 *__restrict__ list,
                                                    int nkeys, const int *__restrict__ plan, int *__restrict__ tU,
                                                    unsigned *__restrict__ tbits) {
  extern __shared__ unsigned bm[];
  __shared__ int tot;
  if (int(blockIdx.x) >= plan[0]) return;  // launched over the bound of tiles
  const int nwords = (nkeys + 31) / 32;
  const int *cells = sorted + tfirst[blockIdx.x];
  const int ncol = tcount[blockIdx.x];
  tile_bitmap(cells, ncol, beg, cnt, list, bm, nwords);
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  int s = 0;
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) {
    s += __popc(bm[i]);
    tbits[int64_t(blockIdx.x) * nwords + i] = bm[i];  // the key set, for k_tile_fill
  }
  atomicAdd(&tot, s);
  __syncthreads();
  if (threadIdx.x == 0) tU[blockIdx.x] = tot;
}

// Key list (ascending) and source matrix [U][tn] (-1 = no pair) of a tile.
// Workgroup (tile, q) owns columns 64q .. 64q + 63: it marks them empty in
// every key row, then four threads per column walk that column's M2L list
// (entries i = sub mod 4) and write the source cell into the key's row.
// The key set comes from k_tile_count's bitmap.
__global__ void __launch_bounds__(256) k_tile_fill(const int *__restrict__ tfirst, const int *__restrict__ tcount,
                                                   const int *__restrict__ sorted, const int *__restrict__ beg,
                                                   const int *__restrict__ cnt, const int2 *__restrict__ list,
                                                   int nkeys, const int *__restrict__ tbase,
                                                   const unsigned *__restrict__ tbits, int *__restrict__ tkeys,
                                                   int tn, int *__restrict__ tmat) {
  extern __shared__ unsigned bm[];
  const int nwords = (nkeys + 31) / 32;
  unsigned *rank = bm + nwords;  // exclusive popcount prefix per word
  const int tile = blockIdx.x, q = blockIdx.y;
  const int *cells = sorted + tfirst[tile];
  const int ncol = tcount[tile];
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) bm[i] = tbits[int64_t(tile) * nwords + i];
  // exclusive prefix of the word popcounts (block scan in chunks of 256 words)
  __shared__ int carry;
  __shared__ int wsum[256 / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int w0 = 0; w0 < nwords; w0 += 256) {
    const int i = w0 + threadIdx.x;
    const int v = i < nwords ? __popc(bm[i]) : 0;
    int incl = v;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(incl, d);
      if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int before = carry;
    for (int k = 0; k < wv; k++) before += wsum[k];
    if (i < nwords) rank[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 255) carry = before + incl;
    __syncthreads();
  }
  const int base = tbase[tile];
  const int U = carry;
  if (q == 0) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) {
      unsigned w = bm[i];
      int r = rank[i];
      while (w) {
        const int bt = __ffs(w) - 1;
        w &= w - 1;
        tkeys[base + r++] = i * 32 + bt;
      }
    }
  }
  for (int i = threadIdx.x; i < U * 64; i += blockDim.x)
    tmat[(int64_t(base) + i / 64) * tn + q * 64 + (i & 63)] = -1;
  __syncthreads();
  const int col = q * 64 + (threadIdx.x >> 2), sub = threadIdx.x & 3;
  if (col < ncol) {
    const int c = cells[col];
    const int b = beg[c], n = cnt[c];
    for (int i0 = sub; i0 < n; i0 += 4 * kTileUnroll) {
      int2 en[kTileUnroll];
#pragma unroll
      for (int k = 0; k < kTileUnroll; k++) {
        const int i = i0 + 4 * k;
        en[k] = i < n ? list[b + i] : make_int2(0, -1);
      }
#pragma unroll
      for (int k = 0; k < kTileUnroll; k++) {
        const int id = en[k].y;
        if (id < 0) continue;
        const unsigned w = bm[id >> 5];
        const int r = rank[id >> 5] + __popc(w & ((1u << (id & 31)) - 1u));
        tmat[(int64_t(base) + r) * tn + col] = en[k].x >> 5;  // source cell (the image is part of the key)
      }
    }
  }
}

// Learned orders are run in a few order classes (ranges of orders, one
// GEMM launch each): of[p] = class of order p, hi[k] = highest order of
// class k.
struct OrderClasses {
  int n;
  unsigned char of[kMaxP + 1], hi[kMaxP + 1];
};

// Each tile's slots grouped by order class (ascending class, key order
// within a class), so that one GEMM launch per class covers a contiguous
// run of every tile's slots.  One wave per tile; per 64 slots, one ballot
// per class gives every slot its rank.  perm[tbase[t] + j] = slot of the
// j-th slot of tile t in that grouping; cls[o * ts + t] = tbase[t] + first
// grouped slot of class o in tile t, cls[(ncls + o) * ts + t] = their count.
__global__ void __launch_bounds__(256) k_order_partition(int ntiles, int ts, const int *__restrict__ tbase,
                                                         const int *__restrict__ tU,
                                                         const unsigned char *__restrict__ tkp, OrderClasses oc,
                                                         int *__restrict__ perm, int *__restrict__ cls) {
  const int ncls = oc.n;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + int(threadIdx.x >> 6);
  if (t >= ntiles) return;  // whole wave
  const int b = tbase[t], U = tU[t];
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  auto cls_of = [&](int u) { return u < U ? int(oc.of[min(int(tkp[b + u]), kMaxP)]) : -1; };
  int cnt = 0;  // lane o < ncls: slots of class o
  for (int u0 = 0; u0 < U; u0 += 64) {
    const int o = cls_of(u0 + lane);
    for (int k = 0; k < ncls; k++) {
      const unsigned long long m = __ballot(o == k);
      if (lane == k) cnt += __popcll(m);
    }
  }
  // exclusive prefix of the counts over the orders
  int incl = cnt;
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  int run = incl - cnt;
  if (lane < ncls) {
    cls[int64_t(lane) * ts + t] = b + run;
    cls[int64_t(ncls + lane) * ts + t] = cnt;
  }
  for (int u0 = 0; u0 < U; u0 += 64) {
    const int o = cls_of(u0 + lane);
    for (int k = 0; k < ncls; k++) {
      const unsigned long long m = __ballot(o == k);
      if (!m) continue;
      const int at = __shfl(run, k);
      if (o == k) perm[b + at + __popcll(m & below)] = b + u0 + lane;
      if (lane == k) run += __popcll(m);
    }
  }
}

// Per order (wave o), the exclusive prefix over the tiles of that order's
// slot counts, and their total: the stream-K schedule of the order's launch.
__global__ void __launch_bounds__(1024) k_order_prefix(int ntiles, int ts, int ncls, int *__restrict__ cls) {
  const int lane = threadIdx.x & 63, o = int(threadIdx.x >> 6);
  if (o >= ncls) return;
  const int *cnt = cls + int64_t(ncls + o) * ts;
  int *pre = cls + int64_t(2 * ncls + o) * ts;
  int carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += 64) {
    const int t = t0 + lane;
    const int v = t < ntiles ? cnt[t] : 0;
    int incl = v;
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(incl, d);
      if (lane >= d) incl += u;
    }
    if (t < ntiles) pre[t] = carry + incl - v;
    carry += __shfl(incl, 63);
  }
  if (lane == 0) cls[int64_t(3 * ncls) * ts + o] = carry;
}

// ---------------------------------------------------------------- operands
// GEMM operands from the scaled FP64 multipoles M~ (P^2 per cell, rows of
// KP): FP32 copy, or bf16 hi + lo split (hi = bf16(x), lo = bf16(x - hi)).
// Row ncells is the zero row; `tail` more zeros follow it, so a k-group load
// of the zero row past KP stays inside the array and reads zeros.
__global__ void __launch_bounds__(256) k_mt(int PP, int KP, int ncells, int tail, const double *__restrict__ M,
                                            float *__restrict__ mt32, __bf16 *__restrict__ mhi,
                                            __bf16 *__restrict__ mlo) {
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int c = int(g / KP), r = int(g % KP);
  if (g >= int64_t(ncells + 1) * KP + tail) return;
  // row ncells is all zeros: the GEMM's loads for empty slots read it
  const float v = (c < ncells && r < PP) ? float(M[int64_t(c) * PP + r]) : 0.0f;
  if (mt32) mt32[g] = v;
  if (mhi) {
    const __bf16 h = __bf16(v);
    mhi[g] = h;
    mlo[g] = __bf16(v - float(h));
  }
}

// One workgroup per key: singular harmonics of d in FP64, then the operator
// in MFMA fragment order.
//  FP32 (16x16x4):  op32[key][kk][r][lane], element (row, kcol):
//                   r = row/16, kk = kcol/4, lane = (kcol%4)*16 + row%16
//  BF16 (16x16x32): ophi/oplo[key][kk][r][lane][8], element (row, kcol):
//                   r = row/16, kk = kcol/32, lane = ((kcol%32)/8)*16 + row%16
__global__ void __launch_bounds__(256) k_ops_gen(int P, int fillP, int R, int K4, int K32, double R0, double cycle,
                                                 const unsigned long long *__restrict__ keys, float *__restrict__ op32,
                                                 __bf16 *__restrict__ ophi, __bf16 *__restrict__ oplo) {
  __shared__ cpx<double> I[(2 * kMaxP) * (2 * kMaxP + 1) / 2];
  const int id = blockIdx.x;
  double dx, dy, dz;
  int a, b;
  key_at(keys, id, R0, cycle, dx, dy, dz, a, b);
  if (threadIdx.x < fillP) singular_column(threadIdx.x, fillP, dx, dy, dz, I);
  __syncthreads();
  if (op32) {
    const int64_t per = int64_t(K4) * R * 64;
    for (int64_t e = threadIdx.x; e < per; e += blockDim.x) {
      const int lane = int(e % 64), r = int((e / 64) % R), kk = int(e / (64 * R));
      const int row = r * 16 + lane % 16, kc = kk * 4 + lane / 16;
      op32[id * per + e] = float(op_entry(P, fillP, row, kc, a, b, I));
    }
  }
  if (ophi) {
    const int64_t per = int64_t(K32) * R * 64 * 8;
    for (int64_t e = threadIdx.x; e < per; e += blockDim.x) {
      const int i = int(e % 8), lane = int((e / 8) % 64), r = int((e / 512) % R), kk = int(e / (512 * R));
      const int row = r * 16 + lane % 16, kc = kk * 32 + (lane / 16) * 8 + i;
      const float v = float(op_entry(P, fillP, row, kc, a, b, I));
      const __bf16 h = __bf16(v);
      ophi[id * per + e] = h;
      oplo[id * per + e] = __bf16(v - float(h));
    }
  }
}

// ---------------------------------------------------------------- GEMM
// One workgroup = one tile of up to tile_width<KPAR>() targets and a block
// of RTB row tiles (blockIdx.y).  8 waves = 8 / KPAR column groups (64
// targets = 4 column tiles of 16) x KPAR k-parities: wave (g, kp) multiplies
// the k-steps kk = kp, kp + KPAR, ... of every key for its 64 targets and all
// RTB row tiles, so each source multipole is fetched once per tile; two
// parities are added through LDS at the end.  With one parity the tile is
// 512 targets wide, so each staged operator serves twice the targets (the
// operator staging is most of the GEMM's fetch).  Per key the block's operator rows are copied
// global -> LDS (global_load_lds, no register round trip) into the second
// of two buffers while the first is multiplied; one wait + barrier per key.
// Each lane's source multipoles for its next k-step are loaded while the
// current one is multiplied.  All keys of a tile accumulate into the same
// MFMA accumulators, so each target's L is written once, without atomics.
constexpr int kGemmMaxRT = 8;  // row tiles per wave (accumulator registers)

template <int RH>
__device__ inline void write_locals(int PP, int rbase, int nr, const int *cells, int ncol, int cg,
                                    const v4f (&acc)[RH][4], int lane, double *L) {
#pragma unroll
  for (int ct = 0; ct < 4; ct++) {
    const int col = cg * 64 + ct * 16 + (lane & 15);
    if (col >= ncol) continue;
    double *Lc = L + int64_t(cells[col]) * PP;
#pragma unroll
    for (int r = 0; r < RH; r++) {
      if (r >= nr) break;  // row tiles of this block only (another block owns the next ones)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = (rbase + r) * 16 + 4 * (lane >> 4) + i;
        if (row < PP) Lc[row] += double(acc[r][ct][i]);
      }
    }
  }
}

// Split-key variant: this workgroup's share of the locals, stored (each
// (cell, row) of a tile belongs to one workgroup per key slice).
template <int RH>
__device__ inline void write_partial(int PP, int rbase, int nr, const int *cells, int ncol, int cg,
                                     const v4f (&acc)[RH][4], int lane, float *Lp) {
#pragma unroll
  for (int ct = 0; ct < 4; ct++) {
    const int col = cg * 64 + ct * 16 + (lane & 15);
    if (col >= ncol) continue;
    float *Lc = Lp + int64_t(cells ? cells[col] : col) * PP;
#pragma unroll
    for (int r = 0; r < RH; r++) {
      if (r >= nr) break;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = (rbase + r) * 16 + 4 * (lane >> 4) + i;
        if (row < PP) Lc[row] = acc[r][ct][i];
      }
    }
  }
}

// Adds the key slices of the split-key GEMM into the locals (fixed order).
__global__ void k_combine_slices(int PP, int64_t ncells, int nslice, const int *__restrict__ sorted, int ncell_tiled,
                                 const float *__restrict__ Lpart, double *__restrict__ L) {
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (g >= int64_t(ncell_tiled) * PP) return;
  const int c = sorted[g / PP], r = int(g % PP);
  double s = 0.0;
  for (int z = 0; z < nslice; z++) s += double(Lpart[(int64_t(z) * ncells + c) * PP + r]);
  L[int64_t(c) * PP + r] += s;
}

// Parity 1 hands its accumulators to parity 0 through LDS (the operator
// buffers are free by then); parity 0 adds them and writes the locals.
template <int RH, int KPAR>
__device__ inline void reduce_parities(v4f (&acc)[RH][4], int kp, int cg, int lane, void *lds) {
  if (KPAR == 1) return;  // one wave per column group: nothing to add
  v4f *red = reinterpret_cast<v4f *>(lds);
  __syncthreads();
  if (kp == 1) {
#pragma unroll
    for (int r = 0; r < RH; r++)
#pragma unroll
      for (int ct = 0; ct < 4; ct++) red[((cg * RH + r) * 4 + ct) * 64 + lane] = acc[r][ct];
  }
  __syncthreads();
  if (kp == 0) {
#pragma unroll
    for (int r = 0; r < RH; r++)
#pragma unroll
      for (int ct = 0; ct < 4; ct++) acc[r][ct] += red[((cg * RH + r) * 4 + ct) * 64 + lane];
  }
}

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// 6:4, lgkmcnt 11:8) that wait for vmcnt <= N and nothing else.
constexpr int waitcnt_vm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
constexpr int kWaitVm0 = waitcnt_vm(0), kWaitVm4 = waitcnt_vm(4), kWaitVm8 = waitcnt_vm(8);

// Row of the full operator layout [kk][row tile][lane] for run c of a
// block's staged copy (c = kk * nrt + local row tile).
__device__ inline int op_run(int c, int nrt, int RT, int rt0) {
  const int kk = c / nrt, rl = c % nrt;
  return kk * RT + rt0 + rl;
}

template <int RH, int KPAR>
__global__ void __launch_bounds__(kGemmThreads) k_m2l_f32(int PP, int RT, int RTB, int K4, int KP,
                                                          const int *__restrict__ tfirst,
                                                          const int *__restrict__ tcount,
                                                          const int *__restrict__ tbase, const int *__restrict__ tU,
                                                          const int *__restrict__ sorted,
                                                          const int *__restrict__ tkeys,
                                                          const int *__restrict__ tmat,
                                                          const float *__restrict__ op32,
                                                          const float *__restrict__ mt,
                                                          const unsigned char *__restrict__ tkp, int kr,
                                                          int64_t ncells, float *__restrict__ Lpart,
                                                          double *__restrict__ L) {
  extern __shared__ float As[];
  constexpr int TN = tile_width<KPAR>(), NG = kGemmWaves / KPAR, PARTS = TN / 64;
  // wave index read through a scalar register: the k-step and key loops
  // below then branch uniformly instead of under execution masks
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cg = w % NG, kp = w / NG;
  const int tile = blockIdx.x;
  const int ncol = tcount[tile];
  // keys of this workgroup: slice blockIdx.z of gridDim.z of the tile's keys
  const int Ut = tU[tile], ub = int(int64_t(Ut) * blockIdx.z / gridDim.z),
            ue = int(int64_t(Ut) * (blockIdx.z + 1) / gridDim.z);
  const int U = ue - ub, base = tbase[tile] + ub;
  const int rt0 = blockIdx.y * RTB;
  const int nrt = min(RTB, RT - rt0);
  const int kq = lane >> 4;
  v4f acc[RH][4];
#pragma unroll
  for (int r = 0; r < RH; r++)
#pragma unroll
    for (int ct = 0; ct < 4; ct++) acc[r][ct] = v4f{0, 0, 0, 0};
  const int64_t per = int64_t(K4) * RT * 64;
  const int NC = K4 * nrt;  // runs of 64 floats per staged block
  const int BUF = NC * 64;  // floats per LDS buffer
  // order of key u (tkp: per-key orders of a variable-order evaluation):
  // k-steps and row tiles of this block that it needs, and its K bound
  auto kcols = [&](int u) { const int p = tkp ? tkp[base + u] : (int)__builtin_sqrtf(float(PP) + 0.5f); return p * p; };
  auto ksteps = [&](int kc) { return (kc + 3) / 4; };
  auto rtiles = [&](int kc) { return max(0, min(nrt, (kc + 15) / 16 - rt0)); };
  if (U > 0) {
    // LDS: operator slots [2 buffers][kr keys][BUF], then the keys' column
    // sources [3 buffers][kr keys][TN]; staging as in the BF16 kernel below
    int *Ts = reinterpret_cast<int *>(As + size_t(2) * kr * BUF);
    const int R = (U + kr - 1) / kr;
    auto stage_ops = [&](int r) {
      for (int u = r * kr; u < min(r * kr + kr, U); u++) {
        float *dst = As + size_t((r & 1) * kr + u - r * kr) * BUF;
        const int kc = kcols(u), ks = ksteps(kc), nr = rtiles(kc);
        const float *Ak = op32 + int64_t(tkeys[base + u]) * per;
        for (int j = w; j < ks * nr; j += kGemmWaves) {
          const int kk = j / nr, rl = j - kk * nr, c = kk * nrt + rl;
          __builtin_amdgcn_global_load_lds(Ak + int64_t(op_run(c, nrt, RT, rt0)) * 64 + lane, dst + c * 64, 4, 0, 0);
        }
      }
    };
    auto stage_src = [&](int r) {
      const int nu = min(kr, U - r * kr);
      for (int j = w; j < PARTS * nu; j += kGemmWaves) {
        const int ui = j / PARTS, q = j % PARTS;
        __builtin_amdgcn_global_load_lds(tmat + (int64_t(base) + r * kr + ui) * TN + q * 64 + lane,
                                         Ts + ((r % 3) * kr + ui) * TN + q * 64, 4, 0, 0);
      }
    };
    auto col_src = [&](int u, int (&sc)[4]) {
      const int *T = Ts + (((u / kr) % 3) * kr + u % kr) * TN + cg * 64 + (lane & 15);
#pragma unroll
      for (int ct = 0; ct < 4; ct++) sc[ct] = T[ct * 16];
    };
    // empty slots and k at or past the key's K bound read the zero row
    // (index ncells): unconditional loads stay in flight until used
    auto load_b = [&](const int (&sc)[4], int kk, int kc, float (&bv)[4]) {
      const int k = kk * 4 + kq;
#pragma unroll
      for (int ct = 0; ct < 4; ct++) bv[ct] = mt[int64_t(sc[ct] >= 0 && k < kc ? sc[ct] : ncells) * KP + k];
    };
    stage_ops(0);
    stage_src(0);
    if (R > 1) stage_src(1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int src[4], nsrc[4];
    float bcur[4], bnext[4];
    col_src(0, src);
    int kc = kcols(0);
    if (kp < ksteps(kc)) load_b(src, kp, kc, bnext);
    for (int r = 0; r < R; r++) {
      if (r + 1 < R) stage_ops(r + 1);
      if (r + 2 < R) stage_src(r + 2);
      __builtin_amdgcn_sched_barrier(0);
      bool b_after = false;
      const int u0 = r * kr, u1 = min(u0 + kr, U);
      for (int u = u0; u < u1; u++) {
        const float *Acur = As + size_t((r & 1) * kr + u - u0) * BUF;
        const int ks = ksteps(kc), nr = rtiles(kc);
        const int kc_next = u + 1 < U ? kcols(u + 1) : 0;
        const bool next = u + 1 < U && kp < ksteps(kc_next);
        if (u + 1 < U) col_src(u + 1, nsrc);
        for (int kk = kp; kk < ks; kk += KPAR) {
#pragma unroll
          for (int ct = 0; ct < 4; ct++) bcur[ct] = bnext[ct];
          if (kk + KPAR < ks) {
            load_b(src, kk + KPAR, kc, bnext);
            b_after = true;
          } else if (next) {
            load_b(nsrc, kp, kc_next, bnext);
            b_after = true;
          }
#pragma unroll
          for (int r2 = 0; r2 < RH; r2++) {
            if (r2 >= nr) break;
            const float a = Acur[(kk * nrt + r2) * 64 + lane];
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
              acc[r2][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bcur[ct], acc[r2][ct], 0, 0, 0);
          }
        }
        if (ks <= kp && next) {
          load_b(nsrc, kp, kc_next, bnext);
          b_after = true;
        }
#pragma unroll
        for (int ct = 0; ct < 4; ct++) src[ct] = nsrc[ct];
        kc = kc_next;
      }
      // staging of this round's start landed (only the 4 loads of the last
      // B prefetch are younger); then every wave is done with this round
      if (b_after)
        __builtin_amdgcn_s_waitcnt(kWaitVm4);
      else
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  reduce_parities<RH, KPAR>(acc, kp, cg, lane, As);
  if (kp == 0) {
    if (Lpart)
      write_partial<RH>(PP, rt0, nrt, sorted + tfirst[tile], ncol, cg, acc, lane,
                        Lpart + int64_t(blockIdx.z) * ncells * PP);
    else
      write_locals<RH>(PP, rt0, nrt, sorted + tfirst[tile], ncol, cg, acc, lane, L);
  }
}

// Keys ub .. ue - 1 of tile bx, row block by.  Output: added into the FP64
// locals L (Lpart null), or stored as FP32 partial locals, by cell
// (Lpart + cell * PP, part_local false) or by tile column (Lpart + col * PP).
// Order of one GEMM launch: pl (k bound pl^2 of every key, unless the
// per-slot orders tkp are given), the row tiles RTo and k-steps K32o it
// covers (of the RT x K32 of the operator layout, which holds order
// sqrt(PP)), and an optional slot permutation (slot u of tile t is
// perm[tbase[t] + u]).
struct OrderPass {
  int pl, RTo, K32o;
  const int *perm;
};

template <int RH, bool SKIP, int KPAR>
__device__ __attribute__((always_inline)) inline void m2l_bf16x3_body(int bx, int by, int ub, int ue, bool part_local,
                                       int PP, int RT, int RTB, int K32, int KP, OrderPass op,
                                       const int *__restrict__ tfirst,
                                       const int *__restrict__ tcount,
                                       const int *__restrict__ tbase,
                                       const int *__restrict__ tU,
                                       const int *__restrict__ sorted,
                                       const int *__restrict__ tkeys,
                                       const int *__restrict__ tmat,
                                       const v8bf *__restrict__ ophi,
                                       const v8bf *__restrict__ oplo,
                                       const __bf16 *__restrict__ mhi,
                                       const __bf16 *__restrict__ mlo,
                                       const unsigned char *__restrict__ tkp, int kr,
                                       int64_t ncells, float *__restrict__ Lpart,
                                       double *__restrict__ L) {
  extern __shared__ v8bf Ab[];
  constexpr int TN = tile_width<KPAR>(), NG = kGemmWaves / KPAR, PARTS = TN / 64;
  // wave index read through a scalar register: the k-step and key loops
  // below then branch uniformly instead of under execution masks
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cg = w % NG, kp = w / NG;
  const int tile = bx;
  const int ncol = tcount[tile];
  // column subtiles of 16 that hold target cells (with SKIP, a partly
  // filled tile skips the MFMAs of its empty subtiles, whose B operands are
  // the zero row)
  const int nct = max(0, min(4, (ncol - cg * 64 + 15) / 16));
  const int U = ue - ub, base = tbase[tile] + ub;
  // a tile without keys in this launch adds nothing to the locals
  if (U <= 0 && !Lpart) return;
  const int rt0 = by * RTB;
  const int nrt = min(RTB, op.RTo - rt0);
  const int kq = lane >> 4;
  v4f acc[RH][4];
#pragma unroll
  for (int r = 0; r < RH; r++)
#pragma unroll
    for (int ct = 0; ct < 4; ct++) acc[r][ct] = v4f{0, 0, 0, 0};
  const int64_t per = int64_t(K32) * RT * 64;  // v8bf per key and part
  const int NC = op.K32o * nrt;                // runs of 64 v8bf per staged part
  const int E = NC * 64;                       // v8bf per staged part
  auto slot = [&](int u) { return op.perm ? op.perm[base + u] : base + u; };
  // order of key u (tkp: per-key orders of a variable-order evaluation):
  // k-steps and row tiles of this block that it needs, and its K bound
  auto kcols = [&](int u) { const int p = tkp ? tkp[slot(u)] : op.pl; return p * p; };
  // a K bound below the operator's order can fall inside a k-group
  const bool varK = tkp || op.pl * op.pl < PP;
  auto ksteps = [&](int kc) { return (kc + 31) / 32; };
  auto rtiles = [&](int kc) { return max(0, min(nrt, (kc + 15) / 16 - rt0)); };
  if (U > 0) {
    // LDS: operator slots [2 buffers][kr keys][hi, lo][E], then the keys'
    // column sources (tile matrix rows) [3 buffers][kr keys][TN].  Round r
    // (keys r kr .. r kr + kr - 1) reads operator buffer r & 1 and source
    // buffer r % 3; at its start it stages round r + 1's operators and round
    // r + 2's sources, so the sources of the next round's first key are
    // already in LDS while this round still runs.
    int *Ts = reinterpret_cast<int *>(Ab + size_t(4) * kr * E);
    const int R = (U + kr - 1) / kr;
    auto stage_ops = [&](int r) {
      for (int u = r * kr; u < min(r * kr + kr, U); u++) {
        v8bf *dst = Ab + size_t((r & 1) * kr + u - r * kr) * 2 * E;
        const int kc = kcols(u), ks = ksteps(kc), nr = rtiles(kc);
        const int64_t k = tkeys[slot(u)];
        const int n1 = ks * nr;
        for (int j = w; j < 2 * n1; j += kGemmWaves) {
          const int part = j >= n1, jj = part ? j - n1 : j;
          const int kk = jj / nr, rl = jj - kk * nr, cc = kk * nrt + rl;
          const v8bf *A = (part ? oplo : ophi) + k * per;
          __builtin_amdgcn_global_load_lds(A + int64_t(op_run(cc, nrt, RT, rt0)) * 64 + lane,
                                           dst + part * E + cc * 64, 16, 0, 0);
        }
      }
    };
    auto stage_src = [&](int r) {
      const int nu = min(kr, U - r * kr);
      for (int j = w; j < PARTS * nu; j += kGemmWaves) {
        const int ui = j / PARTS, q = j % PARTS;
        __builtin_amdgcn_global_load_lds(tmat + int64_t(slot(r * kr + ui)) * TN + q * 64 + lane,
                                         Ts + ((r % 3) * kr + ui) * TN + q * 64, 4, 0, 0);
      }
    };
    auto col_src = [&](int u, int (&sc)[4]) {
      const int *T = Ts + (((u / kr) % 3) * kr + u % kr) * TN + cg * 64 + (lane & 15);
#pragma unroll
      for (int ct = 0; ct < 4; ct++) sc[ct] = T[ct * 16];
    };
    // Empty slots and k-groups at or past the key's K bound read the zero
    // row (index ncells), so every load is unconditional and stays in
    // flight until its k-step: masking at load time would make the wave
    // wait for the data right away.  The one vector that straddles a
    // variable order's bound is masked when it is used (mask_b).
    auto load_b = [&](const int (&sc)[4], int kk, int kc, v8bf (&bh)[4], v8bf (&bl)[4]) {
      const int k0 = kk * 32 + kq * 8;
#pragma unroll
      for (int ct = 0; ct < 4; ct++) {
        const int64_t o = int64_t(sc[ct] >= 0 && k0 < kc ? sc[ct] : ncells) * KP + k0;
        bh[ct] = *reinterpret_cast<const v8bf *>(mhi + o);
        bl[ct] = *reinterpret_cast<const v8bf *>(mlo + o);
      }
    };
    auto mask_b = [&](int kk, int kc, v8bf (&bh)[4], v8bf (&bl)[4]) {
      const int k0 = kk * 32 + kq * 8;
      if (k0 >= kc || k0 + 8 <= kc) return;
#pragma unroll
      for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int i = 0; i < 8; i++)
          if (k0 + i >= kc) {
            bh[ct][i] = __bf16(0.0f);
            bl[ct][i] = __bf16(0.0f);
          }
    };
    stage_ops(0);
    stage_src(0);
    if (R > 1) stage_src(1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int src[4], nsrc[4];
    v8bf bh[4], bl[4], nbh[4], nbl[4];
    col_src(0, src);
    int kc = kcols(0);
    if (kp < ksteps(kc)) load_b(src, kp, kc, nbh, nbl);
    for (int r = 0; r < R; r++) {
      if (r + 1 < R) stage_ops(r + 1);
      if (r + 2 < R) stage_src(r + 2);
      __builtin_amdgcn_sched_barrier(0);
      // B prefetches issued after this round's staging: at most the last
      // one is still in flight when the round ends
      bool b_after = false;
      const int u0 = r * kr, u1 = min(u0 + kr, U);
      for (int u = u0; u < u1; u++) {
        const v8bf *Ahs = Ab + size_t((r & 1) * kr + u - u0) * 2 * E, *Als = Ahs + E;
        const int ks = ksteps(kc), nr = rtiles(kc);
        const int kc_next = u + 1 < U ? kcols(u + 1) : 0;
        const bool next = u + 1 < U && kp < ksteps(kc_next);
        if (u + 1 < U) col_src(u + 1, nsrc);  // in LDS: round r or r + 1
        for (int kk = kp; kk < ks; kk += KPAR) {
#pragma unroll
          for (int ct = 0; ct < 4; ct++) {
            bh[ct] = nbh[ct];
            bl[ct] = nbl[ct];
          }
          // only an order below the operator's has a bound inside the padded K range
          if (varK && kk * 32 + 32 > kc) mask_b(kk, kc, bh, bl);
          if (kk + KPAR < ks) {
            load_b(src, kk + KPAR, kc, nbh, nbl);
            b_after = true;
          } else if (next) {
            load_b(nsrc, kp, kc_next, nbh, nbl);
            b_after = true;
          }
#pragma unroll
          for (int r2 = 0; r2 < RH; r2++) {
            if (r2 >= nr) break;
            const int ai = (kk * nrt + r2) * 64 + lane;
            const v8bf a_hi = Ahs[ai], a_lo = Als[ai];
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
              if (SKIP && ct >= nct) break;
              acc[r2][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, bh[ct], acc[r2][ct], 0, 0, 0);
              acc[r2][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, bl[ct], acc[r2][ct], 0, 0, 0);
              acc[r2][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo, bh[ct], acc[r2][ct], 0, 0, 0);
            }
          }
        }
        if (ks <= kp && next) {
          load_b(nsrc, kp, kc_next, nbh, nbl);
          b_after = true;
        }
#pragma unroll
        for (int ct = 0; ct < 4; ct++) src[ct] = nsrc[ct];
        kc = kc_next;
      }
      // the staging issued at this round's start has landed (loads complete
      // in issue order; only the 8 loads of the last B prefetch are younger),
      // then every wave is done with this round's buffers
      if (b_after)
        __builtin_amdgcn_s_waitcnt(kWaitVm8);
      else
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  reduce_parities<RH, KPAR>(acc, kp, cg, lane, Ab);
  if (kp == 0) {
    if (Lpart)
      write_partial<RH>(PP, rt0, nrt, part_local ? nullptr : sorted + tfirst[tile], ncol, cg, acc, lane, Lpart);
    else
      write_locals<RH>(PP, rt0, nrt, sorted + tfirst[tile], ncol, cg, acc, lane, L);
  }
}

template <int RH, int KPAR>
__global__ void __launch_bounds__(kGemmThreads) k_m2l_bf16x3(int PP, int RT, int RTB, int K32, int KP, OrderPass op,
                                                             const int *__restrict__ tfirst,
                                                             const int *__restrict__ tcount,
                                                             const int *__restrict__ tbase,
                                                             const int *__restrict__ tU,
                                                             const int *__restrict__ sorted,
                                                             const int *__restrict__ tkeys,
                                                             const int *__restrict__ tmat,
                                                             const v8bf *__restrict__ ophi,
                                                             const v8bf *__restrict__ oplo,
                                                             const __bf16 *__restrict__ mhi,
                                                             const __bf16 *__restrict__ mlo,
                                                             const unsigned char *__restrict__ tkp, int kr,
                                                             int64_t ncells, float *__restrict__ Lpart,
                                                             double *__restrict__ L) {
  // keys of this workgroup: slice blockIdx.z of gridDim.z of the tile's keys
  const int Ut = tU[blockIdx.x];
  const int ub = int(int64_t(Ut) * blockIdx.z / gridDim.z), ue = int(int64_t(Ut) * (blockIdx.z + 1) / gridDim.z);
  m2l_bf16x3_body<RH, false, KPAR>(blockIdx.x, blockIdx.y, ub, ue, false, PP, RT, RTB, K32, KP, op, tfirst, tcount, tbase,
                             tU, sorted, tkeys, tmat, ophi, oplo, mhi, mlo, tkp, kr, ncells,
                             Lpart ? Lpart + int64_t(blockIdx.z) * ncells * PP : nullptr, L);
}

// Stream-K schedule: the tiles' key slots, in tile order (slot tbase[t] + u
// is key u of tile t; S slots in all), are split evenly over the gridDim.x
// workgroups of one persistent launch (one per CU).  Workgroup w takes slots
// [w S / G, (w + 1) S / G); each tile it touches gives one FP32 partial of the
// tile's locals, stored by tile column in segment w + t of Lpart, and
// k_sk_combine adds a tile's segments in workgroup order.  Every workgroup
// gets the same number of slots, so none waits on a straggler.  A launch of
// one order of a learned-order evaluation schedules that order's slots:
// tbase is then the exclusive prefix of their per-tile counts tU, tstart[t]
// where tile t's run of them starts in the order-grouped slots, and Sdev
// their total.
template <int RH, int KPAR>
__global__ void __launch_bounds__(kGemmThreads) k_m2l_bf16x3_sk(int PP, int RT, int K32, int KP, OrderPass op,
                                                                int ntiles, int64_t S, const int *__restrict__ Sdev,
                                                                const int *__restrict__ tfirst,
                                                                const int *__restrict__ tcount,
                                                                const int *__restrict__ tbase,
                                                                const int *__restrict__ tstart,
                                                                const int *__restrict__ tU,
                                                                const int *__restrict__ sorted,
                                                                const int *__restrict__ tkeys,
                                                                const int *__restrict__ tmat,
                                                                const v8bf *__restrict__ ophi,
                                                                const v8bf *__restrict__ oplo,
                                                                const __bf16 *__restrict__ mhi,
                                                                const __bf16 *__restrict__ mlo,
                                                                const unsigned char *__restrict__ tkp, int kr,
                                                                int64_t ncells, float *__restrict__ Lseg) {
  const int64_t G = gridDim.x, w = blockIdx.x;
  if (Sdev) S = *Sdev;  // one order's slot count, summed on the device
  const int64_t s0 = w * S / G, s1 = (w + 1) * S / G;
  if (s0 >= s1) return;
  // first tile holding slot s0: the last t with tbase[t] <= s0
  int lo = 0, hi = ntiles - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tbase[mid] <= s0)
      lo = mid;
    else
      hi = mid - 1;
  }
  for (int t = lo; t < ntiles && tbase[t] < s1; t++) {
    const int64_t b = tbase[t];
    const int ub = int(max(s0, b) - b), ue = int(min(s1, b + tU[t]) - b);
    if (ub >= ue) continue;
    m2l_bf16x3_body<RH, false, KPAR>(t, 0, ub, ue, true, PP, RT, op.RTo, K32, KP, op, tfirst, tcount, tstart, tU, sorted,
                               tkeys, tmat, ophi, oplo, mhi, mlo, tkp, kr, ncells,
                               Lseg + (w + t) * int64_t(tile_width<KPAR>()) * PP, nullptr);
    __syncthreads();  // the LDS of this segment is free before the next one stages
  }
}

// Sum of a tile's stream-K segments (workgroups wf .. wl in order) into the
// FP64 locals, rows 0 .. rows - 1 (those the launch computed).  One thread
// per (tile column, row).
__global__ void __launch_bounds__(256) k_sk_combine(int PP, int rows, int ntiles, int64_t S, const int *__restrict__ Sdev,
                                                    int G, int tn, const int *__restrict__ tfirst,
                                                    const int *__restrict__ tcount, const int *__restrict__ tbase,
                                                    const int *__restrict__ tU, const int *__restrict__ sorted,
                                                    const float *__restrict__ Lseg, double *__restrict__ L) {
  const int t = blockIdx.y;
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int col = int(e / rows), row = int(e % rows);
  if (t >= ntiles || col >= tcount[t] || tU[t] == 0) return;
  if (Sdev) S = *Sdev;
  // workgroup of slot s: the largest w with w S / G <= s
  auto wg_of = [&](int64_t sl) { return int(((sl + 1) * G - 1) / S); };
  const int64_t b = tbase[t];
  const int wf = wg_of(b), wl = wg_of(b + tU[t] - 1);
  double acc = 0.0;
  for (int w = wf; w <= wl; w++) acc += double(Lseg[(int64_t(w + t) * tn + col) * PP + row]);
  L[int64_t(sorted[tfirst[t] + col]) * PP + row] += acc;
}

// One problem of a grouped keyed-GEMM launch (all problems of one order in
// a batch): its tiles, operators, cast multipoles and locals.
struct GroupGemm {
  const int *tfirst, *tcount, *tbase, *tU, *sorted, *tkeys, *tmat;
  const v8bf *ophi, *oplo;
  const __bf16 *mhi, *mlo;
  double *L;
  int64_t ncells;
  int ntiles;
};

// All problems of one order in one launch: each problem has ntiles x
// (row blocks) workgroups, one key slice, kr = 1.
template <int RH, int KPAR>
__global__ void __launch_bounds__(kGemmThreads) k_m2l_bf16x3_grouped(const GroupGemm *__restrict__ g,
                                                                     const int *__restrict__ start, int nprob,
                                                                     int PP, int RT, int RTB, int K32, int KP,
                                                                     int pl, int kr) {
  const int b = blockIdx.x, i = group_of(start, nprob, b), lb = b - start[i];
  const GroupGemm &q = g[i];
  m2l_bf16x3_body<RH, true, KPAR>(lb % q.ntiles, lb / q.ntiles, 0, q.tU[lb % q.ntiles], false, PP, RT, RTB, K32, KP,
                            OrderPass{pl, RT, K32, nullptr},
                            q.tfirst, q.tcount, q.tbase,
                            q.tU, q.sorted, q.tkeys, q.tmat, q.ophi, q.oplo, q.mhi, q.mlo, nullptr, kr, q.ncells,
                            nullptr, q.L);
}


// Row tiles per workgroup: two operator buffers within the LDS budget, at
// most kGemmMaxRT accumulator row tiles per wave, and room for the parity
// reduction (4 x RTB x 4 KB).
static int rows_per_block(const Geom &g, bool bf16, int device, int kr, int tn) {
  int max_lds = 0;
  XF_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  const int per_rt = 2 * kr * (bf16 ? g.K32 * 64 * 16 * 2 : g.K4 * 64 * 4);  // both buffers of kr keys
  const int need = std::max(per_rt, 4 * 4 * 64 * 16);                   // or the reduction
  const int avail = max_lds - 3 * kr * tn * 4;                           // less the column sources
  if (need > avail) {
    if (kr == 1) fail(EXAFMM_ERR_UNSUPPORTED, "expansion order too large for the LDS operator stage");
    return 0;
  }
  int rtb = std::min({g.R, kGemmMaxRT, avail / need});
  // balance the row blocks
  const int nblk = ceil_div(g.R, rtb);
  return ceil_div(g.R, nblk);
}

static size_t gemm_lds(const Geom &g, bool bf16, int RTB, int kr, int tn) {
  const size_t ops = size_t(2) * kr * RTB * (bf16 ? g.K32 * 64 * 16 * 2 : g.K4 * 64 * 4) +
                     size_t(3) * kr * tn * sizeof(int);  // operator slots + column sources
  const size_t red = size_t(4) * RTB * 4 * 64 * 16;
  return std::max(ops, red);
}

// One order's launch of a learned-order BF16x3 GEMM: order p over the
// slots of that order (tile t: grouped slots cb[t] .. cb[t] + cnt[t] - 1,
// slot permutation perm).
struct ClassPass {
  int p;          // highest order of the class
  bool variable;  // the class holds several orders: each slot runs at its own (tile_p)
  const int *perm;
  const int *start, *cnt;  // per tile: first order-grouped slot (tbase included) and count
  const int *pre, *tot;    // per tile: exclusive prefix of the counts; their total (stream-K)
};

template <int RH, int KPAR>
static void launch_gemm(Ctx &tc, const Ctx &sc, const Geom &g, int ntiles, bool bf16, int KP, int KPb, int RTB,
                        int kr, int ksplit, float *Lpart, double *Lout, const ClassPass *cp) {
  constexpr int TN = tile_width<KPAR>();
  Ctx &c = tc;
  const int ts = c.m2l_tstride;
  const int *tfirst = c.tiles.p, *tcount = c.tiles.p + ts, *tbase = c.tiles.p + 2 * ts, *tU = c.tiles.p + 3 * ts;
  const int *sorted = c.tiles.p + 4 * ts;
  const Geom go = cp ? geom(cp->p) : g;  // rows and k-steps this launch covers
  const dim3 grid(ntiles, ceil_div(go.R, RTB), ksplit);
  if (bf16) {
    const v8bf *ophi = reinterpret_cast<const v8bf *>(c.ops32.p);
    const size_t per = size_t(g.K32) * g.R * 64 * 8;
    const v8bf *oplo = reinterpret_cast<const v8bf *>(reinterpret_cast<const __bf16 *>(c.ops32.p) +
                                                      per * c.lists.nkeys);
    const __bf16 *mhi = reinterpret_cast<const __bf16 *>(sc.mt32.p);
    const __bf16 *mlo = mhi + bf16_cast_elems(g, sc.t.ncells);
    const size_t lds = gemm_lds(go, true, RTB, kr, TN);
    const OrderPass op{go.P, go.R, go.K32, cp ? cp->perm : nullptr};
    XF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_bf16x3<RH, KPAR>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL((k_m2l_bf16x3<RH, KPAR>), grid, dim3(kGemmThreads), lds, c.s0, g.PP, g.R, RTB, g.K32, KPb, op,
                       tfirst, tcount, cp ? cp->start : tbase, cp ? cp->cnt : tU, sorted, c.tile_keys.p,
                       c.tile_mat.p, ophi, oplo, mhi, mlo, (c.learned && (!cp || cp->variable)) ? c.tile_p.p : nullptr, kr,
                       int64_t(c.t.ncells), ksplit > 1 ? Lpart : nullptr, Lout);
  } else {
    const size_t lds = gemm_lds(g, false, RTB, kr, TN);
    XF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_f32<RH, KPAR>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL((k_m2l_f32<RH, KPAR>), grid, dim3(kGemmThreads), lds, c.s0, g.PP, g.R, RTB, g.K4, KP, tfirst,
                       tcount, tbase, tU, sorted, c.tile_keys.p, c.tile_mat.p, c.ops32.p, sc.mt32.p,
                       c.learned ? c.tile_p.p : nullptr, kr, int64_t(c.t.ncells), ksplit > 1 ? Lpart : nullptr,
                       Lout);
  }
  XF_LAUNCH();
}

struct GemmShape {
  int rtb, kr, ksplit;
  int sk = 0;  // stream-K: one persistent workgroup per CU over the tiles' key slots (BF16x3)
};

static int cu_count(int device) {
  int n = 0;
  XF_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device));
  return std::max(n, 1);
}

// Stream-K launch: all rows per workgroup (R <= kGemmMaxRT), then the
// fixed-order combine of each tile's segments.
template <int RH, int KPAR>
static void launch_gemm_sk(Ctx &c, const Ctx &sc, const Geom &g, int ntiles, int KPb, int kr, double *Lout,
                           const ClassPass *cp) {
  constexpr int TN = tile_width<KPAR>();
  const int ts = c.m2l_tstride;
  const int *tfirst = c.tiles.p, *tcount = c.tiles.p + ts, *tbase = c.tiles.p + 2 * ts, *tU = c.tiles.p + 3 * ts;
  const int *sorted = c.tiles.p + 4 * ts;
  const int G = cu_count(c.cfg.device);
  const int64_t S = c.m2l_totalU;
  const Geom go = cp ? geom(cp->p) : g;
  const OrderPass op{go.P, go.R, go.K32, cp ? cp->perm : nullptr};
  c.m2l_part.ensure(size_t(G + ntiles) * TN * g.PP);
  const v8bf *ophi = reinterpret_cast<const v8bf *>(c.ops32.p);
  const size_t per = size_t(g.K32) * g.R * 64 * 8;
  const v8bf *oplo =
      reinterpret_cast<const v8bf *>(reinterpret_cast<const __bf16 *>(c.ops32.p) + per * c.lists.nkeys);
  const __bf16 *mhi = reinterpret_cast<const __bf16 *>(sc.mt32.p);
  const __bf16 *mlo = mhi + bf16_cast_elems(g, sc.t.ncells);
  const size_t lds = gemm_lds(go, true, go.R, kr, TN);
  const int *pre = cp ? cp->pre : tbase, *start = cp ? cp->start : tbase, *cnt = cp ? cp->cnt : tU;
  const int *Sdev = cp ? cp->tot : nullptr;
  XF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_bf16x3_sk<RH, KPAR>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL((k_m2l_bf16x3_sk<RH, KPAR>), dim3(G), dim3(kGemmThreads), lds, c.s0, g.PP, g.R, g.K32, KPb, op, ntiles,
                     S, Sdev, tfirst, tcount, pre, start, cnt, sorted, c.tile_keys.p, c.tile_mat.p, ophi, oplo, mhi,
                     mlo, (c.learned && (!cp || cp->variable)) ? c.tile_p.p : nullptr, kr, int64_t(c.t.ncells), c.m2l_part.p);
  XF_LAUNCH();
  const int rows = std::min(g.PP, 16 * go.R);
  hipLaunchKernelGGL(k_sk_combine, dim3(ceil_div(int64_t(TN) * rows, 256), ntiles), dim3(256), 0, c.s0, g.PP, rows,
                     ntiles, S, Sdev, G, TN, tfirst, tcount, pre, cnt, sorted, c.m2l_part.p, Lout);
  XF_LAUNCH();
}

static void run_gemm(Ctx &c, const Ctx &sc, const Geom &g, int ntiles, bool bf16, int KP, int KPb, GemmShape sh,
                     double *Lout, const ClassPass *cp = nullptr) {
  if (sh.sk) {
    switch (cp ? geom(cp->p).R : g.R) {
#define XF_GEMM_SK(n)                                                  \
  case n:                                                              \
    if (c.m2l_tn == tile_width<1>())                                   \
      launch_gemm_sk<n, 1>(c, sc, g, ntiles, KPb, sh.kr, Lout, cp);   \
    else                                                               \
      launch_gemm_sk<n, 2>(c, sc, g, ntiles, KPb, sh.kr, Lout, cp);   \
    return;
      XF_GEMM_SK(1)
      XF_GEMM_SK(2)
      XF_GEMM_SK(3)
      XF_GEMM_SK(4)
      XF_GEMM_SK(5)
      XF_GEMM_SK(6)
      XF_GEMM_SK(7)
      XF_GEMM_SK(8)
#undef XF_GEMM_SK
      default: fail(EXAFMM_ERR_UNSUPPORTED, "stream-K GEMM needs all rows in one workgroup");
    }
  }
  float *part = nullptr;
  if (sh.ksplit > 1) {
    c.m2l_part.ensure(size_t(sh.ksplit) * c.t.ncells * g.PP);
    part = c.m2l_part.p;
  }
  switch (sh.rtb) {
#define XF_GEMM_CASE(n)                                                                          \
  case n:                                                                                        \
    if (c.m2l_tn == tile_width<1>())                                                             \
      launch_gemm<n, 1>(c, sc, g, ntiles, bf16, KP, KPb, sh.rtb, sh.kr, sh.ksplit, part, Lout, cp); \
    else                                                                                         \
      launch_gemm<n, 2>(c, sc, g, ntiles, bf16, KP, KPb, sh.rtb, sh.kr, sh.ksplit, part, Lout, cp); \
    break;
    XF_GEMM_CASE(1)
    XF_GEMM_CASE(2)
    XF_GEMM_CASE(3)
    XF_GEMM_CASE(4)
    XF_GEMM_CASE(5)
    XF_GEMM_CASE(6)
    XF_GEMM_CASE(7)
    XF_GEMM_CASE(8)
#undef XF_GEMM_CASE
    default: fail(EXAFMM_ERR_UNSUPPORTED, "expansion order");
  }
  if (sh.ksplit > 1) {
    const int64_t n = int64_t(c.m2l_ncell_tiled) * g.PP;
    hipLaunchKernelGGL(k_combine_slices, dim3(ceil_div(n, 256)), dim3(256), 0, c.s0, g.PP, int64_t(c.t.ncells),
                       sh.ksplit, c.tiles.p + 4 * c.m2l_tstride, c.m2l_ncell_tiled, part, Lout);
    XF_LAUNCH();
  }
}

// Autotuner of the GEMM shape: rows of the operator per workgroup (the
// rows of a tile may be split over several workgroups: more workgroups
// when tiles are few, more re-reads of the source multipoles), keys per
// LDS round (more keys per barrier hide the operator-copy latency of short,
// low-order keys) and key slices per tile (a tile's keys split over
// several workgroups whose FP32 partial locals a second kernel adds in a
// fixed order: more parallelism when tiles are few and keys many).  The first time a (order, precision, order policy, tile
// count class) is seen, every candidate (GEMM and, with key slices, the
// combining kernel) runs twice into a scratch copy of the locals, alone on
// the device, timed with HIP events; the fastest run's shape is kept for
// the context's lifetime.
static GemmShape gemm_shape(Ctx &c, const Ctx &sc, const Geom &g, int ntiles, bool bf16, int KP, int KPb,
                            const ClassPass *cp = nullptr) {
  int tclass = 0;
  while ((1 << tclass) < ntiles) tclass++;
  // a learned-order launch of one order is tuned per order (its rows, k-steps and LDS)
  const Geom go = cp ? geom(cp->p) : g;
  const long long key = ((long long)g.P << 24) | ((long long)bf16 << 20) | ((long long)c.learned << 16) |
                        ((long long)(cp ? cp->p : 0) << 32) | ((long long)(c.m2l_tn == tile_width<1>()) << 40) |
                        tclass;
  auto it = c.gemm_tuned.find(key);
  if (it != c.gemm_tuned.end())
    return GemmShape{int(it->second & 255), int((it->second >> 8) & 255), int((it->second >> 16) & 255),
                     int(it->second >> 24)};
  std::vector<GemmShape> cand;
  // low orders: short keys, so more of them per LDS round
  const int krmax = go.P <= 10 ? 5 : 3;
  for (int kr = 1; kr <= krmax; kr++) {
    const int rmax = rows_per_block(go, bf16, c.cfg.device, kr, c.m2l_tn);
    if (rmax < 1) continue;
    int last = -1;
    for (int nb = ceil_div(go.R, rmax); nb <= ceil_div(go.R, rmax) + 2; nb++) {
      const int r = ceil_div(go.R, nb);
      if (r != last && r <= rmax) {
        cand.push_back(GemmShape{r, kr, 1});
        // key slices only where the workgroups would not fill the device
        const int wg = ntiles * nb;
        for (const int ks : {2, 3, 4, 6, 8})
          if (wg * ks <= 4096 && c.m2l_totalU >= 8 * ks * ntiles) cand.push_back(GemmShape{r, kr, ks});
      }
      last = r;
    }
  }
  // stream-K (BF16x3, all rows per workgroup)
  if (bf16 && go.R <= kGemmMaxRT && std::getenv("EXAFMM_GEMM_NO_SK") == nullptr)
    for (int kr = 1; kr <= 3; kr++)
      if (rows_per_block(go, bf16, c.cfg.device, kr, c.m2l_tn) >= go.R) cand.push_back(GemmShape{go.R, kr, 1, 1});
  GemmShape best = cand[0];
  if (std::getenv("EXAFMM_GEMM_SK_ONLY"))  // the stream-K candidates alone (tests, A/B timing)
    for (const GemmShape &sh : cand)
      if (sh.sk) {
        best = sh;
        break;
      }
  if (c.capturing) fail(EXAFMM_ERR_STATE, "GEMM shape not tuned before recording a graph");
  if (std::getenv("EXAFMM_GEMM_SK_ONLY")) cand.erase(std::remove_if(cand.begin(), cand.end(), [](const GemmShape &x) { return !x.sk; }), cand.end());
  if (cand.empty()) fail(EXAFMM_ERR_UNSUPPORTED, "no GEMM shape");
  if (cand.size() > 1 && c.cfg.autotune) {
    hipStream_t s = c.s0;
    XF_HIP(hipStreamSynchronize(c.s1));  // tune alone on the device
    c.L_tune.ensure(size_t(c.t.ncells) * c.P * c.P);
    size_t need = 1;
    for (const GemmShape &sh : cand)
      need = std::max(need, sh.sk ? size_t(cu_count(c.cfg.device) + ntiles) * c.m2l_tn * g.PP
                                  : size_t(sh.ksplit > 1 ? sh.ksplit : 0) * c.t.ncells * g.PP);
    c.m2l_part.ensure(need);  // no allocation inside the timings
    hipEvent_t e0, e1;
    XF_HIP(hipEventCreate(&e0));
    XF_HIP(hipEventCreate(&e1));
    float bt = 1e30f;
    for (const GemmShape &sh : cand) {
      for (int rep = 0; rep < 2; rep++) {  // best of two runs per candidate
        XF_HIP(hipEventRecord(e0, s));
        run_gemm(c, sc, g, ntiles, bf16, KP, KPb, sh, c.L_tune.p, cp);
        XF_HIP(hipEventRecord(e1, s));
        XF_HIP(hipEventSynchronize(e1));
        float ms = 0;
        XF_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < bt) {
          bt = ms;
          best = sh;
        }
      }
    }
    XF_HIP(hipEventDestroy(e0));
    XF_HIP(hipEventDestroy(e1));
  }
  c.gemm_tuned[key] = best.rtb | (best.kr << 8) | (best.ksplit << 16) | (best.sk << 24);
  return best;
}

static bool ops_cached(const Ctx &c, const std::vector<unsigned long long> &keys, int fillP, int bf16) {
  const bool periodic = c.cfg.images > 0;
  return c.ops_P == c.P && c.ops_fill == fillP && c.ops_bf16 == bf16 && c.ops_keys == keys &&
         (!periodic || (c.ops_R0 == c.t.R0 && c.ops_cycle == c.cfg.cycle));
}

static void ops_remember(Ctx &c, std::vector<unsigned long long> keys, int fillP, int bf16) {
  c.ops_keys = std::move(keys);
  c.ops_P = c.P;
  c.ops_fill = fillP;
  c.ops_bf16 = bf16;
  c.ops_R0 = c.t.R0;
  c.ops_cycle = c.cfg.cycle;
}

// Tiles, key lists and operators: depend only on the tree and the lists,
// so they can be built on a side stream while the upward pass runs.
void m2l_prepare(Ctx &c, hipStream_t s) {
  c.m2l_ready = false;
  c.m2l_ntiles = 0;
  const Geom g = geom(c.P);
  const bool bf16 = c.cfg.m2l_precision == EXAFMM_PREC_BF16X3;
  const int fillP = c.cfg.m2l_truncation == EXAFMM_M2L_FULL ? 2 * c.P - 1 : c.P;
  const int nc = int(c.t.ncells);
  const int nkeys = c.lists.nkeys;
  if (nkeys > (1 << 16)) fail(EXAFMM_ERR_CAPACITY, "more than 65536 distinct M2L translations");

  // 1. tiles: cells with pairs grouped by (level, octant), then key set
  c.t.key.ensure(nc);
  c.t.key2.ensure(nc);
  c.t.val.ensure(nc);
  c.t.val2.ensure(nc);
  // cells sorted by (level, octant, min-hash of key set), BFS order within
  // ties: on adaptive trees this raises the fraction of non-empty GEMM slots
  // (measured 0.50 -> 0.58 on a 2^22-body Plummer set); on uniform trees,
  // where BFS neighbours already share key sets, it changes little.
  const int minhash = 1;
  hipLaunchKernelGGL(k_group_keys, dim3(ceil_div(nc, 256 / kWave)), dim3(256), 0, s, nc, c.t.cx.p, c.t.cparent.p,
                     c.t.clevel.p, c.lists.m2l_beg.p, c.lists.m2l_cnt.p, c.lists.m2l.p, minhash,
                     m2l_fp64_levels(c), c.t.key.p, c.t.val.p);
  XF_LAUNCH();
  unsigned long long *k1 = c.t.key.p, *k2 = c.t.key2.p;
  int *v1 = c.t.val.p, *v2 = c.t.val2.p;
  radix_sort_pairs(k1, v1, k2, v2, nc, minhash ? 0 : 16, 24, c.m2l_rs_hist, s);
  c.m2l_misc.ensure(512);
  XF_HIP(hipMemsetAsync(c.m2l_misc.p, 0, 256 * sizeof(int), s));
  hipLaunchKernelGGL(k_group_hist, dim3(std::min(ceil_div(nc, 256), 64)), dim3(256), 0, s, nc, k1, c.m2l_misc.p);
  XF_LAUNCH();
  // tiles of at most tn cells per group, laid out on the device (no host
  // round trip before the per-tile counts); arrays strided by the bound
  // ntmax: [first | count | base | U] x ntmax, then the sorted cells
  // Tile width: 512 targets halves the operator staging per target where
  // the tiles fill (a free-space tree's octant groups); periodic trees (image
  // keys) and forests (many small problems) have more varied key sets, whose
  // larger per-tile key unions favour 256.
  c.m2l_tn = (c.cfg.images > 0 || c.forest_n > 0) ? tile_width<2>() : tile_width<1>();
  const int tn = c.m2l_tn;
  const int ntmax = ceil_div(nc, tn) + 255;
  c.tiles.ensure(size_t(4) * ntmax + nc);
  int *d_first = c.tiles.p, *d_count = c.tiles.p + ntmax, *d_base = c.tiles.p + 2 * ntmax,
      *d_U = c.tiles.p + 3 * ntmax, *d_sorted = c.tiles.p + 4 * ntmax;
  int *d_plan = c.m2l_misc.p + 257;  // [ntiles, cells in tiles]
  hipLaunchKernelGGL(k_tile_plan, dim3(1), dim3(256), 0, s, c.m2l_misc.p, tn, d_first, d_count, d_plan);
  XF_LAUNCH();
  XF_HIP(hipMemcpyAsync(d_sorted, v1, sizeof(int) * nc, hipMemcpyDeviceToDevice, s));
  const int nwords = (nkeys + 31) / 32;
  c.tile_bits.ensure(size_t(ntmax) * std::max(nwords, 1));
  hipLaunchKernelGGL(k_tile_count, dim3(ntmax), dim3(kTileCountThreads), sizeof(unsigned) * nwords, s, d_first, d_count,
                     d_sorted, c.lists.m2l_beg.p, c.lists.m2l_cnt.p, c.lists.m2l.p, nkeys, d_plan, d_U,
                     c.tile_bits.p);
  XF_LAUNCH();
  scan_exclusive_dev(d_U, d_base, d_plan, ntmax, c.m2l_misc.p + 256, c.m2l_scan, s);
  // one host round trip: tile count, slots, cells in tiles, and the key set
  c.pin_i.ensure(kPinInts);
  c.pin_keys.ensure(std::max(nkeys, 1));
  XF_HIP(hipMemcpyAsync(c.pin_i.p + kPinPlan, c.m2l_misc.p + 256, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
  if (nkeys)
    XF_HIP(hipMemcpyAsync(c.pin_keys.p, c.lists.keys.p, sizeof(unsigned long long) * nkeys, hipMemcpyDeviceToHost,
                          s));
  stream_wait(s);
  const int *plan = c.pin_i.p + kPinPlan;
  std::vector<unsigned long long> keys(c.pin_keys.p, c.pin_keys.p + nkeys);
  const int totalU = plan[0], ntiles = plan[1], off = plan[2];
  c.m2l_tstride = ntmax;
  if (ntiles == 0) {
    c.m2l_ready = true;
    return;
  }
  c.tile_keys.ensure(std::max(totalU, 1));
  c.tile_mat.ensure(size_t(std::max(totalU, 1)) * tn);
  hipLaunchKernelGGL(k_tile_fill, dim3(ntiles, tn / 64), dim3(256), sizeof(unsigned) * 2 * nwords, s, d_first,
                     d_count, d_sorted, c.lists.m2l_beg.p, c.lists.m2l_cnt.p, c.lists.m2l.p, nkeys, d_base,
                     c.tile_bits.p, c.tile_keys.p, tn, c.tile_mat.p);
  XF_LAUNCH();

  // 2. operators (cached by key set)
  const double cyc = c.cfg.images > 0 ? c.cfg.cycle : 0.0;
  if (bf16) {
    const size_t per = size_t(g.K32) * g.R * 64 * 8;
    if (!ops_cached(c, keys, fillP, 1)) {
      c.ops32.ensure((per * nkeys * 2 + 1) / 2 + 1);
      __bf16 *ophi = reinterpret_cast<__bf16 *>(c.ops32.p);
      hipLaunchKernelGGL(k_ops_gen, dim3(nkeys), dim3(256), 0, s, c.P, fillP, g.R, g.K4, g.K32, c.t.R0, cyc,
                         c.lists.keys.p, (float *)nullptr, ophi, ophi + per * nkeys);
      XF_LAUNCH();
      ops_remember(c, std::move(keys), fillP, 1);
    }
  } else {
    const size_t per = size_t(g.K4) * g.R * 64;
    if (!ops_cached(c, keys, fillP, 0)) {
      c.ops32.ensure(per * nkeys);
      hipLaunchKernelGGL(k_ops_gen, dim3(nkeys), dim3(256), 0, s, c.P, fillP, g.R, g.K4, g.K32, c.t.R0, cyc,
                         c.lists.keys.p, c.ops32.p, (__bf16 *)nullptr, (__bf16 *)nullptr);
      XF_LAUNCH();
      ops_remember(c, std::move(keys), fillP, 0);
    }
  }
  c.m2l_ntiles = ntiles;
  c.m2l_ncell_tiled = off;
  c.m2l_totalU = totalU;
  c.m2l_ready = true;
}

// Cast the multipoles and run the keyed GEMM (after m2l_prepare, on s0).
void m2l_fast(Ctx &tc, const Ctx &sc) {
  if (&tc != &sc) fail(EXAFMM_ERR_UNSUPPORTED, "keyed M2L needs source == target tree");
  Ctx &c = tc;
  if (!c.m2l_ready) {
    if (c.capturing) fail(EXAFMM_ERR_STATE, "keyed M2L tiles missing while recording a graph");
    m2l_prepare(c, c.s0);
  }
  hipStream_t s = c.s0;
  const Geom g = geom(c.P);
  const bool bf16 = c.cfg.m2l_precision == EXAFMM_PREC_BF16X3;
  const int nc = int(c.t.ncells);
  const int ntiles = c.m2l_ntiles;
  c.m2l_ntiles_last = ntiles;
  if (!c.graph_locks) c.m2l_ready = false;  // tiles belong to this horizontal pass (or to the graph)
  if (ntiles == 0) return;
  // BF16 rows hold P^2 coefficients padded to a multiple of 8 (one k-group
  // load), not to the MFMA's K of 32 per step: the k-groups past the row
  // read the zero row, which stays in cache
  const int KP = g.K4 * 4, KPb = bf16_row(g);
  if (bf16) {
    const size_t rows = bf16_cast_elems(g, nc);  // per part
    c.mt32.ensure(rows + 1);                     // two bf16 arrays in a float buffer
    __bf16 *mhi = reinterpret_cast<__bf16 *>(c.mt32.p);
    __bf16 *mlo = mhi + rows;
    hipLaunchKernelGGL(k_mt, dim3(ceil_div(int64_t(rows), 256)), dim3(256), 0, s, c.P * c.P, KPb, nc, kBf16Tail,
                       c.M.p, (float *)nullptr, mhi, mlo);
    XF_LAUNCH();
  } else {
    c.mt32.ensure(size_t(nc + 1) * KP);
    hipLaunchKernelGGL(k_mt, dim3(ceil_div(int64_t(nc + 1) * KP, 256)), dim3(256), 0, s, c.P * c.P, KP, nc, 0,
                       c.M.p, c.mt32.p, (__bf16 *)nullptr, (__bf16 *)nullptr);
    XF_LAUNCH();
  }

  // per-key orders of a learned-order evaluation (need the multipoles)
  // (a forest has no common root to scale the tolerance by: fixed order)
  c.learned = c.cfg.order_policy == EXAFMM_ORDER_LEARNED && c.have_model && c.forest_n == 0;
  if (c.learned)
    order_select(c, c.tiles.p, c.tiles.p + c.m2l_tstride, c.tiles.p + 2 * c.m2l_tstride,
                 c.tiles.p + 3 * c.m2l_tstride, c.tiles.p + 4 * c.m2l_tstride, ntiles, c.m2l_totalU, c.m2l_tn, s);

  // 3. GEMM over tiles, shape from the autotuner
  if (c.learned && bf16) {
    // one launch per order class over that class's slots of every tile
    // (grouped by k_order_partition), each with its own tuned shape: a
    // slot's cost then follows its order, and the launches add into the
    // locals in class sequence.  Classes: orders up to 8, then pairs (9-10,
    // 11-12, ...): few launches (each adds a partial-locals pass), each
    // class's LDS stage and row tiles sized to its highest order
    OrderClasses oc{};
    for (int p = 0; p <= kMaxP; p++) {
      const int q = std::min(std::max(p, kMlpPmin), c.P);
      const int k = q <= 8 ? 0 : (q - 7) / 2;
      oc.of[p] = (unsigned char)k;
      oc.hi[k] = (unsigned char)std::max<int>(oc.hi[k], q);
      oc.n = std::max(oc.n, k + 1);
    }
    const int ncls = oc.n;
    const int ts = c.m2l_tstride;
    c.tile_perm.ensure(std::max(c.m2l_totalU, 1));
    // tile_cls: [start | count | prefix] x order x tile stride, then the totals
    c.tile_cls.ensure(size_t(3) * ncls * ts + ncls);
    hipLaunchKernelGGL(k_order_partition, dim3(ceil_div(ntiles, 4)), dim3(256), 0, s, ntiles, ts, c.tiles.p + 2 * ts,
                       c.tiles.p + 3 * ts, c.tile_p.p, oc, c.tile_perm.p, c.tile_cls.p);
    XF_LAUNCH();
    hipLaunchKernelGGL(k_order_prefix, dim3(1), dim3(64 * ncls), 0, s, ntiles, ts, ncls, c.tile_cls.p);
    XF_LAUNCH();
    if (!c.capturing) XF_HIP(hipEventRecord(c.ev[13], s));
    for (int o = ncls - 1; o >= 0; o--) {
      const int *cls = c.tile_cls.p;
      int lo = kMaxP;
      for (int p = kMlpPmin; p <= c.P; p++)
        if (oc.of[p] == o) lo = std::min(lo, p);
      const ClassPass cp{oc.hi[o], lo < oc.hi[o], c.tile_perm.p, cls + size_t(o) * ts, cls + size_t(ncls + o) * ts,
                         cls + size_t(2 * ncls + o) * ts, cls + size_t(3 * ncls) * ts + o};
      const GemmShape sh = gemm_shape(c, sc, g, ntiles, bf16, KP, KPb, &cp);
      run_gemm(c, sc, g, ntiles, bf16, KP, KPb, sh, c.L.p, &cp);
    }
    return;
  }
  const GemmShape sh = gemm_shape(c, sc, g, ntiles, bf16, KP, KPb);
  if (!c.capturing) XF_HIP(hipEventRecord(c.ev[13], s));
  run_gemm(c, sc, g, ntiles, bf16, KP, KPb, sh, c.L.p);
}


// ---------------------------------------------------------------- grouped (batch)
struct GroupCast {
  const double *M;
  __bf16 *mhi, *mlo;
  int64_t ncells;
  int PP, KP;
  int64_t elems;  // elements of each cast array: rows of KP, the zero row, the zero tail
};

// BF16 hi/lo cast of the multipoles of every problem of one order (row
// ncells of each problem is the zero row).
__global__ void __launch_bounds__(256) k_mt_grouped(const GroupCast *__restrict__ g, const int *__restrict__ start,
                                                    int nprob) {
  const int b = blockIdx.x, i = group_of(start, nprob, b);
  const GroupCast q = g[i];
  const int64_t e = int64_t(b - start[i]) * 256 + threadIdx.x;
  if (e >= q.elems) return;
  const int64_t c = e / q.KP;
  const int r = int(e % q.KP);
  const float v = (c < q.ncells && r < q.PP) ? float(q.M[c * q.PP + r]) : 0.0f;
  const __bf16 h = __bf16(v);
  q.mhi[e] = h;
  q.mlo[e] = __bf16(v - float(h));
}

template <int RH, int KPAR>
static void group_attr(size_t lds) {
  XF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_bf16x3_grouped<RH, KPAR>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
}

// Descriptor tables of one order's problems (their tiles and operators from
// m2l_prepare, their cast buffers sized by an earlier evaluation) for one
// GEMM shape (rows per workgroup, keys per LDS round), the same for every
// problem of the launch.
static void group_tables(const std::vector<Ctx *> &probs, GemmGroupPlan &plan, int rtb, int kr) {
  plan.nprob = 0;
  plan.nblocks = plan.cast_blocks = 0;
  if (probs.empty()) return;
  const int P = probs[0]->P;
  const Geom g = geom(P);
  const int KPb = bf16_row(g);  // the cast layout of m2l_fast: rows of KPb, zero row, zero tail
  plan.P = P;
  plan.rtb = rtb;
  plan.kr = kr;
  const int nrb = ceil_div(g.R, plan.rtb);
  plan.tn = probs[0]->m2l_tn;
  plan.lds = gemm_lds(g, true, plan.rtb, kr, plan.tn);
  std::vector<GroupGemm> d;
  std::vector<GroupCast> cd;
  std::vector<int> st, cst;
  int64_t acc = 0, cacc = 0;
  for (Ctx *cp : probs) {
    Ctx &c = *cp;
    if (c.P != P || c.cfg.m2l_precision != EXAFMM_PREC_BF16X3) fail(EXAFMM_ERR_ARG, "mixed orders in a GEMM group");
    if (c.m2l_tn != plan.tn) fail(EXAFMM_ERR_ARG, "mixed tile widths in a GEMM group");
    if (!c.m2l_ready) fail(EXAFMM_ERR_STATE, "keyed M2L tiles missing");
    const int nt = c.m2l_ntiles;
    if (nt == 0) continue;
    const int64_t nc = c.t.ncells;
    const size_t per = size_t(g.K32) * g.R * 64 * 8;
    const __bf16 *ophi = reinterpret_cast<const __bf16 *>(c.ops32.p);
    const size_t elems = bf16_cast_elems(g, nc);
    c.mt32.ensure(elems + 1);  // two bf16 arrays in a float buffer (set-up time, not under capture)
    __bf16 *mhi = reinterpret_cast<__bf16 *>(c.mt32.p);
    GroupGemm q;
    const int ts = c.m2l_tstride;
    q.tfirst = c.tiles.p;
    q.tcount = c.tiles.p + ts;
    q.tbase = c.tiles.p + 2 * ts;
    q.tU = c.tiles.p + 3 * ts;
    q.sorted = c.tiles.p + 4 * ts;
    q.tkeys = c.tile_keys.p;
    q.tmat = c.tile_mat.p;
    q.ophi = reinterpret_cast<const v8bf *>(ophi);
    q.oplo = reinterpret_cast<const v8bf *>(ophi + per * c.lists.nkeys);
    q.mhi = mhi;
    q.mlo = mhi + elems;
    q.L = c.L.p;
    q.ncells = nc;
    q.ntiles = nt;
    d.push_back(q);
    st.push_back(int(acc));
    acc += int64_t(nt) * nrb;
    GroupCast cq{c.M.p, mhi, mhi + elems, nc, g.PP, KPb, int64_t(elems)};
    cd.push_back(cq);
    cst.push_back(int(cacc));
    cacc += ceil_div(int64_t(elems), 256);
  }
  if (acc > (int64_t(1) << 31) - 1 || cacc > (int64_t(1) << 31) - 1) fail(EXAFMM_ERR_CAPACITY, "batch too large");
  plan.nprob = int(d.size());
  plan.nblocks = int(acc);
  plan.cast_blocks = int(cacc);
  if (!plan.nprob) return;
  st.push_back(int(acc));
  cst.push_back(int(cacc));
  plan.desc.ensure(sizeof(GroupGemm) * d.size());
  plan.cast_desc.ensure(sizeof(GroupCast) * cd.size());
  plan.start.ensure(st.size());
  plan.cast_start.ensure(cst.size());
  XF_HIP(hipMemcpy(plan.desc.p, d.data(), sizeof(GroupGemm) * d.size(), hipMemcpyHostToDevice));
  XF_HIP(hipMemcpy(plan.cast_desc.p, cd.data(), sizeof(GroupCast) * cd.size(), hipMemcpyHostToDevice));
  XF_HIP(hipMemcpy(plan.start.p, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
  XF_HIP(hipMemcpy(plan.cast_start.p, cst.data(), sizeof(int) * cst.size(), hipMemcpyHostToDevice));
  switch (plan.rtb) {
    case 1: plan.tn == tile_width<1>() ? group_attr<1, 1>(plan.lds) : group_attr<1, 2>(plan.lds); break;
    case 2: plan.tn == tile_width<1>() ? group_attr<2, 1>(plan.lds) : group_attr<2, 2>(plan.lds); break;
    case 3: plan.tn == tile_width<1>() ? group_attr<3, 1>(plan.lds) : group_attr<3, 2>(plan.lds); break;
    case 4: plan.tn == tile_width<1>() ? group_attr<4, 1>(plan.lds) : group_attr<4, 2>(plan.lds); break;
    case 5: plan.tn == tile_width<1>() ? group_attr<5, 1>(plan.lds) : group_attr<5, 2>(plan.lds); break;
    case 6: plan.tn == tile_width<1>() ? group_attr<6, 1>(plan.lds) : group_attr<6, 2>(plan.lds); break;
    case 7: plan.tn == tile_width<1>() ? group_attr<7, 1>(plan.lds) : group_attr<7, 2>(plan.lds); break;
    case 8: plan.tn == tile_width<1>() ? group_attr<8, 1>(plan.lds) : group_attr<8, 2>(plan.lds); break;
    default: fail(EXAFMM_ERR_UNSUPPORTED, "expansion order");
  }
}

// Autotuner of a grouped GEMM: every (rows per workgroup, keys per LDS
// round) candidate runs on the batch itself (alone on the device, after one
// untimed launch), timed with HIP events; the fastest shape is kept.  The
// runs add into the locals, which the graph's prologue clears.
void m2l_group_setup(const std::vector<Ctx *> &probs, GemmGroupPlan &plan, hipStream_t s, bool tune) {
  if (probs.empty()) return;
  const Geom g = geom(probs[0]->P);
  const int dev = probs[0]->cfg.device;
  std::vector<std::pair<int, int>> cand;  // (rtb, kr)
  for (int kr = 1; kr <= 3; kr++) {
    const int rmax = rows_per_block(g, true, dev, kr, probs[0]->m2l_tn);
    if (rmax < 1) continue;
    int last = -1;
    for (int nb = ceil_div(g.R, rmax); nb <= ceil_div(g.R, rmax) + 1; nb++) {
      const int r = ceil_div(g.R, nb);
      if (r != last && r <= rmax) cand.push_back({r, kr});
      last = r;
    }
  }
  if (cand.empty()) fail(EXAFMM_ERR_UNSUPPORTED, "expansion order too large for the LDS operator stage");
  std::pair<int, int> best = cand[0];
  if (tune && cand.size() > 1) {
    hipEvent_t e0, e1;
    XF_HIP(hipEventCreate(&e0));
    XF_HIP(hipEventCreate(&e1));
    float bt = 1e30f;
    for (const auto &c : cand) {
      group_tables(probs, plan, c.first, c.second);
      m2l_group_launch(plan, s);  // untimed
      XF_HIP(hipEventRecord(e0, s));
      m2l_group_launch(plan, s);
      XF_HIP(hipEventRecord(e1, s));
      XF_HIP(hipEventSynchronize(e1));
      float ms = 0;
      XF_HIP(hipEventElapsedTime(&ms, e0, e1));
      plan.tuned.push_back({c.first, c.second, ms});
      if (ms < bt) {
        bt = ms;
        best = c;
      }
    }
    XF_HIP(hipEventDestroy(e0));
    XF_HIP(hipEventDestroy(e1));
  }
  group_tables(probs, plan, best.first, best.second);
}

void m2l_group_launch(const GemmGroupPlan &plan, hipStream_t s) {
  if (!plan.nprob) return;
  const Geom g = geom(plan.P);
  const GroupCast *cd = reinterpret_cast<const GroupCast *>(plan.cast_desc.p);
  hipLaunchKernelGGL(k_mt_grouped, dim3(plan.cast_blocks), dim3(256), 0, s, cd, plan.cast_start.p, plan.nprob);
  XF_LAUNCH();
  const GroupGemm *d = reinterpret_cast<const GroupGemm *>(plan.desc.p);
  const int KPb = bf16_row(g);
  switch (plan.rtb) {
#define XF_GROUP_CASE(n)                                                                                         \
  case n:                                                                                                         \
    if (plan.tn == tile_width<1>())                                                                              \
      hipLaunchKernelGGL((k_m2l_bf16x3_grouped<n, 1>), dim3(plan.nblocks), dim3(kGemmThreads), plan.lds, s, d,    \
                         plan.start.p, plan.nprob, g.PP, g.R, plan.rtb, g.K32, KPb, g.P, plan.kr);                \
    else                                                                                                          \
      hipLaunchKernelGGL((k_m2l_bf16x3_grouped<n, 2>), dim3(plan.nblocks), dim3(kGemmThreads), plan.lds, s, d,    \
                         plan.start.p, plan.nprob, g.PP, g.R, plan.rtb, g.K32, KPb, g.P, plan.kr);                \
    break;
    XF_GROUP_CASE(1)
    XF_GROUP_CASE(2)
    XF_GROUP_CASE(3)
    XF_GROUP_CASE(4)
    XF_GROUP_CASE(5)
    XF_GROUP_CASE(6)
    XF_GROUP_CASE(7)
    XF_GROUP_CASE(8)
#undef XF_GROUP_CASE
    default: fail(EXAFMM_ERR_UNSUPPORTED, "expansion order");
  }
  XF_LAUNCH();
}

}  // namespace xf
