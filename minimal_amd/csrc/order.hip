// order.hip -- learned per-pair expansion order for the keyed M2L.
//
// No counterpart in the reference (which uses one order P everywhere,
// 3d/kernel.h:233-260).  A small MLP maps four features of an M2L pair
//   f0 = log2(R_source / R_target)
//   f1 = log2(|d| / (R_target + R_source))
//   f2 = level of the target / 8
//   f3 = log10 of the share of the source multipole norm in degrees >= 2
//        (clamped to [-6, 0]; scaled multipoles, so geometry-free)
// to the predicted log10 truncation error of that pair at orders 4..14,
// normalised by (sum |q_source|) / |d|.  The order of a pair is the
// smallest p whose prediction is below the tolerance (14 otherwise, and
// never above the context's order P); a (tile, key) of the keyed GEMM runs
// at the largest order of its pairs.
//
// Network: 4 -> 32 -> 32 -> 11, ReLU, BF16 weights, FP32 accumulation.
// Training data come from the method itself: for sampled pairs of a plain
// evaluation at order 14, the M2L + L2P potential at every order 4..14 is
// compared in FP64 with the direct sum between the two cells.  Training
// runs on the device (forward, backward and Adam kernels below).
#include <algorithm>
#include <cmath>
#include <random>

#include "fmm_internal.h"
#include "fmm_ops.h"

namespace xf {
using namespace xfd;

// parameter layout (floats)
constexpr int oW1 = 0, oB1 = oW1 + kMlpH * kMlpIn, oW2 = oB1 + kMlpH, oB2 = oW2 + kMlpH * kMlpH, oW3 = oB2 + kMlpH,
              oB3 = oW3 + kMlpOut * kMlpH;
static_assert(oB3 + kMlpOut == kMlpParams, "parameter count");

__device__ inline float bf16_round(float v) { return float(__bf16(v)); }

// Forward pass of one sample with the (BF16-valued) weights W.
__device__ inline void mlp_forward(const float *W, const float (&x)[kMlpIn], float (&h1)[kMlpH], float (&h2)[kMlpH],
                                   float (&y)[kMlpOut]) {
#pragma unroll 4
  for (int i = 0; i < kMlpH; i++) {
    float a = W[oB1 + i];
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) a += W[oW1 + i * kMlpIn + k] * x[k];
    h1[i] = fmaxf(a, 0.0f);
  }
  for (int i = 0; i < kMlpH; i++) {
    float a = W[oB2 + i];
#pragma unroll
    for (int k = 0; k < kMlpH; k++) a += W[oW2 + i * kMlpH + k] * h1[k];
    h2[i] = fmaxf(a, 0.0f);
  }
  for (int o = 0; o < kMlpOut; o++) {
    float a = W[oB3 + o];
#pragma unroll
    for (int k = 0; k < kMlpH; k++) a += W[oW3 + o * kMlpH + k] * h2[k];
    y[o] = a;
  }
}

// weights (FP32 master) -> BF16-valued FP32 copy in LDS
__device__ inline void load_weights(const float *__restrict__ w, float *W) {
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) W[i] = bf16_round(w[i]);
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ w, const float *__restrict__ X,
                                                     int64_t n, float *__restrict__ Y) {
  __shared__ float W[kMlpParams];
  load_weights(w, W);
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    float x[kMlpIn], h1[kMlpH], h2[kMlpH], y[kMlpOut];
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) x[k] = X[i * kMlpIn + k];
    mlp_forward(W, x, h1, h2, y);
#pragma unroll
    for (int o = 0; o < kMlpOut; o++) Y[i * kMlpOut + o] = y[o];
  }
}

// Loss = sum over samples and valid outputs of (y - t)^2 / count, with
// t = NaN marking an output without data.  Every gradient entry is a sum
// over samples of a product of two per-sample quantities (weight grads:
// the backpropagated error of a unit times an input of its layer; biases:
// the error times 1).  A workgroup takes chunks of kGradChunk samples
// (chunks b, b + gridDim.x, ...): one thread per sample runs the forward
// and backward pass and stores those quantities in LDS; then each thread
// adds, for the parameters it owns, the chunk's products in sample order.
// No atomics: the gradient is the same bit for bit on every run.
constexpr int kGradChunk = 128;
// per-sample row: x | h1 | h2 | dy | g2 | g1 | 1
constexpr int sX = 0, sH1 = sX + kMlpIn, sH2 = sH1 + kMlpH, sDY = sH2 + kMlpH, sG2 = sDY + kMlpOut,
              sG1 = sG2 + kMlpH, sOne = sG1 + kMlpH, kGradRow = sOne + 1;
constexpr int kGradOwn = (kMlpParams + 255) / 256;  // parameters per thread

// LDS columns (a, b) whose per-sample product sums to parameter q's gradient.
__device__ inline void grad_cols(int q, int &a, int &b) {
  if (q < oB1) {
    a = sG1 + (q - oW1) / kMlpIn;
    b = sX + (q - oW1) % kMlpIn;
  } else if (q < oW2) {
    a = sG1 + (q - oB1);
    b = sOne;
  } else if (q < oB2) {
    a = sG2 + (q - oW2) / kMlpH;
    b = sH1 + (q - oW2) % kMlpH;
  } else if (q < oW3) {
    a = sG2 + (q - oB2);
    b = sOne;
  } else if (q < oB3) {
    a = sDY + (q - oW3) / kMlpH;
    b = sH2 + (q - oW3) % kMlpH;
  } else {
    a = sDY + (q - oB3);
    b = sOne;
  }
}

__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ w, const float *__restrict__ X,
                                                  const float *__restrict__ T, int64_t n, float inv_count,
                                                  float *__restrict__ gpart, float *__restrict__ lpart) {
  __shared__ float W[kMlpParams];
  __shared__ float S[kGradChunk * kGradRow];
  __shared__ float Lsum[256];
  load_weights(w, W);
  float g[kGradOwn];
  int ca[kGradOwn], cb[kGradOwn];
#pragma unroll
  for (int j = 0; j < kGradOwn; j++) {
    g[j] = 0.0f;
    const int q = threadIdx.x + 256 * j;
    ca[j] = cb[j] = sOne;
    if (q < kMlpParams) grad_cols(q, ca[j], cb[j]);
  }
  float lsum = 0.0f;
  for (int64_t c0 = int64_t(blockIdx.x) * kGradChunk; c0 < n; c0 += int64_t(gridDim.x) * kGradChunk) {
    const int ns = int(min<int64_t>(kGradChunk, n - c0));
    if (int(threadIdx.x) < kGradChunk) {
      float *R = S + threadIdx.x * kGradRow;
      if (int(threadIdx.x) < ns) {
        const int64_t s = c0 + threadIdx.x;
        float x[kMlpIn], h1[kMlpH], h2[kMlpH], y[kMlpOut], dy[kMlpOut];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) x[k] = X[s * kMlpIn + k];
        mlp_forward(W, x, h1, h2, y);
        for (int o = 0; o < kMlpOut; o++) {
          const float t = T[s * kMlpOut + o];
          const bool ok = t == t;
          const float e = ok ? y[o] - t : 0.0f;
          lsum += e * e;
          dy[o] = 2.0f * e * inv_count;
        }
        float g2[kMlpH], g1[kMlpH];
        for (int k = 0; k < kMlpH; k++) {
          float a = 0.0f;
          for (int o = 0; o < kMlpOut; o++) a += W[oW3 + o * kMlpH + k] * dy[o];
          g2[k] = h2[k] > 0.0f ? a : 0.0f;
        }
        for (int k = 0; k < kMlpH; k++) {
          float a = 0.0f;
          for (int i = 0; i < kMlpH; i++) a += W[oW2 + i * kMlpH + k] * g2[i];
          g1[k] = h1[k] > 0.0f ? a : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) R[sX + k] = x[k];
        for (int k = 0; k < kMlpH; k++) {
          R[sH1 + k] = h1[k];
          R[sH2 + k] = h2[k];
          R[sG2 + k] = g2[k];
          R[sG1 + k] = g1[k];
        }
        for (int o = 0; o < kMlpOut; o++) R[sDY + o] = dy[o];
        R[sOne] = 1.0f;
      } else {
        for (int k = 0; k < kGradRow; k++) R[k] = 0.0f;  // past the last sample: adds zero
      }
    }
    __syncthreads();
    for (int s = 0; s < kGradChunk; s++) {
      const float *R = S + s * kGradRow;
#pragma unroll
      for (int j = 0; j < kGradOwn; j++) g[j] += R[ca[j]] * R[cb[j]];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kGradOwn; j++) {
    const int q = threadIdx.x + 256 * j;
    if (q < kMlpParams) gpart[int64_t(blockIdx.x) * kMlpParams + q] = g[j];
  }
  // loss: the threads' sums added in thread order
  Lsum[threadIdx.x] = lsum * inv_count;
  __syncthreads();
  if (threadIdx.x == 0) {
    float l = 0.0f;
    for (int i = 0; i < 256; i++) l += Lsum[i];
    lpart[blockIdx.x] = l;
  }
}

// Sum of the partial rows (fixed order) -> gradient and loss.
__global__ void k_mlp_reduce(const float *__restrict__ gpart, const float *__restrict__ lpart, int nparts,
                             float *__restrict__ grad, float *__restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpParams) {
    float g = 0.0f;
    for (int b = 0; b < nparts; b++) g += gpart[int64_t(b) * kMlpParams + i];
    grad[i] = g;
  }
  if (i == 0) {
    float l = 0.0f;
    for (int b = 0; b < nparts; b++) l += lpart[b];
    *loss = l;
  }
}

// Adam on the FP32 master weights.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  const float g = grad[i];
  m[i] = b1 * m[i] + (1.0f - b1) * g;
  v[i] = b2 * v[i] + (1.0f - b2) * g * g;
  const float mh = m[i] / (1.0f - powf(b1, float(step))), vh = v[i] / (1.0f - powf(b2, float(step)));
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- features
// Per cell: log10 of the share of the scaled multipole norm in degrees >= 2.
// One wave per cell: the shares of the squared real-form coefficients,
// read coalesced across the lanes and reduced over the wave.
// feat[c] = (shape, log10 sum|q| of the cell, or -30 without absq): the two
// per-source quantities the order selection reduces over a slot's columns,
// in one 8-byte load per column.
__global__ void __launch_bounds__(256) k_mshape(int PP, int ncells, const double *__restrict__ M,
                                                const double2 *__restrict__ absq, float2 *__restrict__ feat) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncells) return;
  const double *Mc = M + int64_t(c) * PP;
  double tot = 0.0, high = 0.0;
  for (int i = lane; i < PP; i += 64) {
    const double v = Mc[i] * Mc[i];
    tot += v;
    if (i >= 4) high += v;  // real-form index >= 4 <=> degree >= 2
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    tot += __shfl_xor(tot, o, 64);
    high += __shfl_xor(high, o, 64);
  }
  if (lane == 0) {
    const double r = tot > 0.0 ? sqrt(high / tot) : 0.0;
    const double aq = absq ? absq[c].x : 0.0;
    feat[c] = make_float2(float(fmin(0.0, fmax(-6.0, log10(fmax(r, 1e-30))))),
                          aq > 0.0 ? float(log10(aq)) : -30.0f);
  }
}

__device__ inline void pair_features(double dx, double dy, double dz, int a, int b, int level_t, float shape,
                                     float (&x)[kMlpIn]) {
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  x[0] = float(b - a);
  x[1] = float(log2(d / (ldexp(1.0, a) + ldexp(1.0, b))));
  x[2] = float(level_t) * 0.125f;
  x[3] = shape;
}

__device__ inline int order_of(const float (&y)[kMlpOut], float log_tol) {
  for (int o = 0; o < kMlpOut; o++)
    if (y[o] <= log_tol) return kMlpPmin + o;
  return kMlpPmin + kMlpOut - 1;
}

// Per cell (sum |q|, sum q^2): sum |q| is the scale the pair errors are
// normalised by, the root's sum q^2 enters the tolerance's scale.  Leaves
// from their bodies, then level by level from the children.
__global__ void __launch_bounds__(256) k_absq_leaf(int nleaves, const int *__restrict__ leaves,
                                                   const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                                   const double4 *__restrict__ xs, double2 *__restrict__ absq) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nleaves) return;
  const int c = leaves[i];
  double s = 0.0, s2 = 0.0;
  for (int j = cbeg[c], e = cbeg[c] + ccnt[c]; j < e; j++) {
    const double q = xs[j].w;
    s += fabs(q);
    s2 += q * q;
  }
  absq[c] = make_double2(s, s2);
}

__global__ void __launch_bounds__(256) k_absq_up(int lstart, int lcount, const int *__restrict__ cnch,
                                                 const int *__restrict__ cchild, double2 *__restrict__ absq) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= lcount) return;
  const int c = lstart + i, k = cnch[c];
  if (k == 0) return;
  double s = 0.0, s2 = 0.0;
  for (int j = 0; j < k; j++) {
    const double2 v = absq[cchild[c] + j];
    s += v.x;
    s2 += v.y;
  }
  absq[c] = make_double2(s, s2);
}

// Orders of the (tile, key) slots of the keyed GEMM.  All pairs of one
// slot share the key geometry and the target level (tiles hold cells of one
// level); they differ in the source cell, so the slot is evaluated once
// with the largest multipole shape and the largest sum of |q| among its
// sources (the pair that needs the highest order).  The model predicts a
// pair's potential error relative to sum|q_source| / |d|; the tolerance is
// on the absolute error, relative to the tree's potential scale
// S = sqrt((sum q)^2 + sum q^2) / R0 (the size of the potential whether the
// charges share a sign or not): the order is the smallest p with
//   y_p + log10((sum|q_s| / |d|) / S) <= log10_tol,
// so pairs of large, near cells (whose errors dominate a body's total) get
// higher orders than small, distant ones.  Per chunk of 1024
// slots of a tile: each wave takes slots in turn and reduces over the slot's
// 256 columns (coalesced row reads); then one thread per slot runs the MLP
// and caps the order at the context's order.
__global__ void __launch_bounds__(1024) k_order_select(const int *__restrict__ tfirst, const int *__restrict__ tcount,
                                                       const int *__restrict__ tbase, const int *__restrict__ tU,
                                                       const int *__restrict__ sorted, const int *__restrict__ tkeys,
                                                       const int *__restrict__ tmat, const unsigned long long *keys,
                                                       const int *__restrict__ clevel,
                                                       const float2 *__restrict__ feat,
                                                       const double2 *__restrict__ absq,
                                                       const double4 *__restrict__ cx, const double *__restrict__ M,
                                                       const float *__restrict__ w,
                                                       float log_tol, double R0, double cycle, int tn, int pmax,
                                                       unsigned char *__restrict__ tile_p) {
  __shared__ float W[kMlpParams];
  __shared__ float smax_s[1024];
  __shared__ float qmax_s[1024];
  load_weights(w, W);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;  // 16 waves
  const int tile = blockIdx.x;
  const int ncol = tcount[tile], U = tU[tile], base = tbase[tile];
  const int lt = clevel[sorted[tfirst[tile]]];
  const double Rt = cx[sorted[tfirst[tile]]].w;
  // log10 of the potential scale (M[0] = sum q of the root)
  const float lscale = float(log10(fmax(sqrt(M[0] * M[0] + absq[0].y) / cx[0].w, 1e-300)));
  for (int u0 = 0; u0 < U; u0 += 1024) {
    const int nu = min(1024, U - u0);
    // largest source shape per key: a wave per key, two keys (j, j + 16)
    // per pass so both rows' loads are in flight together
    for (int j = wv; j < nu; j += 32) {
      const int j2 = j + 16;
      const bool two = j2 < nu;
      const int *ra = tmat + (int64_t(base) + u0 + j) * tn;
      const int *rb = tmat + (int64_t(base) + u0 + (two ? j2 : j)) * tn;
      float ma = -1e30f, mb = -1e30f, qa = -1e30f, qb = -1e30f;
      for (int c = lane; c < ncol; c += 64) {
        const int sa = ra[c], sb = rb[c];
        if (sa >= 0) {
          const float2 f = feat[sa];
          ma = fmaxf(ma, f.x);
          qa = fmaxf(qa, f.y);
        }
        if (sb >= 0) {
          const float2 f = feat[sb];
          mb = fmaxf(mb, f.x);
          qb = fmaxf(qb, f.y);
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        ma = fmaxf(ma, __shfl_xor(ma, o, 64));
        mb = fmaxf(mb, __shfl_xor(mb, o, 64));
        qa = fmaxf(qa, __shfl_xor(qa, o, 64));
        qb = fmaxf(qb, __shfl_xor(qb, o, 64));
      }
      if (lane == 0) {
        smax_s[j] = ma;
        qmax_s[j] = qa;
        if (two) {
          smax_s[j2] = mb;
          qmax_s[j2] = qb;
        }
      }
    }
    __syncthreads();
    if (int(threadIdx.x) < nu) {
      const int u = u0 + threadIdx.x;
      const float smax = smax_s[threadIdx.x];
      int p = kMlpPmin;
      if (smax > -1e29f) {
        double dx, dy, dz;
        int a, b;
        key_at(keys, tkeys[base + u], R0, cycle, dx, dy, dz, a, b);
        float x[kMlpIn], h1[kMlpH], h2[kMlpH], y[kMlpOut];
        pair_features(dx, dy, dz, a, b, lt, smax, x);
        mlp_forward(W, x, h1, h2, y);
        // |d| in units of the finer radius Rf = Rt / 2^a
        const double d = sqrt(dx * dx + dy * dy + dz * dz) * ldexp(Rt, -a);
        const float shift = qmax_s[threadIdx.x] - float(log10(d)) - lscale;
        p = order_of(y, log_tol - shift);
      }
      tile_p[base + u] = (unsigned char)min(p, pmax);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- training data
// One workgroup (64 threads) per sampled M2L pair (target cell, list entry):
// features, and for p = 4..14 the log10 RMS error of the order-p M2L + L2P
// potential at up to 64 target bodies against the FP64 direct sum over the
// source cell's bodies, normalised by sum |q_source| / |d|.  Scaled
// real-form operator entries (op_entry) at order kMlpPmax, FP64 throughout.
constexpr int kPm = kMlpPmin + kMlpOut - 1;  // 14
constexpr int kPmPP = kPm * kPm;

__global__ void __launch_bounds__(64) k_pair_errors(int npairs, const int *__restrict__ pt,
                                                    const int2 *__restrict__ pe, const unsigned long long *keys,
                                                    double R0, double cycle, const double4 *__restrict__ cx,
                                                    const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                                    const int *__restrict__ clevel, const double4 *__restrict__ xs,
                                                    const double *__restrict__ M, const float2 *__restrict__ shape,
                                                    float *__restrict__ X, float *__restrict__ T) {
  __shared__ cpx<double> I[(2 * kPm - 1) * (2 * kPm) / 2];
  __shared__ double Lp[kMlpOut][kPmPP];
  __shared__ double Ms[kPmPP];
  __shared__ double errs[kMlpOut][64];
  __shared__ double ex2[64];
  __shared__ double qabs[64];
  const int pi = blockIdx.x;
  if (pi >= npairs) return;
  const int lane = threadIdx.x;
  const int ct = pt[pi];
  const int2 en = pe[pi];
  const int cs = en.x >> 5, img = en.x & 31;
  double dx, dy, dz;
  int a, b;
  key_at(keys, en.y, R0, cycle, dx, dy, dz, a, b);
  const int fillP = 2 * kPm - 1;
  if (lane < fillP) singular_column(lane, fillP, dx, dy, dz, I);
  for (int i = lane; i < kPmPP; i += 64) Ms[i] = M[int64_t(cs) * kPmPP + i];
  __syncthreads();
  // L~ at every order: row-major over rows, each column added to the orders
  // that contain both its row and its column degree
  for (int row = lane; row < kPmPP; row += 64) {
    int j = 0;
    while ((j + 1) * (j + 1) <= row) j++;
    double acc[kMlpOut];
#pragma unroll
    for (int o = 0; o < kMlpOut; o++) acc[o] = 0.0;
    for (int col = 0; col < kPmPP; col++) {
      int n = 0;
      while ((n + 1) * (n + 1) <= col) n++;
      const double v = op_entry(kPm, fillP, row, col, a, b, I) * Ms[col];
      const int pmin = max(j, n) + 1;  // orders p > max(j, n) include this term
#pragma unroll
      for (int o = 0; o < kMlpOut; o++)
        if (kMlpPmin + o >= pmin) acc[o] += v;
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; o++) Lp[o][row] = acc[o];
  }
  __syncthreads();
  // targets: up to 64 bodies of the target cell (stride), sources: all
  const double4 Ct = cx[ct], Cs = cx[cs];
  const double sx = ((img % 3) - 1) * cycle, sy = (((img / 3) % 3) - 1) * cycle, sz = ((img / 9) - 1) * cycle;
  const int tb = cbeg[ct], tn = ccnt[ct], sb = cbeg[cs], sn = ccnt[cs];
  const int nt = min(tn, 64);
  const int stride = max(1, tn / 64);
  double qs = 0.0;
  for (int j = lane; j < sn; j += 64) qs += fabs(xs[sb + j].w);
  qabs[lane] = qs;
  double e2[kMlpOut];
#pragma unroll
  for (int o = 0; o < kMlpOut; o++) e2[o] = 0.0;
  double x2 = 0.0;
  if (lane < nt) {
    const double4 B = xs[tb + lane * stride];
    double pex = 0.0;
    for (int j = 0; j < sn; j++) {
      const double4 S = xs[sb + j];
      const double ddx = B.x - S.x - sx, ddy = B.y - S.y - sy, ddz = B.z - S.z - sz;
      const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
      if (r2 > 0.0) pex += S.w / sqrt(r2);
    }
    // L2P at every order: phi = (1/R) sum w_k Re(L~ Y(u))
    const double invR = 1.0 / Ct.w;
    const double ux = (B.x - Ct.x) * invR, uy = (B.y - Ct.y) * invR, uz = (B.z - Ct.z) * invR;
    cpx<double> Y[kPm * (kPm + 1) / 2];
    regular_all<double>(kPm, ux, uy, uz, Y);
    for (int o = 0; o < kMlpOut; o++) {
      const int p = kMlpPmin + o;
      double ph = 0.0;
      for (int jj = 0; jj < p; jj++)
        for (int k = 0; k <= jj; k++) {
          const cpx<double> y = Y[nms(jj, k)];
          const double lr = Lp[o][rf(jj, k, 0)], li = k ? Lp[o][rf(jj, k, 1)] : 0.0;
          ph += (k ? 2.0 : 1.0) * (lr * y.re - li * y.im);
        }
      ph *= invR;
      e2[o] = (ph - pex) * (ph - pex);
    }
    x2 = pex * pex;
  }
#pragma unroll
  for (int o = 0; o < kMlpOut; o++) errs[o][lane] = e2[o];
  ex2[lane] = x2;
  __syncthreads();
  if (lane == 0) {
    double q = 0.0;
    for (int k = 0; k < 64; k++) q += qabs[k];
    const double rf_ = fmin(Ct.w, Cs.w);
    const double dist = sqrt(dx * dx + dy * dy + dz * dz) * rf_;
    const double scale = q / dist;
    for (int o = 0; o < kMlpOut; o++) {
      double s = 0.0;
      for (int k = 0; k < nt; k++) s += errs[o][k];
      const double e = sqrt(s / fmax(nt, 1)) / fmax(scale, 1e-300);
      T[int64_t(pi) * kMlpOut + o] = (q > 0.0 && nt > 0) ? float(log10(fmax(e, 1e-30))) : __builtin_nanf("");
    }
    float x[kMlpIn];
    pair_features(dx, dy, dz, a, b, clevel[ct], shape[cs].x, x);
    for (int k = 0; k < kMlpIn; k++) X[int64_t(pi) * kMlpIn + k] = x[k];
  }
}

// ---------------------------------------------------------------- host side
static void mlp_grad(Ctx &c, const float *w, const float *X, const float *T, int64_t n, float *grad, float *loss,
                     hipStream_t s) {
  // one chunk of samples per workgroup up to 256 workgroups (a fixed
  // function of n: the partial rows, and so the gradient, do not depend on
  // the run)
  const int nb = int(std::min<int64_t>(256, std::max<int64_t>(1, ceil_div(n, kGradChunk))));
  c.mlp_part.ensure(size_t(nb) * (kMlpParams + 1));
  float *gp = c.mlp_part.p, *lp = c.mlp_part.p + size_t(nb) * kMlpParams;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s, w, X, T, n, 1.0f / float(std::max<int64_t>(n, 1)), gp,
                     lp);
  XF_LAUNCH();
  hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kMlpParams, 256)), dim3(256), 0, s, gp, lp, nb, grad, loss);
  XF_LAUNCH();
}

void order_select(Ctx &c, const int *tfirst, const int *tcount, const int *tbase, const int *tU, const int *sorted,
                  int ntiles, int totalU, int tn, hipStream_t s) {
  c.tile_p.ensure(std::max(totalU, 1));
  const int nc = int(c.t.ncells);
  // sum |q| per cell, then the per-cell features
  const Tree &t = c.t;
  c.absq.ensure(nc);
  hipLaunchKernelGGL(k_absq_leaf, dim3(ceil_div(std::max<int64_t>(t.nleaves, 1), 256)), dim3(256), 0, s,
                     int(t.nleaves), t.leaves.p, t.cbeg.p, t.ccnt.p, t.xs.p, c.absq.p);
  XF_LAUNCH();
  for (int l = t.depth - 1; l >= 0; l--) {
    const int cnt = t.lstart[l + 1] - t.lstart[l];
    hipLaunchKernelGGL(k_absq_up, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, t.lstart[l], cnt, t.cnch.p,
                       t.cchild.p, c.absq.p);
    XF_LAUNCH();
  }
  c.mshape.ensure(nc);
  hipLaunchKernelGGL(k_mshape, dim3(ceil_div(nc, 4)), dim3(256), 0, s, c.P * c.P, nc, c.M.p, c.absq.p, c.mshape.p);
  XF_LAUNCH();
  const double cyc = c.cfg.images > 0 ? c.cfg.cycle : 0.0;
  hipLaunchKernelGGL(k_order_select, dim3(ntiles), dim3(1024), 0, s, tfirst, tcount, tbase, tU, sorted,
                     c.tile_keys.p, c.tile_mat.p, c.lists.keys.p, c.t.clevel.p, c.mshape.p, c.absq.p, c.t.cx.p,
                     c.M.p, c.mlp_w.p, c.order_log_tol, c.t.R0, cyc, tn, c.P, c.tile_p.p);
  XF_LAUNCH();
}

// Training: pairs sampled from the M2L lists of the context's current
// evaluation (tree, lists and upward pass at order 14 must exist).
float order_train(Ctx &c, int nsamples, int epochs, float lr, uint64_t seed) {
  if (c.P != kPm) fail(EXAFMM_ERR_ARG, "order-model training needs a context with P = 14");
  if (!c.have_M || !c.lists.valid || c.lists.nkeys == 0)
    fail(EXAFMM_ERR_STATE, "order-model training needs build_tree, upward and a horizontal pass");
  hipStream_t s = c.s0;
  const int64_t nm = c.lists.nm2l;
  const int nc = int(c.t.ncells);
  // host: list offsets -> sampled (target cell, entry index)
  std::vector<int> beg(nc), cnt(nc);
  XF_HIP(hipMemcpyAsync(beg.data(), c.lists.m2l_beg.p, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
  XF_HIP(hipMemcpyAsync(cnt.data(), c.lists.m2l_cnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
  XF_HIP(hipStreamSynchronize(s));
  std::vector<int> owner(nm);
  for (int t = 0; t < nc; t++)
    for (int k = 0; k < cnt[t]; k++) owner[beg[t] + k] = t;
  std::mt19937_64 rng(seed);
  const int n = int(std::min<int64_t>(nsamples, nm));
  std::vector<int> pt(n), pidx(n);
  for (int i = 0; i < n; i++) {
    const int64_t e = int64_t(rng() % uint64_t(nm));
    pidx[i] = int(e);
    pt[i] = owner[e];
  }
  Buf<int> dpt, didx;
  Buf<int2> dpe;
  dpt.ensure(n);
  didx.ensure(n);
  dpe.ensure(n);
  XF_HIP(hipMemcpyAsync(dpt.p, pt.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  XF_HIP(hipMemcpyAsync(didx.p, pidx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  std::vector<int2> pe(n);
  {
    std::vector<int2> all(nm);
    XF_HIP(hipMemcpyAsync(all.data(), c.lists.m2l.p, sizeof(int2) * nm, hipMemcpyDeviceToHost, s));
    XF_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) pe[i] = all[pidx[i]];
  }
  XF_HIP(hipMemcpyAsync(dpe.p, pe.data(), sizeof(int2) * n, hipMemcpyHostToDevice, s));
  c.mshape.ensure(nc);
  hipLaunchKernelGGL(k_mshape, dim3(ceil_div(nc, 4)), dim3(256), 0, s, c.P * c.P, nc, c.M.p, (const double2 *)nullptr,
                     c.mshape.p);
  XF_LAUNCH();
  Buf<float> X, T;
  X.ensure(size_t(n) * kMlpIn);
  T.ensure(size_t(n) * kMlpOut);
  const double cyc = c.cfg.images > 0 ? c.cfg.cycle : 0.0;
  hipLaunchKernelGGL(k_pair_errors, dim3(n), dim3(64), 0, s, n, dpt.p, dpe.p, c.lists.keys.p, c.t.R0, cyc, c.t.cx.p,
                     c.t.cbeg.p, c.t.ccnt.p, c.t.clevel.p, c.t.xs.p, c.M.p, c.mshape.p, X.p, T.p);
  XF_LAUNCH();
  // weights: small random init on the host, Adam on the device
  c.mlp_w.ensure(kMlpParams);
  Buf<float> m, v, grad, loss;
  m.ensure(kMlpParams);
  v.ensure(kMlpParams);
  grad.ensure(kMlpParams);
  loss.ensure(1);
  std::vector<float> w0(kMlpParams, 0.0f);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int i = 0; i < kMlpH * kMlpIn; i++) w0[oW1 + i] = nd(rng) * std::sqrt(2.0f / kMlpIn);
  for (int i = 0; i < kMlpH * kMlpH; i++) w0[oW2 + i] = nd(rng) * std::sqrt(2.0f / kMlpH);
  for (int i = 0; i < kMlpOut * kMlpH; i++) w0[oW3 + i] = nd(rng) * std::sqrt(1.0f / kMlpH);
  for (int o = 0; o < kMlpOut; o++) w0[oB3 + o] = -1.0f - 0.5f * o;  // errors fall with the order
  XF_HIP(hipMemcpyAsync(c.mlp_w.p, w0.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
  XF_HIP(hipMemsetAsync(m.p, 0, sizeof(float) * kMlpParams, s));
  XF_HIP(hipMemsetAsync(v.p, 0, sizeof(float) * kMlpParams, s));
  for (int step = 1; step <= epochs; step++) {
    mlp_grad(c, c.mlp_w.p, X.p, T.p, n, grad.p, loss.p, s);
    hipLaunchKernelGGL(k_adam, dim3(ceil_div(kMlpParams, 256)), dim3(256), 0, s, c.mlp_w.p, grad.p, m.p, v.p, lr,
                       0.9f, 0.999f, 1e-8f, step);
    XF_LAUNCH();
  }
  mlp_grad(c, c.mlp_w.p, X.p, T.p, n, grad.p, loss.p, s);
  float lv = 0.0f;
  XF_HIP(hipMemcpyAsync(&lv, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  // keep the training set and its pairs for inspection
  c.train_pairs.resize(size_t(n) * 3);
  for (int i = 0; i < n; i++) {
    c.train_pairs[3 * i] = pt[i];
    c.train_pairs[3 * i + 1] = pe[i].x >> 5;
    c.train_pairs[3 * i + 2] = pe[i].x & 31;
  }
  c.train_n = n;
  c.train_X.ensure(size_t(n) * kMlpIn);
  c.train_T.ensure(size_t(n) * kMlpOut);
  XF_HIP(hipMemcpyAsync(c.train_X.p, X.p, sizeof(float) * n * kMlpIn, hipMemcpyDeviceToDevice, s));
  XF_HIP(hipMemcpyAsync(c.train_T.p, T.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToDevice, s));
  XF_HIP(hipStreamSynchronize(s));
  c.have_model = true;
  return lv;
}

void mlp_eval_host(Ctx &c, const float *w, const float *X, int64_t n, float *Y) {
  hipStream_t s = c.s0;
  Buf<float> dw, dX, dY;
  dw.ensure(kMlpParams);
  dX.ensure(std::max<int64_t>(n, 1) * kMlpIn);
  dY.ensure(std::max<int64_t>(n, 1) * kMlpOut);
  XF_HIP(hipMemcpyAsync(dw.p, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
  XF_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_forward, dim3(int(std::min<int64_t>(256, std::max<int64_t>(1, ceil_div(n, 256))))),
                     dim3(256), 0, s, dw.p, dX.p, n, dY.p);
  XF_LAUNCH();
  XF_HIP(hipMemcpyAsync(Y, dY.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, s));
  XF_HIP(hipStreamSynchronize(s));
}

void mlp_grad_host(Ctx &c, const float *w, const float *X, const float *T, int64_t n, float *grad, float *loss) {
  hipStream_t s = c.s0;
  Buf<float> dw, dX, dT, dg, dl;
  dw.ensure(kMlpParams);
  dX.ensure(std::max<int64_t>(n, 1) * kMlpIn);
  dT.ensure(std::max<int64_t>(n, 1) * kMlpOut);
  dg.ensure(kMlpParams);
  dl.ensure(1);
  XF_HIP(hipMemcpyAsync(dw.p, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
  XF_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s));
  XF_HIP(hipMemcpyAsync(dT.p, T, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, s));
  mlp_grad(c, dw.p, dX.p, dT.p, n, dg.p, dl.p, s);
  XF_HIP(hipMemcpyAsync(grad, dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
  XF_HIP(hipMemcpyAsync(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost, s));
  XF_HIP(hipStreamSynchronize(s));
}

}  // namespace xf
