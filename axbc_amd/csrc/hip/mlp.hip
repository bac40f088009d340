// Learned pivot proposal: a 3-layer MLP over per-vertex features
// (log degree, LCC, running bc, mean neighbour bc) -> score, with BF16 weights,
// FP32 accumulation, hand-written forward, backward and Adam kernels; the
// 64 x 64 layer and its backward products run on the BF16 matrix cores.
//
// Network:  x (4) -> h1 = relu(W1 x + b1) (64) -> h2 = relu(W2 h1 + b2) (64)
//           -> s = w3 . h2 + b3.
// Training: mean squared error between s and a target per vertex, minibatches
// of vertices drawn by the host; gradients summed per workgroup in LDS and
// added to FP32 gradient buffers; Adam on FP32 master weights, BF16 copies
// used by the forward passes.
// Proposal: w(v) = expm1(s(v)) + eps * mean (degree > 0), turned into the
// fixed-point distribution of the dynamic samplers (kind 7).
//
// No reference counterpart: it stands in for the bc/delta heuristics
// (-z 3/4, src/axbc2d.c:2941-2994) as a pivot proposal.

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <vector>

#include "../../../include/axbc.h"
#include "engine_internal.h"
#include "wave.h"

using namespace axbc;

namespace {

constexpr int kIn = 4, kH = 64;
// parameter layout (floats): W1[64][4], b1[64], W2[64][64], b2[64], w3[64], b3
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
constexpr int kParams = oB3 + 1;  // 4545

struct MlpBufs {
  float *x = nullptr;        // N x 4 features
  float *y = nullptr;        // N targets
  float *score = nullptr;    // N scores
  float *master = nullptr;   // FP32 master weights
  __hip_bfloat16 *wb = nullptr;  // BF16 weights
  float *grad = nullptr, *m = nullptr, *v = nullptr;
  float *loss = nullptr;
  double *prop = nullptr;    // N proposal weights
  double *tw = nullptr;      // N target weights (feature 1 while targets are set)
  double *tred = nullptr;    // per-block (sum, count) of the target feature, then its mean
  bool targets = false;
  uint32_t *idx = nullptr;
  uint32_t idx_cap = 0;
  int step = 0;
  float *gpart = nullptr;    // kGradBlocks x (kParams + 1): per-workgroup gradients and loss
  double *wpart = nullptr;   // per-wave (weight, count) sums of the proposal
};

MlpBufs *bufs_of(axbc_ctx *c) { return reinterpret_cast<MlpBufs *>(c->mlp); }


__device__ __forceinline__ float bf(const __hip_bfloat16 *w, int i) { return __bfloat162float(w[i]); }

// Features of vertex v: log(1 + degree), LCC (or, while targets are set,
// the target sum t(v) = sum of tw over v's neighbours, normalised below),
// log(1 + bc * scale) and log(1 + mean neighbour bc * scale), or, given a
// last dependency row dlast, log(1 + dlast(v) * dscale).  Thread per
// vertex; vertices with more than 64 edges are summed by the whole wave.
__global__ void k_features(uint32_t n, const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                           const float *__restrict__ lcc, const double *__restrict__ bc, double bc_scale,
                           const double *__restrict__ tw, const double *__restrict__ dlast, double dscale,
                           float *x) {
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < n; base += stride) {
    const uint32_t v = base + (uint32_t)lane;
    uint32_t b = 0, e = 0;
    if (v < n) {
      b = col[v];
      e = col[v + 1];
    }
    const bool big = e - b > 64;
    double nsum = 0.0, tsum = 0.0;
    if (!big)
      for (uint32_t k = b; k < e; ++k) {
        nsum += bc[row[k]];
        if (tw) tsum += tw[row[k]];
      }
    unsigned long long todo = __ballot(big);
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t bb = __shfl(b, i), ee = __shfl(e, i);
      double t = 0.0, u = 0.0;
      for (uint32_t k = bb + (uint32_t)lane; k < ee; k += 64) {
        t += bc[row[k]];
        if (tw) u += tw[row[k]];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_xor(t, off);
        u += __shfl_xor(u, off);
      }
      if (lane == i) {
        nsum = t;
        tsum = u;
      }
    }
    if (v < n) {
      const uint32_t d = e - b;
      x[4 * v + 0] = log1pf((float)d);
      x[4 * v + 1] = tw ? (float)tsum : lcc ? lcc[v] : 0.0f;
      x[4 * v + 2] = (float)log1p(bc[v] * bc_scale);
      x[4 * v + 3] = dlast ? (float)log1p(dlast[v] * dscale) : d ? (float)log1p(nsum / (double)d * bc_scale) : 0.0f;
    }
  }
}

constexpr int kRedBlocks = 256;

// Per block, the sum and count of the positive target features, reduced in
// a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_target_partials(uint32_t n, const float *__restrict__ x, double *part) {
  __shared__ double ss[256], sc[256];
  double a = 0.0, c = 0.0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const float t = x[4 * v + 1];
    if (t > 0.f) {
      a += (double)t;
      c += 1.0;
    }
  }
  ss[threadIdx.x] = a;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ss[threadIdx.x] += ss[threadIdx.x + h];
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ss[0];
    part[2 * blockIdx.x + 1] = sc[0];
  }
}

// part[2 kRedBlocks] = mean of the positive target features (1 if none).
__global__ void k_target_mean(double *part) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, c = 0.0;
  for (int i = 0; i < kRedBlocks; ++i) {
    a += part[2 * i];
    c += part[2 * i + 1];
  }
  part[2 * kRedBlocks] = c > 0.0 ? a / c : 1.0;
}

// Target feature: log(1 + t / mean positive t).
__global__ void k_target_norm(uint32_t n, const double *__restrict__ part, float *x) {
  const double mean = part[2 * kRedBlocks];
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    x[4 * v + 1] = (float)log1p((double)x[4 * v + 1] / mean);
}

__global__ void k_scatter_targets(uint32_t m, const uint32_t *__restrict__ vert, const double *__restrict__ w,
                                  double *tw) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) tw[vert[i]] = w[i];
}

// ---- the 64 x 64 layer on the matrix cores ---------------------------------
//
// Layer 2 and its two backward products are 64 x 64 x 64 products per tile
// of 64 vertices (samples), run as 2 x 2 tiles of 32 x 32 on the BF16 MFMA
// (v_mfma_f32_32x32x8_bf16: eight k-steps of 8 per tile) with FP32
// accumulation.  Their inputs are BF16: the weights (kept in BF16 anyway)
// and the activations h1 and gradients dz2, rounded to nearest even.
// Operands sit in LDS as BF16 bits, k contiguous, rows padded to kLd.
// Layer 1 (4 -> 64), the output unit and the small gradients stay on the
// vector ALUs in FP32.

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kLd = kH + 4;  // row pitch (BF16 elements) of the LDS operands: 136 B, 8-B aligned rows

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// acc += A[0..32)[0..64) x B[0..64)[0..32) for the 32 x 32 output tile, A
// given by rows (k contiguous), B by its columns (Bt: column n is row n of
// Bt, k contiguous).  Lane l supplies rows / columns l % 32, k-quad l / 32.
__device__ __forceinline__ f32x16 mfma_tile(const unsigned short *A, const unsigned short *Bt, f32x16 acc) {
  const int l = threadIdx.x & 63;
  const unsigned short *a = A + (l & 31) * kLd + ((l >> 5) << 2);
  const unsigned short *b = Bt + (l & 31) * kLd + ((l >> 5) << 2);
#pragma unroll
  for (int ks = 0; ks < kH; ks += 8)
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(*(const bf16x4 *)(a + ks), *(const bf16x4 *)(b + ks), acc, 0, 0,
                                                    0);
  return acc;
}

// Output element of register r in lane l of a 32 x 32 MFMA tile: row
// 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
__device__ __forceinline__ int tile_row(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// FP32 copies of the small parameters (W1, b1, b2, w3, b3)
struct SmallW {
  float w1[kH * kIn], b1[kH], b2[kH], w3[kH], b3;
};

// Small parameters as FP32, W2 as BF16 bits by rows ([out][in]) and, when
// w2t is given, by columns ([in][out]).
__device__ void load_weights(const __hip_bfloat16 *__restrict__ wb, SmallW &sw, unsigned short *w2,
                             unsigned short *w2t) {
  const unsigned short *raw = reinterpret_cast<const unsigned short *>(wb);
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) sw.w1[i] = bf(wb, oW1 + i);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    sw.b1[i] = bf(wb, oB1 + i);
    sw.b2[i] = bf(wb, oB2 + i);
    sw.w3[i] = bf(wb, oW3 + i);
  }
  if (threadIdx.x == 0) sw.b3 = bf(wb, oB3);
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
    const int j = i / kH, k = i % kH;  // W2[j][k]: unit j of layer 2, input k
    w2[j * kLd + k] = raw[oW2 + i];
    if (w2t) w2t[k * kLd + j] = raw[oW2 + i];
  }
  __syncthreads();
}

// Layer 1 of one vertex: h1 = relu(W1 x + b1)
__device__ __forceinline__ float layer1(const SmallW &sw, const float xv[kIn], int j) {
  float a = sw.b1[j];
#pragma unroll
  for (int k = 0; k < kIn; ++k) a += sw.w1[j * kIn + k] * xv[k];
  return a > 0.0f ? a : 0.0f;
}

// Scores: 128-thread workgroups, each wave a tile of 64 vertices per step
// (the loop is uniform over the workgroup, so the barriers are reached by
// both waves alike).
constexpr int kFwdThreads = 128;

__global__ __launch_bounds__(kFwdThreads) void k_mlp_forward(uint32_t n, const float *__restrict__ x,
                                                             const __hip_bfloat16 *__restrict__ wb, float *score) {
  __shared__ SmallW sw;
  __shared__ __align__(16) unsigned short W2s[kH * kLd];
  __shared__ __align__(16) unsigned short H1s[kFwdThreads / 64][64 * kLd];
  __shared__ float Z[kFwdThreads / 64][64][kH + 1];
  load_weights(wb, sw, W2s, nullptr);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (uint32_t base = blockIdx.x * kFwdThreads; base < n; base += gridDim.x * kFwdThreads) {
    const uint32_t v = base + threadIdx.x;
    float xv[kIn] = {0.f, 0.f, 0.f, 0.f};
    if (v < n)
#pragma unroll
      for (int k = 0; k < kIn; ++k) xv[k] = x[4 * v + k];
#pragma unroll 8
    for (int j = 0; j < kH; ++j) H1s[w][l * kLd + j] = bf16_bits(layer1(sw, xv, j));
    __syncthreads();
    // z2 = H1 W2^T: rows = vertices of the wave, columns = units
#pragma unroll
    for (int vb = 0; vb < 2; ++vb)
#pragma unroll
      for (int ob = 0; ob < 2; ++ob) {
        f32x16 acc = {};
        acc = mfma_tile(&H1s[w][vb * 32 * kLd], &W2s[ob * 32 * kLd], acc);
#pragma unroll
        for (int r = 0; r < 16; ++r) Z[w][vb * 32 + tile_row(r, l)][ob * 32 + (l & 31)] = acc[r];
      }
    __syncthreads();
    float s = sw.b3;
#pragma unroll 16
    for (int j = 0; j < kH; ++j) {
      const float a = Z[w][l][j] + sw.b2[j];
      s += sw.w3[j] * (a > 0.0f ? a : 0.0f);
    }
    if (v < n) score[v] = s;
    __syncthreads();
  }
}

// Gradient of 1/M sum (s - y)^2 over the samples idx[0..m).  A workgroup is
// one wave of 64 threads and takes tiles of 64 samples, lane t sample t:
//   z2 = H1 W2^T                 (MFMA; H1 = BF16 h1 of the tile)
//   dz2 = [z2 > 0] ds w3^T       (ds = 2 (s - y) / m per sample)
//   dh1 = DZ2 W2                 (MFMA; DZ2 = BF16 dz2), dz1 = [h1 > 0] dh1
//   gW2 += DZ2^T H1              (MFMA, accumulated over the tiles in registers)
// and the small gradients by owner lanes over the tile's samples (lane j:
// b2[j], w3[j], b1[j], W1[j][.]).  One atomic add per parameter at the end.
constexpr int kTile = 64;

// Workgroups of the gradient pass: each writes its whole gradient (and its
// loss) to its own row of gpart, and k_grad_reduce adds the rows in
// workgroup order, so a gradient does not depend on the run.
constexpr int kGradBlocks = 1024;
constexpr int kGradRow = kParams + 1;

__global__ __launch_bounds__(64) void k_mlp_grad(uint32_t m, const uint32_t *__restrict__ idx,
                                                 const float *__restrict__ x, const float *__restrict__ y,
                                                 const __hip_bfloat16 *__restrict__ wb, float *gpart) {
  __shared__ SmallW sw;
  __shared__ __align__(16) unsigned short W2s[kH * kLd];  // [unit j][input k]
  __shared__ __align__(16) unsigned short W2t[kH * kLd];  // [input k][unit j]
  __shared__ __align__(16) unsigned short H1s[kTile * kLd], H1t[kH * kLd];  // [sample][k], [k][sample]
  __shared__ __align__(16) unsigned short DZs[kTile * kLd], DZt[kH * kLd];  // [sample][j], [j][sample]
  __shared__ float Z[kTile][kH + 1];  // z2 (pre-bias), then dz1
  __shared__ float X[kTile][kIn], DS[kTile];
  load_weights(wb, sw, W2s, W2t);
  const int t = threadIdx.x;
  f32x16 gW2[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) gW2[a][b] = f32x16{};
  float gW1[kIn] = {0.f, 0.f, 0.f, 0.f}, gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, lsum = 0.f;
  const uint32_t ntiles = (m + kTile - 1) / kTile;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t si = tile * kTile + t;
    const bool live = si < m;
    uint32_t v = 0;
    float xv[kIn] = {0.f, 0.f, 0.f, 0.f};
    if (live) {
      v = idx[si];
#pragma unroll
      for (int k = 0; k < kIn; ++k) xv[k] = x[4 * v + k];
    }
    // layer 1 (a sample outside the range has h1 = 0 and ds = 0)
    uint64_t pos1 = 0;  // h1 > 0 per unit
#pragma unroll 8
    for (int j = 0; j < kH; ++j) {
      const float h = live ? layer1(sw, xv, j) : 0.0f;
      pos1 |= (uint64_t)(h > 0.0f) << j;
      const unsigned short hb = bf16_bits(h);
      H1s[t * kLd + j] = hb;
      H1t[j * kLd + t] = hb;
    }
#pragma unroll
    for (int k = 0; k < kIn; ++k) X[t][k] = xv[k];
    __syncthreads();
#pragma unroll
    for (int vb = 0; vb < 2; ++vb)
#pragma unroll
      for (int ob = 0; ob < 2; ++ob) {
        f32x16 acc = {};
        acc = mfma_tile(&H1s[vb * 32 * kLd], &W2s[ob * 32 * kLd], acc);
#pragma unroll
        for (int r = 0; r < 16; ++r) Z[vb * 32 + tile_row(r, t)][ob * 32 + (t & 31)] = acc[r];
      }
    __syncthreads();
    // output, loss and dz2 of sample t
    float s = sw.b3;
#pragma unroll 16
    for (int j = 0; j < kH; ++j) {
      const float a = Z[t][j] + sw.b2[j];
      s += sw.w3[j] * (a > 0.0f ? a : 0.0f);
    }
    float ds = 0.0f;
    if (live) {
      const float r = s - y[v];
      lsum += r * r;
      ds = 2.0f * r / (float)m;
    }
    DS[t] = ds;
#pragma unroll 8
    for (int j = 0; j < kH; ++j) {
      const float dz = Z[t][j] + sw.b2[j] > 0.0f ? ds * sw.w3[j] : 0.0f;
      const unsigned short db = bf16_bits(dz);
      DZs[t * kLd + j] = db;
      DZt[j * kLd + t] = db;
    }
    __syncthreads();
    // owner lane j: b2[j] and w3[j] over the tile's samples (FP32)
    for (int q = 0; q < kTile; ++q) {
      const float a = Z[q][t] + sw.b2[t];
      if (a > 0.0f) {
        gb2 += DS[q] * sw.w3[t];
        gw3 += DS[q] * a;
      }
      if (t == 0) gb3 += DS[q];
    }
    // gW2 += DZ2^T H1: rows = units j, columns = inputs k
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) gW2[jb][kb] = mfma_tile(&DZt[jb * 32 * kLd], &H1t[kb * 32 * kLd], gW2[jb][kb]);
    __syncthreads();  // Z is overwritten below
    // dh1 = DZ2 W2: rows = samples, columns = inputs k
#pragma unroll
    for (int sb = 0; sb < 2; ++sb)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        f32x16 acc = {};
        acc = mfma_tile(&DZs[sb * 32 * kLd], &W2t[kb * 32 * kLd], acc);
#pragma unroll
        for (int r = 0; r < 16; ++r) Z[sb * 32 + tile_row(r, t)][kb * 32 + (t & 31)] = acc[r];
      }
    __syncthreads();
    // dz1 of sample t, in place
#pragma unroll 16
    for (int k = 0; k < kH; ++k)
      if (!((pos1 >> k) & 1)) Z[t][k] = 0.0f;
    __syncthreads();
    // owner lane j: b1[j], W1[j][.]
    for (int q = 0; q < kTile; ++q) {
      const float dz1 = Z[q][t];
      gb1 += dz1;
#pragma unroll
      for (int k = 0; k < kIn; ++k) gW1[k] += dz1 * X[q][k];
    }
    __syncthreads();
  }
  // every entry of the row is written: each parameter has one owner lane
  float *grad = gpart + (size_t)blockIdx.x * kGradRow;
#pragma unroll
  for (int jb = 0; jb < 2; ++jb)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        grad[oW2 + (jb * 32 + tile_row(r, t)) * kH + kb * 32 + (t & 31)] = gW2[jb][kb][r];
#pragma unroll
  for (int k = 0; k < kIn; ++k) grad[oW1 + t * kIn + k] = gW1[k];
  grad[oB1 + t] = gb1;
  grad[oB2 + t] = gb2;
  grad[oW3 + t] = gw3;
  if (t == 0) grad[oB3] = gb3;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (t == 0) grad[kParams] = lsum / (float)m;
}

// grad[i] += the rows' entries i in workgroup order (i = kParams: the loss,
// stored instead of added).
__global__ void k_grad_reduce(const float *__restrict__ gpart, int rows, float *grad, float *loss) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= kParams; i += gridDim.x * blockDim.x) {
    float t = 0.f;
    for (int r = 0; r < rows; ++r) t += gpart[(size_t)r * kGradRow + i];
    if (i < kParams)
      grad[i] += t;
    else
      *loss = t;
  }
}

__global__ void k_adam(int n, float *master, __hip_bfloat16 *wb, float *grad, float *m, float *v, float lr,
                       float b1, float b2, float eps, int step) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = grad[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    master[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    wb[i] = __float2bfloat16(master[i]);
    grad[i] = 0.0f;
  }
}

__global__ void k_to_bf16(int n, const float *__restrict__ src, __hip_bfloat16 *dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = __float2bfloat16(src[i]);
}

// The score predicts log(1 + utility / mean utility) of a root, so the
// proposal weight of a vertex of positive degree is the predicted relative
// utility expm1(score) (clamped to [0, e^30]) plus eps times its mean over
// those vertices; vertices without edges cannot be roots and get weight 0.
__device__ __forceinline__ double score_weight(float s) {
  return expm1(fmin(fmax((double)s, 0.0), 30.0));
}

__global__ void k_proposal(uint32_t n, const uint32_t *__restrict__ col, const float *__restrict__ score,
                           double floor_w, double *w) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    w[v] = col[v + 1] > col[v] ? score_weight(score[v]) + floor_w : 0.0;
}

// Per wave (weights, vertices of positive degree) into part[2 gw ..], then
// k_weight_total adds the waves' pairs in wave order: the proposal's
// normalisation does not depend on the run.
__global__ void k_weight_sum(uint32_t n, const uint32_t *__restrict__ col, const float *__restrict__ score,
                             double *part) {
  double a = 0.0, c = 0.0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (col[v + 1] > col[v]) {
      a += score_weight(score[v]);
      c += 1.0;
    }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    c += __shfl_xor(c, off);
  }
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (lane_id() == 0) {
    part[2 * gw] = a;
    part[2 * gw + 1] = c;
  }
}

// One workgroup of 256: thread t adds waves t, t + 256, ... in order, then a
// fixed tree over the threads.
__global__ __launch_bounds__(256) void k_weight_total(const double *__restrict__ part, uint32_t waves, double *sum) {
  __shared__ double sa[256], sc[256];
  double a = 0.0, c = 0.0;
  for (uint32_t w = threadIdx.x; w < waves; w += 256) {
    a += part[2 * w];
    c += part[2 * w + 1];
  }
  sa[threadIdx.x] = a;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sa[threadIdx.x] += sa[threadIdx.x + h];
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sum[0] = sa[0];
    sum[1] = sc[0];
  }
}

int grid_of(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

}  // namespace

namespace axbc {
void mlp_free(axbc_ctx *c) {
  MlpBufs *b = bufs_of(c);
  if (!b) return;
  void *d[] = {b->x,   b->y,    b->score, b->master, b->wb,   b->grad,  b->m,    b->v,
               b->loss, b->idx, b->prop,  b->tw,     b->tred, b->gpart, b->wpart};
  for (void *p : d)
    if (p) (void)hipFree(p);
  delete b;
  c->mlp = nullptr;
}
}  // namespace axbc

extern "C" {

int axbc_mlp_init(axbc_ctx *c, const float *params) {
  if (!c || !params) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(c->device));
  MlpBufs *b = bufs_of(c);
  if (!b) {
    const size_t n = c->n;
    {
      int rc;
      if ((rc = budget_check(c, 32 * n, "the MLP buffers"))) return rc;
    }
    b = new MlpBufs();
    c->mlp = b;
    HIP_TRY(hipMalloc((void **)&b->x, sizeof(float) * 4 * n));
    HIP_TRY(hipMalloc((void **)&b->y, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&b->score, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&b->master, sizeof(float) * kParams));
    HIP_TRY(hipMalloc((void **)&b->wb, sizeof(__hip_bfloat16) * kParams));
    HIP_TRY(hipMalloc((void **)&b->grad, sizeof(float) * kParams));
    HIP_TRY(hipMalloc((void **)&b->m, sizeof(float) * kParams));
    HIP_TRY(hipMalloc((void **)&b->v, sizeof(float) * kParams));
    HIP_TRY(hipMalloc((void **)&b->loss, sizeof(float)));
    HIP_TRY(hipMalloc((void **)&b->prop, sizeof(double) * n));
    c->bytes += 32 * n;
  }
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(b->master, params, sizeof(float) * kParams, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(b->grad, 0, sizeof(float) * kParams, s));
  HIP_TRY(hipMemsetAsync(b->m, 0, sizeof(float) * kParams, s));
  HIP_TRY(hipMemsetAsync(b->v, 0, sizeof(float) * kParams, s));
  hipLaunchKernelGGL(k_to_bf16, dim3(32), dim3(256), 0, s, kParams, b->master, b->wb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  b->step = 0;
  return AXBC_OK;
}

int axbc_mlp_param_count(void) { return kParams; }

int axbc_mlp_features(axbc_ctx *c, double bc_scale) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  const float *lcc = c->sampler && c->sampler->lcc_ready ? c->sampler->lcc : nullptr;
  const double *tw = b->targets ? b->tw : nullptr;
  hipLaunchKernelGGL(k_features, dim3(grid_of(c->n)), dim3(256), 0, c->stream, c->n, c->col, c->row, lcc, c->bc,
                     bc_scale, tw, (const double *)nullptr, 0.0, b->x);
  if (tw) {
    hipLaunchKernelGGL(k_target_partials, dim3(kRedBlocks), dim3(256), 0, c->stream, c->n, b->x, b->tred);
    hipLaunchKernelGGL(k_target_mean, dim3(1), dim3(64), 0, c->stream, b->tred);
    hipLaunchKernelGGL(k_target_norm, dim3(grid_of(c->n)), dim3(256), 0, c->stream, c->n, b->tred, b->x);
  }
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

// The four inputs of the specification: log(1 + degree), LCC (built here
// when the context has none), log(1 + bc * bc_scale) and the last
// dependency values, log(1 + delta_k(v) * delta_scale) with delta_k root k's
// row of the last batch.
int axbc_mlp_features_spec(axbc_ctx *c, double bc_scale, int k, double delta_scale) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  int rc;
  if (!(c->sampler && c->sampler->lcc_ready)) {
    uint64_t total = 0;
    if ((rc = axbc_lcc_build(c, nullptr, nullptr, nullptr, &total))) return rc;
  }
  const double *dl = nullptr;
  if ((rc = batch_delta_row(c, k, &dl))) return rc;
  hipLaunchKernelGGL(k_features, dim3(grid_of(c->n)), dim3(256), 0, c->stream, c->n, c->col, c->row,
                     (const float *)c->sampler->lcc, c->bc, bc_scale, (const double *)nullptr, dl, delta_scale, b->x);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

int axbc_mlp_targets(axbc_ctx *c, const uint32_t *vertex, const double *weight, uint32_t count) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  if (count && (!vertex || !weight)) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  for (uint32_t i = 0; i < count; ++i)
    if (vertex[i] >= c->n) return fail(AXBC_EINVAL, "vertex out of range%s (line %d)", "", __LINE__);
  b->targets = count > 0;
  if (!count) return AXBC_OK;
  hipStream_t s = c->stream;
  if (!b->tw) {
    {
      int rc;
      if ((rc = budget_check(c, 8ull * c->n, "the MLP target weights"))) return rc;
    }
    HIP_TRY(hipMalloc((void **)&b->tw, sizeof(double) * c->n));
    HIP_TRY(hipMalloc((void **)&b->tred, sizeof(double) * (2 * kRedBlocks + 1)));
    c->bytes += 8ull * c->n;
  }
  uint32_t *dv = nullptr;
  double *dw = nullptr;
  HIP_TRY(hipMalloc((void **)&dv, sizeof(uint32_t) * count));
  HIP_TRY(hipMalloc((void **)&dw, sizeof(double) * count));
  HIP_TRY(hipMemsetAsync(b->tw, 0, sizeof(double) * c->n, s));
  HIP_TRY(hipMemcpyAsync(dv, vertex, sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dw, weight, sizeof(double) * count, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_scatter_targets, dim3(grid_of(count)), dim3(256), 0, s, count, dv, dw, b->tw);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  (void)hipFree(dv);
  (void)hipFree(dw);
  return AXBC_OK;
}

int axbc_mlp_set_data(axbc_ctx *c, const float *x, const float *y) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  if (x) HIP_TRY(hipMemcpyAsync(b->x, x, sizeof(float) * 4 * c->n, hipMemcpyHostToDevice, c->stream));
  if (y) HIP_TRY(hipMemcpyAsync(b->y, y, sizeof(float) * c->n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_mlp_get_data(axbc_ctx *c, float *x, float *score) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  if (x) HIP_TRY(hipMemcpyAsync(x, b->x, sizeof(float) * 4 * c->n, hipMemcpyDeviceToHost, c->stream));
  if (score) HIP_TRY(hipMemcpyAsync(score, b->score, sizeof(float) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

// Gradients (summed into an FP32 host array) of the MSE over the samples.
int axbc_mlp_gradient(axbc_ctx *c, const uint32_t *idx, uint32_t m, float *grad_host, float *loss_host) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b || !idx || m == 0) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  hipStream_t s = c->stream;
  if (m > b->idx_cap) {
    if (b->idx) HIP_TRY(hipFree(b->idx));
    HIP_TRY(hipMalloc((void **)&b->idx, sizeof(uint32_t) * m));
    b->idx_cap = m;
  }
  if (!b->gpart) HIP_TRY(hipMalloc((void **)&b->gpart, sizeof(float) * kGradBlocks * kGradRow));
  HIP_TRY(hipMemcpyAsync(b->idx, idx, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
  const uint32_t tiles = (m + 63) / 64;
  const int rows = tiles < (uint32_t)kGradBlocks ? (int)tiles : kGradBlocks;
  hipLaunchKernelGGL(k_mlp_grad, dim3(rows), dim3(64), 0, s, m, b->idx, b->x, b->y, b->wb, b->gpart);
  hipLaunchKernelGGL(k_grad_reduce, dim3((kParams + 256) / 256), dim3(256), 0, s, b->gpart, rows, b->grad, b->loss);
  HIP_TRY(hipGetLastError());
  if (grad_host) HIP_TRY(hipMemcpyAsync(grad_host, b->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, s));
  if (loss_host) HIP_TRY(hipMemcpyAsync(loss_host, b->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return AXBC_OK;
}

// One Adam step with the gradients accumulated by axbc_mlp_gradient.
int axbc_mlp_adam(axbc_ctx *c, float lr) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  b->step += 1;
  hipLaunchKernelGGL(k_adam, dim3(32), dim3(256), 0, c->stream, kParams, b->master, b->wb, b->grad, b->m, b->v, lr,
                     0.9f, 0.999f, 1e-8f, b->step);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

int axbc_mlp_params(axbc_ctx *c, float *master_host) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b || !master_host) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(master_host, b->master, sizeof(float) * kParams, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_mlp_forward(axbc_ctx *c) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  const uint64_t fwd_groups = (c->n + kFwdThreads - 1) / kFwdThreads;
  hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_groups < 4096 ? (unsigned)fwd_groups : 4096), dim3(kFwdThreads), 0,
                     c->stream, c->n, b->x, b->wb, b->score);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

// Proposal distribution from the scores: w = expm1(s) + eps * mean over the
// vertices of positive degree (0 for the others), as fixed-point weights
// scanned for axbc_weights_pick (kind 7).
int axbc_mlp_proposal(axbc_ctx *c, double eps, uint64_t *total) {
  MlpBufs *b = c ? bufs_of(c) : nullptr;
  if (!b) return fail(AXBC_ESTATE, "mlp not initialised%s (line %d)", "", __LINE__);
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  double hsum[2] = {0, 0};
  const int wgroups = grid_of(c->n);
  const uint32_t waves = (uint32_t)wgroups * (256 / kWave);
  // per-wave pairs, then the total in the last two entries
  if (!b->wpart) HIP_TRY(hipMalloc((void **)&b->wpart, sizeof(double) * (2 * 4096 * (256 / kWave) + 2)));
  double *dsum = b->wpart + 2 * 4096 * (256 / kWave);
  hipLaunchKernelGGL(k_weight_sum, dim3(wgroups), dim3(256), 0, c->stream, c->n, c->col, b->score, b->wpart);
  hipLaunchKernelGGL(k_weight_total, dim3(1), dim3(256), 0, c->stream, b->wpart, waves, dsum);
  HIP_TRY(hipMemcpyAsync(hsum, dsum, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (!(hsum[1] > 0.0)) return fail(AXBC_ESTATE, "no vertex of positive degree%s (line %d)", "", __LINE__);
  double *w = b->prop;
  hipLaunchKernelGGL(k_proposal, dim3(grid_of(c->n)), dim3(256), 0, c->stream, c->n, c->col, b->score,
                     eps * hsum[0] / hsum[1], w);
  HIP_TRY(hipGetLastError());
  return axbc_weights_from(c, w, total);
}

}  // extern "C"
