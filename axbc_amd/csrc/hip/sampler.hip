// Pivot samplers on the device: distributions built with device prefix
// scans, draws resolved by a search over the scanned distribution.
//
//  - high-degree-neighbourhood tables (-z 1, src/axbc2d.c:1758-1845,
//    2758-2786, 2895-2928): pref[v] = wrapping uint32 prefix of
//    (uint32)deg(v)^2, its running maximum pmax (so that the reference's
//    "first index whose prefix >= r" over a non-monotone wrapped prefix
//    becomes a binary search over a monotone array), and sdn[v] = wrapping
//    sum of the neighbours' degrees.
//  - dynamic distributions (-z 3 bc, -z 4 delta of the last round, -z 5
//    1/(1+bc); src/axbc2d.c:2929-3060) as 64-bit fixed-point weights
//    W[v] = floor(w[v] * 2^s), one shift s = 61 - e for the whole array
//    where wmax * N = m * 2^e, so that the prefix is exact in any order.
//    A draw r in [0, total] picks the first v with prefix(v) >= r.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../../include/axbc.h"
#include "engine_internal.h"
#include "wave.h"

using namespace axbc;

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

struct AddU64 {
  __device__ static uint64_t op(uint64_t a, uint64_t b) { return a + b; }
  __device__ static uint64_t id() { return 0; }
};
struct AddU32 {
  __device__ static uint32_t op(uint32_t a, uint32_t b) { return a + b; }
  __device__ static uint32_t id() { return 0; }
};
struct MaxU32 {
  __device__ static uint32_t op(uint32_t a, uint32_t b) { return a > b ? a : b; }
  __device__ static uint32_t id() { return 0; }
};

// Block-wide inclusive scan of one value per thread (Hillis-Steele in LDS).
template <typename T, typename Op>
__device__ T block_incl_scan(T x, T *sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    T y = t >= off ? sh[t - off] : Op::id();
    __syncthreads();
    if (t >= off) sh[t] = Op::op(y, sh[t]);
    __syncthreads();
  }
  T r = sh[t];
  __syncthreads();
  return r;
}

// Pass 1: per-tile reduction.
template <typename T, typename Op>
__global__ __launch_bounds__(kScanBlock) void k_tile_reduce(const T *__restrict__ in, uint64_t n, T *tiles) {
  __shared__ T sh[kScanBlock];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  T acc = Op::id();
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) acc = Op::op(acc, in[base + k]);
  T inc = block_incl_scan<T, Op>(acc, sh);
  if (threadIdx.x == kScanBlock - 1) tiles[blockIdx.x] = inc;
}

// Pass 2: exclusive scan of the tile totals in one block (any count).
template <typename T, typename Op>
__global__ __launch_bounds__(kScanBlock) void k_tiles_scan(T *tiles, uint32_t ntiles) {
  __shared__ T sh[kScanBlock];
  __shared__ T prev[kScanBlock];
  T carry = Op::id();
  for (uint32_t b = 0; b < ntiles; b += kScanBlock) {
    const uint32_t i = b + threadIdx.x;
    const T x = i < ntiles ? tiles[i] : Op::id();
    const T inc = block_incl_scan<T, Op>(x, sh);
    prev[threadIdx.x] = inc;
    __syncthreads();
    const T before = threadIdx.x ? prev[threadIdx.x - 1] : Op::id();
    const T total = prev[kScanBlock - 1];
    if (i < ntiles) tiles[i] = Op::op(carry, before);
    __syncthreads();
    carry = Op::op(carry, total);
  }
}

// Pass 3: inclusive scan inside each tile seeded with the tile's prefix.
template <typename T, typename Op>
__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const T *__restrict__ in, uint64_t n,
                                                          const T *__restrict__ tiles, T *out) {
  __shared__ T sh[kScanBlock];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T acc = Op::id();
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : Op::id();
    acc = Op::op(acc, v[k]);
  }
  T inc = block_incl_scan<T, Op>(acc, sh);
  __shared__ T prev[kScanBlock];
  prev[threadIdx.x] = inc;
  __syncthreads();
  T run = Op::op(tiles[blockIdx.x], threadIdx.x ? prev[threadIdx.x - 1] : Op::id());
  for (int k = 0; k < kScanItems; ++k) {
    run = Op::op(run, v[k]);
    if (base + k < n) out[base + k] = run;
  }
}

template <typename T, typename Op>
int device_scan(const T *in, T *out, uint64_t n, T *tiles, hipStream_t s) {
  const uint32_t ntiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
  if (ntiles == 0) return 0;
  hipLaunchKernelGGL((k_tile_reduce<T, Op>), dim3(ntiles), dim3(kScanBlock), 0, s, in, n, tiles);
  hipLaunchKernelGGL((k_tiles_scan<T, Op>), dim3(1), dim3(kScanBlock), 0, s, tiles, ntiles);
  hipLaunchKernelGGL((k_tile_scan<T, Op>), dim3(ntiles), dim3(kScanBlock), 0, s, in, n, tiles, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---- high-degree-neighbourhood tables ------------------------------------

// Thread per vertex; vertices with more than 64 edges are summed by the
// whole wave (wrapping uint32 sums are order-free).
__global__ void k_hdn_terms(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row, uint32_t n,
                            uint32_t *deg2, uint32_t *sdn) {
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < n; base += stride) {
    const uint32_t v = base + (uint32_t)lane;
    uint32_t b = 0, e = 0;
    if (v < n) {
      b = col[v];
      e = col[v + 1];
      const uint64_t d = e - b;
      deg2[v] = (uint32_t)(d * d);  // (LOCINT)pow(deg, 2): exact square, low 32 bits
    }
    const bool big = e - b > 64;
    uint32_t s = 0;
    if (!big)
      for (uint32_t k = b; k < e; ++k) {
        const uint32_t w = row[k];
        s += col[w + 1] - col[w];
      }
    unsigned long long todo = __ballot(big);
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t bb = __shfl(b, i), ee = __shfl(e, i);
      uint32_t t = 0;
      for (uint32_t k = bb + (uint32_t)lane; k < ee; k += 64) {
        const uint32_t w = row[k];
        t += col[w + 1] - col[w];
      }
      t = (uint32_t)wave_sum_u64(t);
      if (lane == i) s = t;
    }
    if (v < n) sdn[v] = s;
  }
}

// ---- local clustering coefficient (-z 2, -z 6) ------------------------------
//
// lcc_func (src/axbc2d.c:1847-1935): for vertex i of degree d >= 2,
// counter = #{(j, x): j in adj(i), x in adj(j), x != i, x in adj(i)} as a
// wrapping uint32, lcc = (float)counter / (float)(uint32)(d*(d-1)).  Each
// (i, j) pair intersects two sorted adjacency lists: lanes walk the shorter
// one and binary-search the longer one.

__device__ __forceinline__ bool in_sorted(const uint32_t *__restrict__ a, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && a[lo] == x;
}

// |adj(i) cap adj(j) minus {i}|, one partial count per lane
__device__ __forceinline__ uint32_t pair_count(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                               uint32_t i, uint32_t j) {
  const uint32_t *ai = row + col[i], *aj = row + col[j];
  const uint32_t di = col[i + 1] - col[i], dj = col[j + 1] - col[j];
  const uint32_t *walk = dj <= di ? aj : ai, *look = dj <= di ? ai : aj;
  const uint32_t nw = dj <= di ? dj : di, nl = dj <= di ? di : dj;
  uint32_t c = 0;
  for (uint32_t k = lane_id(); k < nw; k += kWave) {
    const uint32_t x = walk[k];
    if (x != i && in_sorted(look, nl, x)) ++c;
  }
  return c;
}

__global__ __launch_bounds__(256) void k_lcc_small(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                   uint32_t n, uint32_t hub_deg, uint32_t *counter) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t i = gw; i < n; i += nwv) {
    const uint32_t d = col[i + 1] - col[i];
    if (d < 2 || d > hub_deg) continue;
    uint32_t c = 0;
    for (uint32_t k = 0; k < d; ++k) c += pair_count(col, row, i, row[col[i] + k]);
    c = (uint32_t)wave_sum_u64(c);
    if (lane_id() == 0) counter[i] = c;
  }
}

// Hubs: one wave per (hub, edge) item; items enumerated by the prefix of
// the hubs' degrees.
__global__ __launch_bounds__(256) void k_lcc_hubs(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                  const uint32_t *__restrict__ hubs, const uint64_t *__restrict__ hoff,
                                                  uint32_t nh, uint32_t *counter) {
  const uint64_t total = hoff[nh];
  const uint64_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint64_t t = gw; t < total; t += nwv) {
    uint32_t lo = 0, hi = nh;  // hoff[lo] <= t < hoff[lo+1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (hoff[mid] <= t)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t i = hubs[lo];
    const uint32_t j = row[col[i] + (uint32_t)(t - hoff[lo])];
    const uint32_t c = (uint32_t)wave_sum_u64(pair_count(col, row, i, j));
    if (lane_id() == 0 && c) atomicAdd(&counter[i], c);
  }
}

// Whether every list is strictly increasing, loop-free and mirrored (j in
// adj(i) <=> i in adj(j)); bit 0 of *bad set otherwise.  A wave per vertex.
__global__ __launch_bounds__(256) void k_lcc_check(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                   uint32_t n, unsigned int *bad) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t i = gw; i < n; i += nwv) {
    const uint32_t b = col[i], d = col[i + 1] - b;
    bool ok = true;
    for (uint32_t k = lane_id(); k < d; k += kWave) {
      const uint32_t j = row[b + k];
      ok &= j != i && j < n && (k == 0 || row[b + k - 1] < j) && in_sorted(row + col[j], col[j + 1] - col[j], i);
    }
    if (__ballot(!ok) && lane_id() == 0) atomicOr(bad, 1u);
  }
}

// Owner of an undirected edge: the endpoint of larger (degree, id).
__device__ __forceinline__ bool ranks_below(const uint32_t *__restrict__ col, uint32_t s, uint32_t o) {
  const uint32_t ds = col[s + 1] - col[s], d_o = col[o + 1] - col[o];
  return ds < d_o || (ds == d_o && s < o);
}

// Owners of degree below kLccBit: for each owned edge (o, s) lanes walk
// adj(s) and binary-search the short adj(o); both counters get the count.
__global__ __launch_bounds__(256) void k_lcc_light(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                   uint32_t n, uint32_t bit_deg, uint32_t *counter) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t o = gw; o < n; o += nwv) {
    const uint32_t b = col[o], d = col[o + 1] - b;
    if (d < 2 || d >= bit_deg) continue;
    uint32_t co = 0;
    for (uint32_t k = 0; k < d; ++k) {
      const uint32_t sv = row[b + k];
      if (!ranks_below(col, sv, o)) continue;
      const uint32_t *as = row + col[sv];
      const uint32_t ds = col[sv + 1] - col[sv];
      uint32_t c = 0;
      for (uint32_t l = lane_id(); l < ds; l += kWave) c += in_sorted(row + b, d, as[l]) ? 1u : 0u;
      c = (uint32_t)wave_sum_u64(c);
      co += c;
      if (lane_id() == 0 && c) atomicAdd(&counter[sv], c);
    }
    if (lane_id() == 0 && co) atomicAdd(&counter[o], co);
  }
}

// Owners of kLccBit .. kLccHash edges: a workgroup per owner keeps an
// open-addressing hash of adj(o) in LDS (load factor <= 1/2, linear
// probing; the lists hold distinct ids), and its waves take the owned
// neighbours s in turn, lanes walking adj(s) and probing the table.
constexpr uint32_t kLccHash = 4096;
constexpr uint32_t kHashSlots = 2 * kLccHash;
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hash_slot(uint32_t x, uint32_t mask) { return (x * 2654435761u) & mask; }

__global__ __launch_bounds__(256) void k_lcc_hash(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                  const uint32_t *__restrict__ owner, uint32_t nown,
                                                  uint32_t *counter) {
  __shared__ uint32_t tab[kHashSlots];
  const uint32_t wid = threadIdx.x / kWave, nwg = blockDim.x / kWave;
  for (uint32_t oi = blockIdx.x; oi < nown; oi += gridDim.x) {
    const uint32_t o = owner[oi];
    const uint32_t b = col[o], d = col[o + 1] - b;
    uint32_t size = 64;
    while (size < 2 * d) size <<= 1;
    const uint32_t mask = size - 1;
    for (uint32_t t = threadIdx.x; t < size; t += blockDim.x) tab[t] = kHashEmpty;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < d; k += blockDim.x) {
      const uint32_t x = row[b + k];
      uint32_t h = hash_slot(x, mask);
      while (atomicCAS(&tab[h], kHashEmpty, x) != kHashEmpty) h = (h + 1) & mask;
    }
    __syncthreads();
    uint32_t co = 0;
    for (uint32_t k = wid; k < d; k += nwg) {
      const uint32_t sv = row[b + k];
      if (!ranks_below(col, sv, o)) continue;
      uint32_t c = 0;
      for (uint32_t l = col[sv] + lane_id(); l < col[sv + 1]; l += kWave) {
        const uint32_t x = row[l];
        uint32_t h = hash_slot(x, mask);
        uint32_t y;
        while ((y = tab[h]) != kHashEmpty && y != x) h = (h + 1) & mask;
        c += y == x ? 1u : 0u;
      }
      c = (uint32_t)wave_sum_u64(c);
      co += c;
      if (lane_id() == 0 && c) atomicAdd(&counter[sv], c);
    }
    if (lane_id() == 0 && co) atomicAdd(&counter[o], co);
    __syncthreads();
  }
}

// Bitmaps of the adjacency of a batch of owners (one N-bit row per slot):
// set (on = 1) or clear (on = 0) the words their neighbours touch.
__global__ __launch_bounds__(256) void k_lcc_bits(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                  const uint32_t *__restrict__ owner, uint32_t nslots, uint64_t words,
                                                  uint32_t *bits, int on) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t sl = gw; sl < nslots; sl += nwv) {
    const uint32_t o = owner[sl];
    uint32_t *bm = bits + (size_t)sl * words;
    for (uint32_t k = col[o] + lane_id(); k < col[o + 1]; k += kWave) {
      const uint32_t x = row[k];
      if (on)
        atomicOr(&bm[x >> 5], 1u << (x & 31));
      else
        bm[x >> 5] = 0u;
    }
  }
}

// Owned edges of a batch of bitmap owners, one wave per (slot, neighbour)
// item: lanes walk adj(s) and test the owner's bits.
__global__ __launch_bounds__(256) void k_lcc_heavy(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                   const uint32_t *__restrict__ owner,
                                                   const uint64_t *__restrict__ ioff, uint32_t nslots, uint64_t words,
                                                   const uint32_t *__restrict__ bits, uint32_t *counter) {
  const uint64_t total = ioff[nslots] - ioff[0];
  const uint64_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint64_t t = gw; t < total; t += nwv) {
    const uint64_t tt = t + ioff[0];
    uint32_t lo = 0, hi = nslots;  // ioff[lo] <= tt < ioff[lo+1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ioff[mid] <= tt)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t o = owner[lo];
    const uint32_t sv = row[col[o] + (uint32_t)(tt - ioff[lo])];
    if (!ranks_below(col, sv, o)) continue;
    const uint32_t *bm = bits + (size_t)lo * words;
    uint32_t c = 0;
    for (uint32_t k = col[sv] + lane_id(); k < col[sv + 1]; k += kWave) {
      const uint32_t x = row[k];
      c += (bm[x >> 5] >> (x & 31)) & 1u;
    }
    c = (uint32_t)wave_sum_u64(c);
    if (lane_id() == 0 && c) {
      atomicAdd(&counter[o], c);
      atomicAdd(&counter[sv], c);
    }
  }
}

__global__ void k_lcc_value(const uint32_t *__restrict__ col, uint32_t n, const uint32_t *__restrict__ counter,
                            float *lcc) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t d = col[i + 1] - col[i];
    lcc[i] = d < 2 ? 0.0f : (float)counter[i] / (float)(uint32_t)(d * (d - 1u));
  }
}

// ---- dynamic fixed-point weights -----------------------------------------

__device__ __forceinline__ double weight_of(int kind, uint32_t v, const double *__restrict__ bc,
                                            const double *__restrict__ delta) {
  // kind 2 passes the LCC values (float) through `delta` reinterpreted,
  // kind 7 passes arbitrary weights (double) through `delta`
  double w = kind == 2 ? (double)((const float *)delta)[v]
             : kind == 3 ? bc[v] : kind == 4 || kind == 7 ? delta[v] : 1.0 / (1.0 + bc[v]);
  return (isfinite(w) && w > 0.0) ? w : 0.0;
}

__global__ void k_weight_max(int kind, const double *__restrict__ bc, const double *__restrict__ delta,
                             uint32_t n, unsigned long long *wmax_bits) {
  double m = 0.0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    m = fmax(m, weight_of(kind, v, bc, delta));
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  // non-negative doubles order like their bit patterns
  if (lane_id() == 0) atomicMax(wmax_bits, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_weight_fixed(int kind, const double *__restrict__ bc, const double *__restrict__ delta,
                               uint32_t n, const unsigned long long *__restrict__ wmax_bits, uint64_t *W) {
  const double wmax = __longlong_as_double((long long)*wmax_bits);
  int e = 0;
  frexp(wmax * (double)n, &e);
  const int s = 61 - e;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const double w = wmax > 0.0 ? weight_of(kind, v, bc, delta) : 0.0;
    W[v] = w > 0.0 ? (uint64_t)ldexp(w, s) : 0ull;
  }
}

// First index with P[i] >= r over a non-decreasing array; n-1 when none.
template <typename T>
__global__ void k_first_geq(const T *__restrict__ P, uint32_t n, T r, uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t lo = 0, hi = n;  // answer in [lo, hi]
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (P[mid] >= r)
      hi = mid;
    else
      lo = mid + 1;
  }
  *out = lo < n ? lo : n - 1;
}

// Draw windows: lrand48 at positions 0..count-1 ahead of state X (each
// value advances X = (0x5DEECE66D X + 0xB) mod 2^48 and is X >> 17; the map
// to position p is the affine map raised to p + 1 by squaring, in 64-bit
// arithmetic -- exact mod 2^48), and the vertex a first draw there picks:
// table 1, r = lr % (total + 1) as uint32 (uniform_int(total, 0), the raw
// value when total + 1 wraps to 0) and the first i with the running maximum
// of the wrapped deg^2 prefix >= r; table 2, r = lr / 2147483647 * total
// (uniform_double, clamped to total) and the first i with the LCC weight
// prefix >= r; n-1 when none.
constexpr uint64_t kLcgA = 0x5DEECE66Dull, kLcgC = 0xBull, kLcgMask = (1ull << 48) - 1;

__global__ void k_draw_window(int table, uint64_t state, uint64_t total, uint32_t count,
                              const uint32_t *__restrict__ pmax, const uint64_t *__restrict__ lp, uint32_t n,
                              uint32_t *lr_out, uint32_t *v_out) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += gridDim.x * blockDim.x) {
    uint64_t A = 1, Cc = 0, ap = kLcgA, cp = kLcgC;
    for (uint64_t e = (uint64_t)p + 1; e; e >>= 1) {
      if (e & 1) {
        A = ap * A;
        Cc = ap * Cc + cp;
      }
      cp = ap * cp + cp;
      ap = ap * ap;
    }
    const uint64_t x = (A * state + Cc) & kLcgMask;
    const uint32_t lr = (uint32_t)(x >> 17);
    uint32_t lo = 0, hi = n;
    if (table == 1) {
      const uint32_t m = (uint32_t)total + 1u;
      const uint32_t r = m ? lr % m : lr;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (pmax[mid] >= r)
          hi = mid;
        else
          lo = mid + 1;
      }
    } else {
      const double rd = (double)lr / 2147483647.0 * (double)total;
      const uint64_t r = rd >= (double)total ? total : (uint64_t)rd;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (lp[mid] >= r)
          hi = mid;
        else
          lo = mid + 1;
      }
    }
    lr_out[p] = lr;
    v_out[p] = lo < n ? lo : n - 1;
  }
}

int grid_n(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

}  // namespace

namespace axbc {

int sampler_bufs(axbc_ctx *c, SamplerBufs **out) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  if (c->sampler) {
    *out = c->sampler;
    return AXBC_OK;
  }
  HIP_TRY(hipSetDevice(c->device));
  const size_t n = c->n;
  const size_t ntiles = (n + kScanTile - 1) / kScanTile + 1;
  {
    int rc;
    if ((rc = budget_check(c, 28 * n + 8 * ntiles, "the sampler tables"))) return rc;
  }
  SamplerBufs *sb = new SamplerBufs();
  c->sampler = sb;
  HIP_TRY(hipMalloc(&sb->tmp, 8 * n));
  HIP_TRY(hipMalloc(&sb->u32a, 4 * n));
  HIP_TRY(hipMalloc(&sb->u32b, 4 * n));
  HIP_TRY(hipMalloc(&sb->u32c, 4 * n));
  HIP_TRY(hipMalloc(&sb->u64a, 8 * n));
  HIP_TRY(hipMalloc(&sb->tiles, 8 * ntiles));
  HIP_TRY(hipMalloc((void **)&sb->wmax, sizeof(unsigned long long)));
  HIP_TRY(hipMalloc((void **)&sb->pick, sizeof(uint32_t)));
  HIP_TRY(hipHostMalloc((void **)&sb->pick_host, sizeof(uint32_t), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void **)&sb->host64, 2 * sizeof(uint64_t), hipHostMallocDefault));
  c->bytes += 28 * n + 8 * ntiles;
  *out = sb;
  return AXBC_OK;
}

void sampler_free(axbc_ctx *c) {
  SamplerBufs *sb = c->sampler;
  if (!sb) return;
  void *d[] = {sb->tmp, sb->u32a, sb->u32b, sb->u32c, sb->u64a, sb->tiles, sb->wmax, sb->pick,
               sb->lcc, sb->lcc_cnt, sb->lw, sb->lp};
  for (void *p : d)
    if (p) (void)hipFree(p);
  if (sb->pick_host) (void)hipHostFree(sb->pick_host);
  if (sb->host64) (void)hipHostFree(sb->host64);
  if (sb->wstream) {
    (void)hipStreamSynchronize(sb->wstream);
    (void)hipStreamDestroy(sb->wstream);
  }
  if (sb->win) {
    (void)hipFree(sb->win);
    c->bytes -= std::min<size_t>(c->bytes, 8ull * sb->win_cap);
  }
  delete sb;
  c->sampler = nullptr;
}

}  // namespace axbc

extern "C" {

int axbc_hdn_build(axbc_ctx *c, uint32_t *pmax_host, uint32_t *sdn_host, uint32_t *sum) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if (!sum) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  const uint32_t n = c->n;
  hipStream_t s = c->stream;
  uint32_t *deg2 = (uint32_t *)sb->tmp, *pref = (uint32_t *)sb->u32a, *pmax = (uint32_t *)sb->u32b;
  uint32_t *sdn = (uint32_t *)sb->u32c;
  hipLaunchKernelGGL(k_hdn_terms, dim3(grid_n(n)), dim3(256), 0, s, c->col, c->row, n, deg2, sdn);
  if (device_scan<uint32_t, AddU32>(deg2, pref, n, (uint32_t *)sb->tiles, s))
    return fail(AXBC_EHIP, "scan launch failed%s (line %d)", "", __LINE__);
  if (device_scan<uint32_t, MaxU32>(pref, pmax, n, (uint32_t *)sb->tiles, s))
    return fail(AXBC_EHIP, "scan launch failed%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(&sb->host64[0], pref + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (pmax_host) HIP_TRY(hipMemcpyAsync(pmax_host, pmax, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  if (sdn_host) HIP_TRY(hipMemcpyAsync(sdn_host, sdn, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *sum = (uint32_t)(sb->host64[0] & 0xFFFFFFFFull);
  sb->hdn_ready = 1;
  return AXBC_OK;
}

int axbc_hdn_pick(axbc_ctx *c, uint32_t r, uint32_t *u, uint32_t *sdn_u) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if (!sb->hdn_ready || !u) return fail(AXBC_ESTATE, "tables not built%s (line %d)", "", __LINE__);
  hipStream_t s = c->stream;
  hipLaunchKernelGGL((k_first_geq<uint32_t>), dim3(1), dim3(64), 0, s, (const uint32_t *)sb->u32b, c->n, r,
                     sb->pick);
  HIP_TRY(hipMemcpyAsync(sb->pick_host, sb->pick, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *u = *sb->pick_host;
  if (sdn_u) {
    HIP_TRY(hipMemcpyAsync(sb->pick_host, (const uint32_t *)sb->u32c + *u, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *sdn_u = *sb->pick_host;
  }
  return AXBC_OK;
}

int axbc_draw_window(axbc_ctx *c, int table, uint64_t state, uint64_t total, uint32_t count, uint32_t *lr,
                     uint32_t *v) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if ((table != 1 && table != 2) || !lr || !v || count == 0 || count > (1u << 20))
    return fail(AXBC_EINVAL, "table 1 or 2, 1..2^20 draws%s (line %d)", "", __LINE__);
  if ((table == 1 && !sb->hdn_ready) || (table == 2 && !sb->lcc_ready))
    return fail(AXBC_ESTATE, "tables not built%s (line %d)", "", __LINE__);
  if (state >> 48) return fail(AXBC_EINVAL, "a 48-bit generator state%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(c->device));
  if (!sb->wstream) HIP_TRY(hipStreamCreateWithFlags(&sb->wstream, hipStreamNonBlocking));
  if (count > sb->win_cap) {
    const uint32_t cap = std::max<uint32_t>(count, 1024);
    const size_t old = 8ull * sb->win_cap;
    if ((rc = budget_check(c, 8ull * cap - old, "the draw window"))) return rc;
    if (sb->win) HIP_TRY(hipFree(sb->win));
    sb->win = nullptr;
    sb->win_cap = 0;
    c->bytes -= std::min<size_t>(c->bytes, old);
    HIP_TRY(hipMalloc((void **)&sb->win, 8ull * cap));
    sb->win_cap = cap;
    c->bytes += 8ull * cap;
  }
  hipStream_t s = sb->wstream;
  hipLaunchKernelGGL(k_draw_window, dim3(grid_n(count)), dim3(256), 0, s, table, state, total, count,
                     (const uint32_t *)sb->u32b, (const uint64_t *)sb->lp, c->n, sb->win, sb->win + sb->win_cap);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(lr, sb->win, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(v, sb->win + sb->win_cap, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return AXBC_OK;
}

static int weights_impl(axbc_ctx *c, int kind, const double *src, uint64_t *total);

int axbc_weights_from(axbc_ctx *c, const double *w_device, uint64_t *total) {
  return weights_impl(c, 7, w_device, total);
}

int axbc_weights(axbc_ctx *c, int kind, uint64_t *total) {
  if (kind != 3 && kind != 4 && kind != 5) return fail(AXBC_EINVAL, "weight kind%s (line %d)", "", __LINE__);
  return weights_impl(c, kind, c ? c->delta : nullptr, total);
}

static int weights_impl(axbc_ctx *c, int kind, const double *src, uint64_t *total) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  const uint32_t n = c->n;
  hipStream_t s = c->stream;
  uint64_t *W = (uint64_t *)sb->tmp, *P = (uint64_t *)sb->u64a;
  HIP_TRY(hipMemsetAsync(sb->wmax, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_weight_max, dim3(grid_n(n)), dim3(256), 0, s, kind, c->bc, src, n, sb->wmax);
  hipLaunchKernelGGL(k_weight_fixed, dim3(grid_n(n)), dim3(256), 0, s, kind, c->bc, src, n, sb->wmax, W);
  if (device_scan<uint64_t, AddU64>(W, P, n, (uint64_t *)sb->tiles, s))
    return fail(AXBC_EHIP, "scan launch failed%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(&sb->host64[0], P + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  sb->total = sb->host64[0];
  if (total) *total = sb->total;
  sb->weights_ready = 1;
  return AXBC_OK;
}

int axbc_lcc_build(axbc_ctx *c, float *lcc_host, uint64_t *lw_host, uint64_t *lp_host, uint64_t *total) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  const uint32_t n = c->n;
  hipStream_t s = c->stream;
  if (!sb->lcc) {
    if ((rc = budget_check(c, 24ull * n, "the LCC tables"))) return rc;
    HIP_TRY(hipMalloc((void **)&sb->lcc, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&sb->lcc_cnt, sizeof(uint32_t) * n));
    HIP_TRY(hipMalloc((void **)&sb->lw, sizeof(uint64_t) * n));
    HIP_TRY(hipMalloc((void **)&sb->lp, sizeof(uint64_t) * n));
    c->bytes += 24ull * n;
  }
  constexpr uint32_t kLccHub = 4096;
  std::vector<uint32_t> hubs;
  std::vector<uint64_t> hoff(1, 0);
  for (uint32_t v = 0; v < n; ++v) {
    const uint32_t d = c->h_col[v + 1] - c->h_col[v];
    if (d > kLccHub) {
      hubs.push_back(v);
      hoff.push_back(hoff.back() + d);
    }
  }
  uint32_t *dh = nullptr;
  uint64_t *doff = nullptr;
  HIP_TRY(hipMalloc((void **)&dh, sizeof(uint32_t) * (hubs.size() + 1)));
  HIP_TRY(hipMalloc((void **)&doff, sizeof(uint64_t) * hoff.size()));
  if (!hubs.empty()) HIP_TRY(hipMemcpyAsync(dh, hubs.data(), sizeof(uint32_t) * hubs.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(doff, hoff.data(), sizeof(uint64_t) * hoff.size(), hipMemcpyHostToDevice, s));
  if (sb->lcc_sym < 0) {  // once per context: may every edge be intersected once?
    HIP_TRY(hipMemsetAsync(sb->lcc_cnt, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_lcc_check, dim3(4096), dim3(256), 0, s, c->col, c->row, n, sb->lcc_cnt);
    uint32_t bad = 1;
    HIP_TRY(hipMemcpyAsync(&bad, sb->lcc_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    sb->lcc_sym = bad ? 0 : 1;
  }
  HIP_TRY(hipMemsetAsync(sb->lcc_cnt, 0, sizeof(uint32_t) * n, s));
  if (sb->lcc_sym == 1) {
    // every undirected edge once, from its owner (larger degree, id):
    // short owners binary-search, owners of kLccBit .. kLccHash edges probe
    // an LDS hash of their adjacency, larger ones test a bitmap of it, built
    // for a batch of them at a time
    constexpr uint32_t kLccBit = 256;
    constexpr uint64_t kBitBytes = 512ull << 20;
    hipLaunchKernelGGL(k_lcc_light, dim3(4096), dim3(256), 0, s, c->col, c->row, n, kLccBit, sb->lcc_cnt);
    // the LDS hash pays once an owner's bitmap (n bits) no longer stays in
    // cache: from 2^21 vertices (256 KB bitmaps) on
    const uint32_t hash_max = n >= (1u << 21) ? kLccHash : kLccBit - 1;
    std::vector<uint32_t> heavy, mid;
    std::vector<uint64_t> ioff(1, 0);
    for (uint32_t v = 0; v < n; ++v) {
      const uint32_t d = c->h_col[v + 1] - c->h_col[v];
      if (d > hash_max) {
        heavy.push_back(v);
        ioff.push_back(ioff.back() + d);
      } else if (d >= kLccBit) {
        mid.push_back(v);
      }
    }
    if (!mid.empty()) {
      uint32_t *dm = nullptr;
      if (hipMalloc((void **)&dm, sizeof(uint32_t) * mid.size()) != hipSuccess) {
        (void)hipFree(dh);
        (void)hipFree(doff);
        return fail(AXBC_ENOMEM, "LCC owner list%s (line %d)", "", __LINE__);
      }
      HIP_TRY(hipMemcpyAsync(dm, mid.data(), sizeof(uint32_t) * mid.size(), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_lcc_hash, dim3((uint32_t)std::min<size_t>(mid.size(), 2048)), dim3(256), 0, s, c->col,
                         c->row, dm, (uint32_t)mid.size(), sb->lcc_cnt);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(s));
      (void)hipFree(dm);
    }
    if (!heavy.empty()) {
      const uint64_t words = ((uint64_t)n + 31) / 32;
      const uint32_t slots = (uint32_t)std::max<uint64_t>(
          1, std::min<uint64_t>(heavy.size(), kBitBytes / (words * sizeof(uint32_t))));
      uint32_t *dhv = nullptr, *bits = nullptr;
      uint64_t *dio = nullptr;
      hipError_t e1 = hipMalloc((void **)&dhv, sizeof(uint32_t) * heavy.size());
      hipError_t e2 = hipMalloc((void **)&dio, sizeof(uint64_t) * ioff.size());
      hipError_t e3 = hipMalloc((void **)&bits, sizeof(uint32_t) * words * slots);
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (dhv) (void)hipFree(dhv);
        if (dio) (void)hipFree(dio);
        if (bits) (void)hipFree(bits);
        (void)hipFree(dh);
        (void)hipFree(doff);
        return fail(AXBC_ENOMEM, "LCC bitmaps%s (line %d)", "", __LINE__);
      }
      HIP_TRY(hipMemcpyAsync(dhv, heavy.data(), sizeof(uint32_t) * heavy.size(), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(dio, ioff.data(), sizeof(uint64_t) * ioff.size(), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemsetAsync(bits, 0, sizeof(uint32_t) * words * slots, s));
      for (size_t b0 = 0; b0 < heavy.size(); b0 += slots) {
        const uint32_t nb = (uint32_t)std::min<size_t>(slots, heavy.size() - b0);
        hipLaunchKernelGGL(k_lcc_bits, dim3(grid_n((uint64_t)nb * kWave)), dim3(256), 0, s, c->col, c->row,
                           dhv + b0, nb, words, bits, 1);
        hipLaunchKernelGGL(k_lcc_heavy, dim3(4096), dim3(256), 0, s, c->col, c->row, dhv + b0, dio + b0, nb, words,
                           bits, sb->lcc_cnt);
        hipLaunchKernelGGL(k_lcc_bits, dim3(grid_n((uint64_t)nb * kWave)), dim3(256), 0, s, c->col, c->row,
                           dhv + b0, nb, words, bits, 0);
      }
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(s));
      (void)hipFree(dhv);
      (void)hipFree(dio);
      (void)hipFree(bits);
    }
  } else {
    hipLaunchKernelGGL(k_lcc_small, dim3(4096), dim3(256), 0, s, c->col, c->row, n, kLccHub, sb->lcc_cnt);
    if (!hubs.empty())
      hipLaunchKernelGGL(k_lcc_hubs, dim3(4096), dim3(256), 0, s, c->col, c->row, dh, doff, (uint32_t)hubs.size(),
                         sb->lcc_cnt);
  }
  hipLaunchKernelGGL(k_lcc_value, dim3(grid_n(n)), dim3(256), 0, s, c->col, n, sb->lcc_cnt, sb->lcc);
  // fixed-point weights of the LCC values and their prefix
  HIP_TRY(hipMemsetAsync(sb->wmax, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_weight_max, dim3(grid_n(n)), dim3(256), 0, s, 2, c->bc, (const double *)sb->lcc, n, sb->wmax);
  hipLaunchKernelGGL(k_weight_fixed, dim3(grid_n(n)), dim3(256), 0, s, 2, c->bc, (const double *)sb->lcc, n, sb->wmax,
                     sb->lw);
  if (device_scan<uint64_t, AddU64>(sb->lw, sb->lp, n, (uint64_t *)sb->tiles, s))
    return fail(AXBC_EHIP, "scan launch failed%s (line %d)", "", __LINE__);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&sb->host64[0], sb->lp + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  if (lcc_host) HIP_TRY(hipMemcpyAsync(lcc_host, sb->lcc, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  if (lw_host) HIP_TRY(hipMemcpyAsync(lw_host, sb->lw, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
  if (lp_host) HIP_TRY(hipMemcpyAsync(lp_host, sb->lp, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipFree(dh));
  HIP_TRY(hipFree(doff));
  sb->ltotal = sb->host64[0];
  if (total) *total = sb->ltotal;
  sb->lcc_ready = 1;
  return AXBC_OK;
}

int axbc_weights_at(axbc_ctx *c, uint32_t v, uint64_t *w) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if (!sb->weights_ready || !w || v >= c->n) return fail(AXBC_ESTATE, "weights not built%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(&sb->host64[1], (const uint64_t *)sb->tmp + v, sizeof(uint64_t), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *w = sb->host64[1];
  return AXBC_OK;
}

int axbc_weights_copy(axbc_ctx *c, uint64_t *w) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if (!sb->weights_ready || !w) return fail(AXBC_ESTATE, "weights not built%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(w, sb->tmp, sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_weights_pick(axbc_ctx *c, uint64_t r, uint32_t *v, double *p) {
  SamplerBufs *sb = nullptr;
  int rc = sampler_bufs(c, &sb);
  if (rc) return rc;
  if (!sb->weights_ready || !v || !p) return fail(AXBC_ESTATE, "weights not built%s (line %d)", "", __LINE__);
  hipStream_t s = c->stream;
  const uint64_t *W = (const uint64_t *)sb->tmp, *P = (const uint64_t *)sb->u64a;
  hipLaunchKernelGGL((k_first_geq<uint64_t>), dim3(1), dim3(64), 0, s, P, c->n, r, sb->pick);
  HIP_TRY(hipMemcpyAsync(sb->pick_host, sb->pick, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *v = *sb->pick_host;
  HIP_TRY(hipMemcpyAsync(&sb->host64[1], W + *v, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *p = (double)sb->host64[1] / (double)sb->total;
  return AXBC_OK;
}

}  // extern "C"
