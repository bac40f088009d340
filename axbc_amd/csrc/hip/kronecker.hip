// Kronecker (R-MAT) edge generation on the device, with the output of the
// host generator (kronecker.c) element for element, followed by the device
// preprocessing of ingest.hip.
//
// The algorithm is the Graph500 generator's as AxBC uses it
// (src/axbc2d.c:518-527 -> generator/make_graph.c:271-353):
//   - the L'Ecuyer-Blouin-Couture MRG (m = 2^31-1, multipliers 107374182 and
//     104480) with random access by skipping the state by A^e mod m
//     (generator/splittable_mrg.c:22-60, 217-231); the skip table
//     A^(v * 256^b) is built here by repeated multiplication;
//   - the 2x2 recursion: every node re-derives its stream from the root
//     stream skipped by (0, (src+N)/n, (tgt+N)/n), draws its four quadrant
//     counts (one categorical draw per edge up to 20 edges, otherwise three
//     sequential binomials) and folds the lower-left quadrant onto the
//     upper-right one on the diagonal (generator/graph_generator.c:95-260);
//     a node with one edge walks down by categorical draws; a node of one
//     vertex emits its edges;
//   - the rand_sort vertex permutation (generator/permutation_gen.c:148-224):
//     per vertex a key in [0, 2n+128) from stream (1, i, 0), a stable sort
//     by key, and a Fisher-Yates shuffle of every bucket from stream
//     (1, key, 100);
//   - (min, max) orientation of every permuted edge
//     (generator/make_graph.c:63-75).
// The recursion runs breadth-first, one thread per node with at least two
// edges; every node carries the index of its first edge, so edge positions
// do not depend on the order in which the threads run.
//
// Floating point: the categorical and binomial draws compare doubles, so
// this file is compiled without contraction of a*b+c into fused
// multiply-adds, and every operation rounds as the host's SSE2 arithmetic
// does.  The binomial draws call log(): the device's log and glibc's may
// differ in the last bit for some arguments, which changes a draw only
// when a comparison falls within that bit; the tests compare the device
// output with the host generator's and with the reference generator's
// fixtures.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../../include/axbc.h"
#include "engine_internal.h"

using namespace axbc;

namespace {

constexpr uint64_t kM = 0x7FFFFFFFull;
constexpr uint64_t kMul1 = 107374182ull;
constexpr uint64_t kMul5 = 104480ull;
constexpr int kSkipTables = 24;  // byte positions of a 192-bit exponent

struct Mrg {
  uint32_t z[5];
};

__host__ __device__ __forceinline__ uint32_t fold(uint64_t t) {
  t = (t & kM) + (t >> 31);
  t = (t & kM) + (t >> 31);
  return (uint32_t)(t >= kM ? t - kM : t);
}

// ---- host: the skip table -------------------------------------------------

struct Mat5 {
  uint32_t a[5][5];
};

Mat5 mat5_mul(const Mat5 &x, const Mat5 &y) {
  Mat5 r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 5; ++k) acc = fold(acc + (uint64_t)x.a[i][k] * y.a[k][j]);
      r.a[i][j] = (uint32_t)acc;
    }
  return r;
}

// tab[(b * 256 + v) * 25 + 5 i + k] = (A^(v * 256^b))[i][k]
std::vector<uint32_t> skip_table() {
  std::vector<uint32_t> tab((size_t)kSkipTables * 256 * 25);
  Mat5 step = {};
  step.a[0][0] = (uint32_t)kMul1;
  step.a[0][4] = (uint32_t)kMul5;
  for (int i = 1; i < 5; ++i) step.a[i][i - 1] = 1;
  for (int b = 0; b < kSkipTables; ++b) {
    Mat5 cur = {};
    for (int i = 0; i < 5; ++i) cur.a[i][i] = 1;
    for (int v = 0; v < 256; ++v) {
      std::copy(&cur.a[0][0], &cur.a[0][0] + 25, &tab[((size_t)b * 256 + v) * 25]);
      cur = mat5_mul(cur, step);
    }
    step = cur;  // A^(256^(b+1)) = (A^(256^b))^256
  }
  return tab;
}

Mrg seed_from_user(uint64_t seed1, uint64_t seed2) {
  Mrg s;
  s.z[0] = (uint32_t)((seed1 & 0x3FFFFFFF) + 1);
  s.z[1] = (uint32_t)(((seed1 >> 30) & 0x3FFFFFFF) + 1);
  s.z[2] = (uint32_t)((seed2 & 0x3FFFFFFF) + 1);
  s.z[3] = (uint32_t)(((seed2 >> 30) & 0x3FFFFFFF) + 1);
  s.z[4] = (uint32_t)(((seed2 >> 60) << 4) + (seed1 >> 60) + 1);
  return s;
}

// ---- device: the stream --------------------------------------------------

__device__ __forceinline__ void mrg_apply(const uint32_t *__restrict__ m, Mrg &st) {
  uint32_t r[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) acc = fold(acc + (uint64_t)m[5 * i + k] * st.z[k]);
    r[i] = (uint32_t)acc;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) st.z[i] = r[i];
}

// state * A^(e_high * 2^128 + e_mid * 2^64 + e_low)
__device__ void mrg_skip(const uint32_t *__restrict__ tab, Mrg &st, uint64_t e_high, uint64_t e_mid,
                         uint64_t e_low) {
  const uint64_t words[3] = {e_low, e_mid, e_high};
  for (int w = 0; w < 3; ++w) {
    uint64_t e = words[w];
    for (int byte = 0; e; ++byte, e >>= 8) {
      const unsigned v = (unsigned)(e & 0xFF);
      if (v) mrg_apply(tab + ((size_t)(w * 8 + byte) * 256 + v) * 25, st);
    }
  }
}

__device__ __forceinline__ uint32_t mrg_uint(Mrg &st) {
  const uint32_t fresh = fold(kMul1 * st.z[0] + kMul5 * (uint64_t)st.z[4]);
  st.z[4] = st.z[3];
  st.z[3] = st.z[2];
  st.z[2] = st.z[1];
  st.z[1] = st.z[0];
  st.z[0] = fresh;
  return fresh;
}

__device__ __forceinline__ double mrg_double(Mrg &st) {
  const double inv = 1.0 / 2147483647.0;
  const double inv2 = (1.0 / 2147483647.0) * (1.0 / 2147483647.0);
  const double hi = (double)mrg_uint(st);
  const double lo = (double)mrg_uint(st);
  return hi * inv + lo * inv2;
}

// unbiased draw in [0, n) by rejection
__device__ uint64_t mrg_up_to(Mrg &st, uint64_t n) {
  if (n <= 1) return 0;
  if (n <= kM) {
    const uint64_t limit = (kM / n) * n;
    for (;;) {
      const uint64_t x = mrg_uint(st);
      if (x < limit) return x % n;
    }
  }
  const uint64_t span = 0x3FFFFFFF00000001ull;  // (2^31-1)^2
  const uint64_t limit = (span / n) * n;
  for (;;) {
    uint64_t x = (uint64_t)mrg_uint(st) * kM;
    x += mrg_uint(st);
    if (x < limit) return x % n;
  }
}

// ---- device: binomial variates ---------------------------------------------

// Stirling-series correction of ln k! (Hoermann 1993, Table 1 for k <= 9).
__device__ double stirling_fc(int k) {
  switch (k) {
    case 0: return 0.08106146679532726;
    case 1: return 0.04134069595540929;
    case 2: return 0.02767792568499834;
    case 3: return 0.02079067210376509;
    case 4: return 0.01664469118982119;
    case 5: return 0.01387612882307075;
    case 6: return 0.01189670994589177;
    case 7: return 0.01041126526197209;
    case 8: return 0.009255462182712733;
    case 9: return 0.008330563433362871;
    default: break;
  }
  const double c12 = 1. / 12, c360 = 1. / 360, c1260 = 1. / 1260;
  const double q = 1. / (k + 1);
  return (c12 - (c360 - c1260 * q * q) * q * q) * q;
}

// Geometric waiting times (Devroye ch. X) for n p < 10.
__device__ int64_t binomial_geometric(uint64_t n, double p, Mrg &st) {
  const double r = 1. - 2. / p;
  const double ir = 1. / r;
  const double ir2 = ir * ir;
  const double ir4 = ir2 * ir2;
  const double ir6 = ir4 * ir2;
  const double ln1mp = 2. * ir * (1 + ir2 / 3. + ir4 / 5. + ir6 / 7.);
  const double inv_ln1mp = 1. / ln1mp;
  int count = 0, total = 0;
  do {
    total += (int)ceil(log(mrg_double(st)) * inv_ln1mp);
    ++count;
  } while ((uint64_t)total <= n);
  return count - 1;
}

// BTRD transformed rejection (Hoermann 1993, steps 0-3.4), n <= 10^9, n p >= 10;
// step 1 divides by (0.5 + |u|) as the reference generator does.
__device__ int64_t binomial_btrd(int n, double p, Mrg &st) {
  const int m = (int)floor((n + 1) * p);
  const double r = p / (1. - p);
  const double nr = (n + 1) * r;
  const double npq = n * p * (1. - p);
  const double sq = sqrt(npq);
  const double b = 1.15 + 2.53 * sq;
  const double a = -.0873 + .0248 * b + .01 * p;
  const double c = n * p + .5;
  const double alpha = (2.83 + 5.1 / b) * sq;
  const double vr = .92 - 4.2 / b;
  const double urvr = .86 * vr;
  for (;;) {
    double v = mrg_double(st);
    if (v <= urvr) {
      const double u = v / vr - .43;
      return (int)floor((2 * a / (.5 + fabs(u)) + b) * u + c);
    }
    const double w = mrg_double(st);
    double u;
    if (v >= vr) {
      u = w - .5;
    } else {
      u = v / vr - .93;
      u = (u > 0. ? 1. : u < 0. ? -1. : 0.) * .5 - u;
      v = vr * w;
    }
    const double us = .5 - fabs(u);
    const int k = (int)floor((2 * a / us + b) * u + c);
    if (k < 0 || k > n) continue;
    v *= alpha / (a / (us * us) + b);
    const int km = k >= m ? k - m : m - k;
    if (km > 15) {
      const double rho = (km / npq) * (((km / 3 + .625) * km + 1. / 6) / npq + .5);
      const double t = -km * km / (2 * npq);
      v = log(v);
      if (v < t - rho) return k;
      if (v > t + rho) continue;
      const int nm = n - m + 1;
      const double h = (m + .5) * log((m + 1) / (r * nm)) + stirling_fc(m) + stirling_fc(n - m);
      const int nk = n - k + 1;
      const double bound = h + (n + 1) * log((double)nm / nk) + (k + .5) * log(nk * r / (k + 1)) -
                           stirling_fc(k) - stirling_fc(n - k);
      if (v <= bound) return k;
    } else {
      double f = 1.;
      if (m < k) {
        for (int i = m; i != k; ++i) f *= nr / i - r;
      } else if (m > k) {
        for (int i = k; i != m; ++i) v *= nr / i + r;
      }
      if (v <= f) return k;
    }
  }
}

// Binomial(n, p) with the host generator's case split: p > 1/2 by symmetry,
// n p < 10 by waiting times, n > 10^9 in parts of 10^9 from the left.
__device__ int64_t binomial(int64_t n_in, double p, Mrg &st) {
  if (p == 0.) return 0;
  bool flip = false;
  if (p > .5) {
    flip = true;
    p = 1. - p;
    if (p == 0.) return n_in;
  }
  uint64_t n = (uint64_t)n_in;
  int64_t got = 0;
  for (;;) {
    if ((double)n * p < 10) {
      got += binomial_geometric(n, p, st);
      break;
    }
    if (n > 1000000000ull) {
      // the left part of 10^9 is a binomial of its own (its own n p test)
      if ((double)1000000000ull * p < 10)
        got += binomial_geometric(1000000000ull, p, st);
      else
        got += binomial_btrd(1000000000, p, st);
      n -= 1000000000ull;
      continue;
    }
    got += binomial_btrd((int)n, p, st);
    break;
  }
  return flip ? n_in - got : got;
}

// ---- device: the recursion ---------------------------------------------------

struct GenNode {
  uint64_t first, count;
  uint32_t src, tgt;
};

struct GenParams {
  Mrg root;
  double ini[4];
  uint64_t total;  // vertices (2^scale)
};

__device__ __forceinline__ int categorical4(const GenParams &g, Mrg &st) {
  double r = mrg_double(st);
  for (int j = 0; j < 3; ++j) {
    if (r < g.ini[j]) return j;
    r -= g.ini[j];
  }
  return 3;
}

__device__ __forceinline__ Mrg node_state(const GenParams &g, const uint32_t *__restrict__ tab, uint64_t src,
                                          uint64_t tgt, uint64_t nverts) {
  Mrg s = g.root;
  mrg_skip(tab, s, 0, (src + g.total) / nverts, (tgt + g.total) / nverts);
  return s;
}

__device__ __forceinline__ uint64_t pack(uint64_t s, uint64_t t) { return (s << 32) | t; }

// A node of one edge: categorical steps down to a single vertex pair.
__device__ void walk_one(const GenParams &g, const uint32_t *__restrict__ tab, uint64_t first, uint64_t src,
                         uint64_t tgt, uint64_t nverts, uint64_t *out) {
  Mrg st = node_state(g, tab, src, tgt, nverts);
  uint64_t n = nverts, s = src, t = tgt;
  while (n > 1) {
    const int q = categorical4(g, st);
    int so = q / 2, to = q % 2;
    if (s == t && so > to) {
      const int tmp = so;
      so = to;
      to = tmp;
    }
    n /= 2;
    s += n * so;
    t += n * to;
  }
  out[first] = pack(s, t);
}

// One level of the recursion: nodes with at least two edges and two
// vertices.  Kids with one vertex emit their edges, kids with one edge walk
// down, the others go to the next level.
__global__ void k_gen_level(GenParams g, const uint32_t *__restrict__ tab, const GenNode *__restrict__ cur,
                            uint64_t ncur, uint64_t nverts, GenNode *next, unsigned long long *nnext, uint64_t cap,
                            uint64_t *out, unsigned int *overflow) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ncur;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const GenNode nd = cur[i];
    Mrg st = node_state(g, tab, nd.src, nd.tgt, nverts);
    int64_t cnt[4];
    if (nd.count <= 20) {
      cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
      for (uint64_t e = 0; e < nd.count; ++e) cnt[categorical4(g, st)]++;
    } else {
      int64_t left = (int64_t)nd.count;
      double rest = 1.;
      for (int q = 0; q < 3; ++q) {
        cnt[q] = binomial(left, g.ini[q] / rest, st);
        left -= cnt[q];
        rest -= g.ini[q];
      }
      cnt[3] = left;
    }
    if (nd.src == nd.tgt) {
      cnt[1] += cnt[2];
      cnt[2] = 0;
    }
    // counts outside [0, count] would write past the node's edges: stop
    bool bad = false;
    for (int q = 0; q < 4; ++q) bad |= cnt[q] < 0 || (uint64_t)cnt[q] > nd.count;
    if (bad) {
      atomicOr(overflow, 2u);
      continue;
    }
    const uint64_t half = nverts / 2;
    uint64_t fei = nd.first;
    for (int q = 0; q < 4; ++q) {
      if (!cnt[q]) continue;
      const uint64_t ks = nd.src + half * (uint64_t)(q / 2), kt = nd.tgt + half * (uint64_t)(q % 2);
      const uint64_t kc = (uint64_t)cnt[q];
      if (half == 1) {
        for (uint64_t e = 0; e < kc; ++e) out[fei + e] = pack(ks, kt);
      } else if (kc == 1) {
        walk_one(g, tab, fei, ks, kt, half, out);
      } else {
        const unsigned long long slot = atomicAdd(nnext, 1ull);
        if (slot < cap)
          next[slot] = GenNode{fei, kc, (uint32_t)ks, (uint32_t)kt};
        else
          atomicOr(overflow, 1u);
      }
      fei += kc;
    }
  }
}

// A graph of one edge: the root walks down.
__global__ void k_gen_one(GenParams g, const uint32_t *__restrict__ tab, uint64_t *out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) walk_one(g, tab, 0, 0, 0, g.total, out);
}

// ---- device: the vertex permutation ------------------------------------------

__global__ void k_perm_keys(GenParams g, const uint32_t *__restrict__ tab, uint64_t n, uint64_t nkeys,
                            uint64_t *key, uint32_t *id) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    Mrg s = g.root;
    mrg_skip(tab, s, 1, i, 0);
    key[i] = mrg_up_to(s, nkeys);
    id[i] = (uint32_t)i;
  }
}

// Fisher-Yates shuffle of every bucket (equal keys) with more than one id,
// one thread per bucket, from stream (1, key, 100).
__global__ void k_perm_shuffle(GenParams g, const uint32_t *__restrict__ tab, uint64_t n,
                               const uint64_t *__restrict__ key, uint32_t *id) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[p];
    if (p > 0 && key[p - 1] == k) continue;
    uint64_t e = p + 1;
    while (e < n && key[e] == k) ++e;
    const uint64_t cnt = e - p;
    if (cnt < 2) continue;
    Mrg s = g.root;
    mrg_skip(tab, s, 1, k, 100);
    uint32_t *a = id + p;
    for (uint64_t i = cnt - 1; i > 0; --i) {
      const uint64_t j = mrg_up_to(s, i + 1);
      if (i != j) {
        const uint32_t t = a[i];
        a[i] = a[j];
        a[j] = t;
      }
    }
  }
}

// Permuted endpoints, (min, max) orientation, packed as min << 32 | max.
__global__ void k_perm_edges(const uint32_t *__restrict__ perm, uint64_t ned, uint64_t *edges) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < ned;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = edges[e];
    const uint64_t a = perm[x >> 32], b = perm[x & 0xFFFFFFFFull];
    edges[e] = a < b ? pack(a, b) : pack(b, a);
  }
}

int grid_of(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 16384));
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(uint64_t count) {
    if (hipMalloc((void **)&p, sizeof(T) * std::max<uint64_t>(count, 1)) != hipSuccess)
      return fail(AXBC_ENOMEM, "generator allocation of %s failed (line %d)", "device memory", __LINE__);
    return AXBC_OK;
  }
};

}  // namespace

namespace axbc {

// Generated edges (min << 32 | max) into edges[0, nedges) on the device.
int kronecker_device_edges(hipStream_t s, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                           const double initiator[4], uint64_t *edges) {
  if (scale < 1 || scale > 30) return fail(AXBC_EINVAL, "scale must be in [1, 30]%s (line %d)", "", __LINE__);
  GenParams g;
  g.root = seed_from_user(seed1, seed2);
  for (int i = 0; i < 4; ++i) g.ini[i] = initiator[i];
  g.total = 1ull << scale;
  const std::vector<uint32_t> htab = skip_table();
  DevBuf<uint32_t> tab;
  int rc;
  if ((rc = tab.alloc(htab.size()))) return rc;
  HIP_TRY(hipMemcpyAsync(tab.p, htab.data(), sizeof(uint32_t) * htab.size(), hipMemcpyHostToDevice, s));
  // every slot is written by the recursion; zeros keep the permutation's
  // reads in range even if a level stopped early
  if (nedges) HIP_TRY(hipMemsetAsync(edges, 0, sizeof(uint64_t) * nedges, s));
  // the recursion, level by level; a level holds nodes of >= 2 edges, so at
  // most nedges / 2 of them
  if (nedges == 1) hipLaunchKernelGGL(k_gen_one, dim3(1), dim3(64), 0, s, g, tab.p, edges);
  if (nedges >= 2) {
    const uint64_t cap = nedges / 2 + 1;
    DevBuf<GenNode> a, b;
    DevBuf<unsigned long long> cnt;
    DevBuf<unsigned int> ovf;
    if ((rc = a.alloc(cap)) || (rc = b.alloc(cap)) || (rc = cnt.alloc(1)) || (rc = ovf.alloc(1))) return rc;
    const GenNode root{0, nedges, 0, 0};
    HIP_TRY(hipMemcpyAsync(a.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(ovf.p, 0, sizeof(unsigned int), s));
    uint64_t ncur = 1;
    GenNode *cur = a.p, *nxt = b.p;
    for (uint64_t nverts = g.total; nverts >= 2 && ncur > 0; nverts /= 2) {
      HIP_TRY(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_gen_level, dim3(grid_of(ncur)), dim3(256), 0, s, g, tab.p, cur, ncur, nverts, nxt,
                         cnt.p, cap, edges, ovf.p);
      HIP_TRY(hipGetLastError());
      unsigned long long hn = 0;
      unsigned int hovf = 0;
      HIP_TRY(hipMemcpyAsync(&hn, cnt.p, sizeof(hn), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(&hovf, ovf.p, sizeof(hovf), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (hovf)
        return fail(AXBC_ESTATE, "generator recursion failed (code %s) (line %d)", hovf == 1 ? "1" : "2", __LINE__);
      ncur = hn;
      std::swap(cur, nxt);
    }
  }
  // the vertex permutation
  const uint64_t n = g.total, nkeys = 2 * n + 128;
  DevBuf<uint64_t> key, key2;
  DevBuf<uint32_t> id, id2;
  if ((rc = key.alloc(n)) || (rc = key2.alloc(n)) || (rc = id.alloc(n)) || (rc = id2.alloc(n))) return rc;
  hipLaunchKernelGGL(k_perm_keys, dim3(grid_of(n)), dim3(256), 0, s, g, tab.p, n, nkeys, key.p, id.p);
  HIP_TRY(hipGetLastError());
  int end_bit = 1;
  while (end_bit < 64 && (nkeys - 1) >> end_bit) ++end_bit;
  size_t tb = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, key2.p, id.p, id2.p, (int)n, 0, end_bit, s));
  DevBuf<unsigned char> tmp;
  if ((rc = tmp.alloc(tb))) return rc;
  // stable: equal keys keep ascending ids, as the counting sort does
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, key2.p, id.p, id2.p, (int)n, 0, end_bit, s));
  hipLaunchKernelGGL(k_perm_shuffle, dim3(grid_of(n)), dim3(256), 0, s, g, tab.p, n, key2.p, id2.p);
  if (nedges)
    hipLaunchKernelGGL(k_perm_edges, dim3(grid_of(nedges)), dim3(256), 0, s, id2.p, nedges, edges);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return AXBC_OK;
}

}  // namespace axbc

extern "C" {

int axbc_kronecker_generate(int device, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                            const double initiator[4], uint64_t *edges) {
  if (!initiator || (nedges && !edges)) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(device));
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf<uint64_t> d;
  int rc;
  if ((rc = d.alloc(nedges))) return rc;
  if ((rc = kronecker_device_edges(s, scale, nedges, seed1, seed2, initiator, d.p))) return rc;
  if (nedges) HIP_TRY(hipMemcpyAsync(edges, d.p, sizeof(uint64_t) * nedges, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return AXBC_OK;
}

}  // extern "C"
