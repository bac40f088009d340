// Batched rounds for high-diameter graphs: one workgroup per root.
//
// On graphs of thousands of levels whose roots' frontiers rarely meet (the
// 4096^2 lattice: ~8000 levels, a few thousand vertices per level and root)
// the bit-packed sweep shares almost no adjacency reads between roots and
// pays several grid-wide kernel boundaries per level.  Here each root of the
// batch owns one workgroup (one CU) for its whole round: the levels are
// separated by workgroup barriers only, so no level waits for another root
// or for a kernel launch, and up to 256 roots (one per CU) run side by side.
//
// Per root (lane l) the state is root-major, so that the neighbours of a
// vertex share cache lines:
//   st[l][v]  = level tag << 32 | sigma   (tag = tag0 + level this batch;
//               any smaller tag means unvisited, so nothing is cleared
//               between batches)
//   dl[l][v]  = delta (FP64), valid for visited v
//   lst[l][]  = the visited vertices in level order, off[l][L] = start of L
//               (graphs of at most 2^28 vertices: an entry of a thread's pair
//               of lists of at most 4 edges each gets, once expanded, the
//               slots of its list that lead to DAG children in its top 4
//               bits, so that the backward pass reads the state words of
//               those children only)
// Forward, level L -> L+1 (src/cuda_kernels.cu:588-658): every frontier
// vertex u pushes sigma(u) to each neighbour w; the first visitor claims w
// with a compare-and-swap of the whole word (tag and sigma at once) and
// appends it, later visitors of the same level add to sigma with a 32-bit
// atomic on the word's low half (uint32, wrapping: order-independent).
// Backward, depth d (src/cuda_kernels.cu:719-772, bc_func_mono's do-while
// at src/axbc2d.c:1613-1625): every vertex v of level d folds
//   delta(v) = reach(v) + sum_{w in adj(v), lvl(w) = d+1} ((delta(w)+1)/sigma(w)) * sigma(v)
// left to right in adjacency order, with no atomics: the same operations in
// the same order as the one-root rounds, hence the same bits.
// The deltas then go to the batch's root-minor rows (delta = reach where the
// root did not reach), where the stop test and the bc update of the
// bit-packed sweep take them unchanged.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_internal.h"
#include "wave.h"

using namespace axbc;

namespace {

constexpr int kWgBlock = 1024;  // threads per root (16 waves on one CU)
constexpr int kEdgeRun = 8;     // edge loads in flight per thread
constexpr int kIdBits = 28;     // vertex id bits of a list entry that carries a child mask
constexpr uint32_t kIdMask = (1u << kIdBits) - 1u;

// The state words change under atomics from every wave of the workgroup:
// they are read and changed with workgroup-scope atomics, and the barrier
// between levels (a workgroup-scope fence) orders one level's atomics before
// the next level's reads.  Nothing here synchronises beyond the workgroup.
__device__ __forceinline__ uint64_t ld_wg(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t ld_wg(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Both list offsets of a vertex in one 8-byte load (4-byte aligned).
__device__ __forceinline__ uint2 ld_pair(const uint32_t *__restrict__ p) {
  uint2 r;
  __builtin_memcpy(&r, p, sizeof(r));
  return r;
}

// The ids of a list [b, b + d) with d <= 4 into slots O .. O+3 of w, from at
// most two 16-byte aligned loads ([b, b + 4) lies in [q, q + 8) for
// q = b & ~3); slots past d get 0.
// ned_al = the list array's length when it starts 16-byte aligned, else 0:
// lists whose window would pass the end (or any list, unaligned) load entry
// by entry.
template <int O>
__device__ __forceinline__ void short_list(const uint32_t *__restrict__ row, uint64_t ned_al, uint32_t b, uint32_t d,
                                           uint32_t (&w)[kEdgeRun]) {
  const uint32_t q = b & ~3u;
  if ((uint64_t)q + 8 > ned_al) {
#pragma unroll
    for (int t = 0; t < 4; ++t) w[O + t] = (uint32_t)t < d ? row[b + t] : 0u;
    return;
  }
  // both windows lie inside the array here; loaded unconditionally, as a
  // conditional vector load is split into one load per slot
  uint4 lo = *(const uint4 *)(row + q);
  uint4 hi = *(const uint4 *)(row + q + 4);
  // without this the compiler turns the selects below back into one 4-byte
  // load per slot at the selected address
  asm volatile("" : "+v"(lo.x), "+v"(lo.y), "+v"(lo.z), "+v"(lo.w), "+v"(hi.x), "+v"(hi.y), "+v"(hi.z));
  uint32_t x0 = lo.x, x1 = lo.y, x2 = lo.z, x3 = lo.w, x4 = hi.x, x5 = hi.y, x6 = hi.z;
  const uint32_t sh = b - q;
  if (sh & 2u) {
    x0 = x2;
    x1 = x3;
    x2 = x4;
    x3 = x5;
    x4 = x6;
  }
  if (sh & 1u) {
    x0 = x1;
    x1 = x2;
    x2 = x3;
    x3 = x4;
  }
  w[O + 0] = d > 0 ? x0 : 0u;
  w[O + 1] = d > 1 ? x1 : 0u;
  w[O + 2] = d > 2 ? x2 : 0u;
  w[O + 3] = d > 3 ? x3 : 0u;
}

// Forward over one run of edges (slot t valid when bit t of `valid`, its
// frontier vertex's path count su[t]): every neighbour word, then every
// claim, then the appends and additions, so the run costs three memory
// round trips.  A vertex unvisited in this batch is claimed with tag and
// path count at once (on failure word[t] receives the winner's word); later
// visitors of the same level add to the word's low half (uint32, wrapping).
// Returns the run's DAG edges as a slot mask; *all gains the reach of the
// claimed vertices.
__device__ __forceinline__ uint32_t push_run(uint64_t *st, uint32_t *lst, uint32_t e, uint32_t *here,
                                             const uint32_t *__restrict__ reach, uint32_t tag0, uint32_t tnext,
                                             const uint32_t (&w)[kEdgeRun], uint32_t valid,
                                             const uint32_t (&su)[kEdgeRun], unsigned long long *all) {
  uint64_t word[kEdgeRun];
  bool won[kEdgeRun];
  uint32_t dag = 0;  // bit t: slot t is a DAG edge
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t) word[t] = (valid >> t) & 1u ? ld_wg(st + w[t]) : ~0ull;
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t)
    won[t] = (uint32_t)(word[t] >> 32) < tag0 &&
             __hip_atomic_compare_exchange_strong(st + w[t], &word[t], ((uint64_t)tnext << 32) | su[t],
                                                  __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t) {
    if (won[t]) {
      lst[e + atomicAdd(here, 1u)] = w[t];
      dag |= 1u << t;
      if (reach) *all += reach[w[t]];
    } else if ((uint32_t)(word[t] >> 32) == tnext) {
      __hip_atomic_fetch_add((uint32_t *)(st + w[t]), su[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      dag |= 1u << t;
    }
  }
  return dag;
}

// Backward over one run of edges: slots below `first1` belong to the first
// vertex (path count sv0, sum acc0), the others to the second; each sum
// takes its terms in slot order, i.e. in its own adjacency order.
__device__ __forceinline__ void pull_run(const uint64_t *st, const double *dl, uint32_t tnext,
                                         const uint32_t (&w)[kEdgeRun], uint32_t valid, int first1, double sv0,
                                         double sv1, double &acc0, double &acc1, uint32_t &terms) {
  uint64_t word[kEdgeRun];
  double x[kEdgeRun];
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t) word[t] = (valid >> t) & 1u ? st[w[t]] : 0ull;
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t) x[t] = (uint32_t)(word[t] >> 32) == tnext ? dl[w[t]] : 0.0;
#pragma unroll
  for (int t = 0; t < kEdgeRun; ++t)
    if ((uint32_t)(word[t] >> 32) == tnext) {
      const double xt = (x[t] + 1.0) / (double)(uint32_t)word[t];
      if (t < first1)
        acc0 += xt * sv0;
      else
        acc1 += xt * sv1;
      ++terms;
    }
}

__device__ __forceinline__ unsigned long long block_sum(unsigned long long x, unsigned long long *slot) {
  x = wave_sum_u64(x);
  if (lane_id() == 0 && x) atomicAdd(slot, x);
  return x;
}

// One workgroup per lane l < nb; reach null means every reach is 0 (the
// common case without the 1-degree reduction) and saves those loads.
// cnt[4 l + (0 visited, 1 reach sum over levels > 0, 2 edges of visited
// vertices, 3 DAG edges)], nlev[l] = deepest non-empty level.  A root with
// more than lcap - 2 levels sets *ovf and leaves its round unfinished (the
// caller grows lcap and runs the batch again).
__global__ __launch_bounds__(kWgBlock) void k_wgsweep(uint32_t n, const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ row,
                                                      const uint32_t *__restrict__ reach,
                                                      const uint32_t *__restrict__ src, uint32_t tag0, uint64_t *pst,
                                                      double *pdl, uint32_t *plst, uint32_t *ploff, uint32_t lcap,
                                                      unsigned long long *cnt, int32_t *nlev, uint32_t *ovf,
                                                      unsigned long long *kb, uint64_t ned_al) {
  const uint32_t l = blockIdx.x;
  const uint32_t tid = threadIdx.x;
  uint64_t *st = pst + (size_t)l * n;
  double *dl = pdl + (size_t)l * n;
  uint32_t *lst = plst + (size_t)l * n;
  uint32_t *off = ploff + (size_t)l * lcap;
  // list entries carry child masks where the ids leave 4 bits free
  const bool packm = (uint64_t)n <= (uint64_t)kIdMask + 1u;
  const uint32_t v0 = src[l];
  __shared__ uint32_t s_cnt[3];
  __shared__ unsigned long long s_acc[4];
  if (tid == 0) {
    __hip_atomic_store(st + v0, ((uint64_t)tag0 << 32) | 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    dl[v0] = reach ? (double)reach[v0] : 0.0;
    lst[0] = v0;
    off[0] = 0;
    off[1] = 1;
    s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
    s_acc[0] = s_acc[1] = s_acc[2] = s_acc[3] = 0;
  }
  __syncthreads();
  unsigned long long er = 0, dag = 0, all = 0;
  uint32_t b = 0, e = 1;
  int L = 0;
  bool over = false;
  // Forward.  Level L's counter is s_cnt[L % 3]; thread 0 clears the next
  // level's before this level's barrier (its last reader passed the barrier
  // of level L-1), so one barrier per level suffices.
  for (;;) {
    const uint32_t tnext = tag0 + (uint32_t)L + 1u;
    uint32_t *here = &s_cnt[L % 3];
    if (tid == 0) s_cnt[(L + 1) % 3] = 0;
    // two frontier vertices per thread; lists of at most 4 edges (all of a
    // lattice) in two 4-slot halves of one run, longer ones as one list
    // walked kEdgeRun at a time
    for (uint32_t i0 = b + tid; i0 < e; i0 += 2 * kWgBlock) {
      const uint32_t i1 = i0 + kWgBlock;
      const bool h1 = i1 < e;
      const uint32_t u0 = lst[i0], u1 = h1 ? lst[i1] : 0u;
      const uint32_t s0 = (uint32_t)ld_wg(st + u0), s1 = h1 ? (uint32_t)ld_wg(st + u1) : 0u;
      const uint2 c0 = ld_pair(col + u0), c1 = h1 ? ld_pair(col + u1) : make_uint2(0u, 0u);
      const uint32_t b0 = c0.x, d0 = c0.y - c0.x, b1 = c1.x, d1 = c1.y - c1.x;
      er += (unsigned long long)d0 + d1;
      uint32_t w[kEdgeRun], su[kEdgeRun];
      if (d0 <= 4 && d1 <= 4) {
        short_list<0>(row, ned_al, b0, d0, w);
        short_list<4>(row, ned_al, b1, d1, w);
#pragma unroll
        for (int t = 0; t < kEdgeRun; ++t) su[t] = t < 4 ? s0 : s1;
        const uint32_t bits = push_run(st, lst, e, here, reach, tag0, tnext, w,
                                       ((1u << d0) - 1u) | (((1u << d1) - 1u) << 4), su, &all);
        dag += (uint32_t)__popc(bits);
        if (packm) {
          // this thread's entries (read above; no other thread reads them
          // until the backward pass, which pairs the same entries)
          lst[i0] = u0 | ((bits & 15u) << kIdBits);
          if (h1) lst[i1] = u1 | ((bits >> 4) << kIdBits);
        }
        continue;
      }
      for (uint32_t p0 = 0; p0 < d0 + d1; p0 += kEdgeRun) {
        uint32_t valid = 0;
#pragma unroll
        for (int t = 0; t < kEdgeRun; ++t) {
          const uint32_t q = p0 + t;
          w[t] = q < d0 ? row[b0 + q] : (q < d0 + d1 ? row[b1 + (q - d0)] : 0u);
          su[t] = q < d0 ? s0 : s1;
          valid |= (q < d0 + d1 ? 1u : 0u) << t;
        }
        dag += (uint32_t)__popc(push_run(st, lst, e, here, reach, tag0, tnext, w, valid, su, &all));
      }
    }
    __syncthreads();
    const uint32_t c = *here;
    if (c == 0) break;
    b = e;
    e += c;
    ++L;
    if ((uint32_t)L + 1u >= lcap) {
      over = true;
      break;
    }
    if (tid == 0) off[L + 1] = e;
  }
  block_sum(er, &s_acc[2]);
  block_sum(dag, &s_acc[3]);
  block_sum(all, &s_acc[1]);
  __syncthreads();
  if (tid == 0) {
    cnt[4 * l + 0] = e;
    cnt[4 * l + 1] = s_acc[1];
    cnt[4 * l + 2] = s_acc[2];
    cnt[4 * l + 3] = s_acc[3];
    nlev[l] = L;
    if (over) atomicOr(ovf, 1u);
  }
  if (over) return;
  // Backward.  The deepest level keeps delta = reach; then depths L-1 .. 1,
  // or depth 0 alone when the deepest level is 1 (the do-while); the root's
  // delta stays reach otherwise.
  const uint32_t vlast = e;
  unsigned long long bytes = 0;
  if (L >= 1)
    for (uint32_t i = b + tid; i < e; i += kWgBlock) {
      const uint32_t v = packm ? lst[i] & kIdMask : lst[i];
      dl[v] = reach ? (double)reach[v] : 0.0;
    }
  __syncthreads();
  const int dmin = L == 1 ? 0 : 1;
  for (int d = L - 1; d >= dmin; --d) {
    const uint32_t db = off[d], de = off[d + 1];
    const uint32_t tnext = tag0 + (uint32_t)d + 1u;
    // two vertices per thread, their edges walked as in the forward sweep;
    // each vertex folds its terms in its own adjacency order
    for (uint32_t i0 = db + tid; i0 < de; i0 += 2 * kWgBlock) {
      const uint32_t i1 = i0 + kWgBlock;
      const bool h1 = i1 < de;
      const uint32_t e0 = lst[i0], e1 = h1 ? lst[i1] : 0u;
      const uint32_t v0 = packm ? e0 & kIdMask : e0, v1 = packm ? e1 & kIdMask : e1;
      const double sv0 = (double)(uint32_t)st[v0], sv1 = h1 ? (double)(uint32_t)st[v1] : 1.0;
      double acc0 = reach ? (double)reach[v0] : 0.0, acc1 = reach && h1 ? (double)reach[v1] : 0.0;
      const uint2 c0 = ld_pair(col + v0), c1 = h1 ? ld_pair(col + v1) : make_uint2(0u, 0u);
      const uint32_t b0 = c0.x, d0 = c0.y - c0.x, b1 = c1.x, d1 = c1.y - c1.x;
      uint32_t terms = 0;
      uint32_t w[kEdgeRun];
      if (d0 <= 4 && d1 <= 4) {
        short_list<0>(row, ned_al, b0, d0, w);
        short_list<4>(row, ned_al, b1, d1, w);
        // only the slots that lead to DAG children (the others cannot hold a
        // vertex of level d+1)
        const uint32_t k0 = packm ? e0 >> kIdBits : 15u, k1 = packm ? e1 >> kIdBits : 15u;
        pull_run(st, dl, tnext, w, (k0 & ((1u << d0) - 1u)) | ((k1 & ((1u << d1) - 1u)) << 4), 4, sv0, sv1, acc0, acc1,
                 terms);
      } else {
        for (uint32_t p0 = 0; p0 < d0 + d1; p0 += kEdgeRun) {
          uint32_t valid = 0;
#pragma unroll
          for (int t = 0; t < kEdgeRun; ++t) {
            const uint32_t q = p0 + t;
            w[t] = q < d0 ? row[b0 + q] : (q < d0 + d1 ? row[b1 + (q - d0)] : 0u);
            valid |= (q < d0 + d1 ? 1u : 0u) << t;
          }
          // slots of the first vertex: those below d0 - p0
          const int first1 = p0 >= d0 ? 0 : (int)min(d0 - p0, (uint32_t)kEdgeRun);
          pull_run(st, dl, tnext, w, valid, first1, sv0, sv1, acc0, acc1, terms);
        }
      }
      dl[v0] = acc0;
      if (h1) dl[v1] = acc1;
      // per vertex: list 4 + word 8 + offsets 8 + reach 4 + delta 8; per
      // edge: id 4 + word 8; per term: delta 8
      bytes += 32ull * (1u + (h1 ? 1u : 0u)) + 12ull * (d0 + d1) + 8ull * terms;
    }
    __syncthreads();
  }
  if (kb) {
    // forward: per visited vertex list 4 + word 8 + offsets 8 (+ reach 4 and
    // the list write 4 when claimed); per edge id 4 + word 8; per DAG edge
    // the atomic 8
    if (tid == 0) bytes += 28ull * vlast + 12ull * s_acc[2] + 8ull * s_acc[3];
    bytes = wave_sum_u64(bytes);
    if (lane_id() == 0 && bytes) atomicAdd(kb + (size_t)((blockIdx.x * (kWgBlock / kWave) + tid / kWave) & (kKbSlots - 1)) * kKbStride, bytes);
  }
}

// The batch's root-minor delta rows from the per-root arrays: delta for the
// visited (vertex, root) pairs of lanes < nb, reach otherwise; seen[v] = stamp
// where some root reached v.  A tile of 32 vertices goes through LDS, so
// both the per-root reads and the row writes are coalesced.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_wgout(uint32_t n, int nb, uint32_t tag0, const uint64_t *__restrict__ pst,
                                                  const double *__restrict__ pdl, const uint32_t *__restrict__ reach,
                                                  double *dlt, uint32_t *seen, uint32_t stamp) {
  // pst / pdl start at the first sweep lane of these rows
  constexpr int B = 64 * NW;
  constexpr int TV = 32;
  __shared__ double tile[TV][B + 1];
  __shared__ double s_r[TV];
  __shared__ uint32_t s_vis[TV];
  for (uint32_t vb = blockIdx.x * TV; vb < n; vb += gridDim.x * TV) {
    if (threadIdx.x < TV) {
      s_vis[threadIdx.x] = 0;
      s_r[threadIdx.x] = vb + threadIdx.x < n ? (double)reach[vb + threadIdx.x] : 0.0;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < (uint32_t)(TV * nb); idx += blockDim.x) {
      const uint32_t l = idx / TV, vi = idx % TV, v = vb + vi;
      if (v < n) {
        const uint64_t word = pst[(size_t)l * n + v];
        const bool vis = (uint32_t)(word >> 32) >= tag0;
        tile[vi][l] = vis ? pdl[(size_t)l * n + v] : s_r[vi];
        if (vis) s_vis[vi] = 1;
      }
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < (uint32_t)(TV * B); idx += blockDim.x) {
      const uint32_t vi = idx / B, l = idx % B, v = vb + vi;
      // lanes past nb hold reach too, so that every row of a vertex no root
      // reached equals reach (the bit-packed sweep relies on it)
      if (v < n) dlt[(size_t)v * B + l] = (int)l < nb ? tile[vi][l] : s_r[vi];
    }
    if (threadIdx.x < TV && vb + threadIdx.x < n && s_vis[threadIdx.x]) seen[vb + threadIdx.x] = stamp;
    __syncthreads();
  }
}

// One root's delta from its state words and deltas: reach where it did not
// reach.
__global__ void k_wgdelta(uint32_t n, uint32_t tag0, const uint64_t *__restrict__ st, const double *__restrict__ dl,
                          const uint32_t *__restrict__ reach, double *out) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    out[v] = (uint32_t)(st[v] >> 32) >= tag0 ? dl[v] : (double)reach[v];
}

// One root's levels (-1 unreached) and path counts from its state words.
__global__ void k_wglevels(uint32_t n, uint32_t tag0, const uint64_t *__restrict__ st, int32_t *lvl,
                           uint32_t *sigma) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint64_t word = st[v];
    const uint32_t tg = (uint32_t)(word >> 32);
    const bool vis = tg >= tag0;
    lvl[v] = vis ? (int32_t)(tg - tag0) : -1;
    sigma[v] = vis ? (uint32_t)word : 0u;
  }
}

}  // namespace

namespace axbc {

hipError_t wg_sweep(hipStream_t s, int nb, uint32_t n, const uint32_t *col, const uint32_t *row,
                    const uint32_t *reach, const uint32_t *src, uint32_t tag0, uint64_t *pst, double *pdl,
                    uint32_t *plst, uint32_t *ploff, uint32_t lcap, unsigned long long *cnt, int32_t *nlev,
                    uint32_t *ovf, unsigned long long *kb, uint64_t ned) {
  // short lists load 16-byte windows of row only when it starts 16-byte aligned
  const uint64_t ned_al = ((uintptr_t)row & 15u) ? 0ull : ned;
  hipLaunchKernelGGL(k_wgsweep, dim3(nb), dim3(kWgBlock), 0, s, n, col, row, reach, src, tag0, pst, pdl, plst, ploff,
                     lcap, cnt, nlev, ovf, kb, ned_al);
  return hipGetLastError();
}

hipError_t wg_out(hipStream_t s, int nw, uint32_t n, int nb, int lane0, uint32_t tag0, const uint64_t *pst,
                  const double *pdl, const uint32_t *reach, double *dlt, uint32_t *seen, uint32_t stamp) {
  const int grid = grid_for((n + 31) / 32, 4096);
  pst += (size_t)lane0 * n;
  pdl += (size_t)lane0 * n;
  if (nw == 1)
    hipLaunchKernelGGL(k_wgout<1>, dim3(grid), dim3(kBlock), 0, s, n, nb, tag0, pst, pdl, reach, dlt, seen, stamp);
  else
    hipLaunchKernelGGL(k_wgout<2>, dim3(grid), dim3(kBlock), 0, s, n, nb, tag0, pst, pdl, reach, dlt, seen, stamp);
  return hipGetLastError();
}

hipError_t wg_delta(hipStream_t s, uint32_t n, uint32_t tag0, const uint64_t *st, const double *dl,
                    const uint32_t *reach, double *out) {
  hipLaunchKernelGGL(k_wgdelta, dim3(grid_for(n)), dim3(kBlock), 0, s, n, tag0, st, dl, reach, out);
  return hipGetLastError();
}

hipError_t wg_levels(hipStream_t s, uint32_t n, uint32_t tag0, const uint64_t *st, int32_t *lvl, uint32_t *sigma) {
  hipLaunchKernelGGL(k_wglevels, dim3(grid_for(n)), dim3(kBlock), 0, s, n, tag0, st, lvl, sigma);
  return hipGetLastError();
}

}  // namespace axbc
