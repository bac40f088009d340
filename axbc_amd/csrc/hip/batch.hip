// Bit-packed multi-source sampled-Brandes rounds (B = 64 or 128 roots per
// batch) for MI355X.
//
// One lane of a wave owns one root (two for B = 128), so every per-root row
// of a vertex (sigma u32, delta f64, x = (delta+1)/sigma f64) is read and
// written as one coalesced row, and the adjacency of a vertex is read once
// for the whole batch.  Per vertex the batch keeps bit words (NW = B/64
// uint64 each): VIS (roots that reached it so far), NX (roots reaching it
// at the current level) and a double-buffered dense frontier FR.  The
// frontier of level L is a list of (vertex, word) entries; all levels stay
// in one list (q, qw, qoff) for the backward sweep.
//
// Per level L:
//   push    frontier entries (u, f) of L-1 mark every w in adj(u) with
//           f & ~VIS[w] (atomicOr on NX; the first marker appends w)
//   sigma   every new w pulls sigma_s[w] = sum of sigma_s[u] over u in
//           radj(w) with s in FR[u] & NX[w]   (uint32, wrapping)
//   close   FR/VIS/list words for level L, per-root counters, x rows
// Per depth d (deepest first): VIS loses the level d+1 bits, then every
// entry (v, f) of level d pulls, per root s in f,
//   delta_s[v] = reach[v] + sum_{w in adj(v), s not in VIS[w]} x_s[w] * sigma_s[v]
// in adjacency order (the same left fold as the sequential oracle), no
// atomics.  Vertices with more edges than one chunk are split into chunks
// processed by separate waves and combined in chunk order.
// The bc update applies the batch's roots to every vertex in root order
// (bc = ((bc + t_0) + t_1) + ...), i.e. exactly the sequence of one-root
// rounds, and stops at the first root after which the adaptive stop test
// (src/axbc2d.c:2846-2870) holds, evaluated exactly per root.
//
// Reference semantics restated: src/cuda_kernels.cu:588-658, 719-772,
// 269-302, 339-373; src/axbc2d.c:1514-1659 (per root) and 2834-3156 (loop).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <tuple>
#include <utility>
#include <vector>

#include "../../../include/axbc.h"
#include "engine_internal.h"
#include "wave.h"

using namespace axbc;

namespace axbc {

constexpr uint32_t kChunkMin = 64;  // smallest edges-per-work-item for split vertices
#ifndef AXBC_BURST
#define AXBC_BURST 3
#endif
constexpr int kBurst = AXBC_BURST;  // row loads in flight per wave in the pull loops
// Waves per SIMD asked of the compiler for the frontier kernels (a 64-VGPR
// budget): these loops are bound by memory latency, and 8 resident waves
// with 2 row loads each beat 4 waves with 8 (measured on SCALE 22, 128
// roots); k_bclose keeps the compiler's choice.
constexpr int kWavesPerSimd = 8;
constexpr int kBGrid = 2048;      // workgroups of the batch level kernels
constexpr int kGraphSteps = 8;    // levels (depths) per graph
// Per-wave counter slots.  Counters that many waves would add to the same
// few addresses (per-root visits, reach sums, edges, DAG edges; the level's
// edge sums for the direction test) are kept per wave instead: a wave owns
// slot gw (its index in the grid) and adds into it with plain loads and
// stores, and one reduction sums the slots (integers, so the order does
// not matter).  Device-scope atomics on one address from thousands of waves
// serialise, which made every level cost tens of microseconds whatever its
// size.
constexpr int kSlotWaves = kBGrid * kBlock / kWave;  // waves of the largest level grid
constexpr int kAppendBuf = 256;   // list ids buffered per wave before one tail atomic
constexpr int kLevelSlot = 8;     // 64-bit fields per wave in the level slots of k_bclose
// Sparse levels: at most kSparseRatio roots per list entry on average (high-
// diameter graphs, where the roots' frontiers rarely meet).  Their entries
// of at most kSparseDeg edges run one (entry, root) pair per lane instead of
// one root per lane, which would leave most lanes idle.
constexpr int kSparseRatio = 4;
constexpr int kSparseAvgDeg = 8;   // and at most this many edges per entry on average
constexpr uint32_t kSparseDeg = 32;
// A batch runs its levels in persistent cooperative kernels (when enabled
// and no vertex is split) once the previous batch had more levels than this.
constexpr int kPersistLevels = 64;
// Workgroups per CU of a persistent launch: every workgroup's agent-scope
// fences at each grid barrier write back its XCD's L2, so fewer workgroups
// make cheaper barriers.
constexpr int kPersistPerCu = 1;

constexpr int kHandSlices = 8;   // item counters of a split kernel: one per XCD
constexpr int kHandStride = 16;  // 128 bytes between counters

struct BCtl {
  uint32_t tail;       // list length
  uint32_t done;
  uint32_t empty_lev;
  uint32_t overflow;
  uint32_t bu;         // the current level runs bottom-up (decided by the previous close)
  uint32_t wg_done;
  uint32_t stop_round; // first round after which the stop test holds (nb if none)
  uint32_t pad;
  unsigned long long split_claim;  // (entry index << 40) | chunk base
  unsigned long long cand_edges;   // in-edges of vertices not yet reached by every root
  // chunk-item hand-out of k_bdep_split (dep, cleared by k_bstrip) and of
  // k_bpull_split (pull, cleared at each level's end): counter c (one per
  // XCD, each on its own 128-byte line) hands out slice c of the items
  unsigned long long hand_dep[kHandSlices * kHandStride];
  unsigned long long hand_pull[kHandSlices * kHandStride];
  // 64-entry groups of k_bdep's dense levels (cleared by k_bstrip) and
  // 64-vertex groups of the bottom-up pull (cleared at each level's end)
  unsigned long long hand_ent[kHandSlices * kHandStride];
  unsigned long long hand_vert[kHandSlices * kHandStride];
  // device-resident loop state, so that every level (depth) runs the same
  // launch sequence and a prebuilt graph can replay it
  int level;                       // forward level the next kernels run
  uint32_t mark_base;              // level stamps of `mark`: mark_base + level
  uint32_t stamp;                  // batch stamp of `seen`
  int depth;                       // backward depth the dependency kernels run (set by k_bstrip)
  int dmin;                        // last backward depth (0 only for single-level roots)
  int sdepth;                      // backward depth the next k_bstrip runs (set by k_bdep)
  uint32_t pad3;
  unsigned long long all[2];       // the batch's roots per word
  unsigned long long rootmask[2];  // roots whose deepest level is 1 (depth 0)
};

struct BatchState {
  int nw = 1;        // 64-bit words per vertex
  int B = 64;        // roots per batch (64 * nw)
  uint32_t n = 0;
  uint64_t *vis = nullptr, *nx = nullptr, *fr[2] = {nullptr, nullptr};
  uint32_t *sig = nullptr;
  double *dlt = nullptr, *xr = nullptr;
  uint32_t *mark = nullptr;  // B = 128: level stamp of the appender
  uint32_t *seen = nullptr;  // batch stamp of visited vertices
  uint32_t *q = nullptr;
  uint64_t *qw = nullptr;
  uint32_t qcap = 0;
  uint32_t *qoff = nullptr;
  uint64_t *act = nullptr;   // per level: OR of the level's words
  // split vertices: entry list index, vertex, chunk base per split entry
  uint32_t *sp_e = nullptr;
  uint64_t *sp_base = nullptr;
  uint32_t sp_cap = 0;
  double *part = nullptr;
  uint64_t part_rows = 0;
  unsigned long long *cnt = nullptr;  // per root: nvisited, all, e_r, e_dag
  unsigned long long *wslot = nullptr;  // 4 x kSlotWaves x B per-wave partials of cnt
  unsigned long long *lslot = nullptr;  // per-wave level sums of k_bclose (kLevelSlot each)
  uint8_t *lmode = nullptr;             // per level: 1 when sparse (few roots per entry)
  uint32_t *src = nullptr;
  double *prob = nullptr;
  uint32_t *r0 = nullptr;            // reach of each root
  // src/prob/r0/cnt and the root rows are in lane order; lane_of[k] is the
  // lane of the batch's k-th root (device copy for the stop test and update)
  uint32_t *lane_of = nullptr;
  std::vector<uint32_t> h_lane_of;
  uint32_t *check = nullptr;
  BCtl *ctl = nullptr, *h_ctl = nullptr;
  unsigned long long *h_cnt = nullptr;
  uint64_t *h_act = nullptr;
  uint32_t stamp = 0;
  uint32_t epoch = 0;  // level stamps for `mark`, unique across batches
  bool have = false;
  bool had_prev = false;  // a sweep ran in this state before the current one
  int lev_guess = 12;
  uint32_t chunk = 256;  // edges per work item of split vertices (tunable)
  // sigma rows of the vertices with more than this many pull edges are zero
  // at a batch start (all rows are after the fill)
  uint32_t sig_zero_above = 0;
  uint32_t pull16 = 32;  // bottom-up when candidate in-edges <= pull16/16 x push edges (0: never)
  hipGraphExec_t gfwd = nullptr, gbwd = nullptr;  // prebuilt level / depth steps
  uint64_t gkey = 0;
  size_t bytes = 0;
  int levels = 0;            // levels of the last batch's list (qoff[0 .. levels])
  int coop_fwd = 0, coop_bwd = 0;  // co-resident grids of the persistent kernels (0: not computed)
  uint32_t *prog = nullptr;          // per workgroup of a persistent launch: barriers reached
  uint32_t *bar = nullptr;           // grid barrier words of the persistent launches
  bool plan_only = false;            // balloc adds up sizes without allocating
  bool charged = false;              // bytes counted in the context's total
  uint32_t h_bar[4] = {0, 0, 0, 0};
  // scratch of the getters and the utility, allocated on first use and kept
  void *row_out = nullptr;   // N x 8 bytes: one root's delta, or its lvl + sigma
  uint32_t *util_v = nullptr;
  double *util_w = nullptr, *util_part = nullptr, *util_out = nullptr;
  uint32_t util_cap = 0;
  // workgroup sweep (one workgroup per root, wgsweep.hip), allocated on
  // first use: per lane the state words, deltas, level list and level
  // offsets (wlcap per lane), the deepest level per lane and the overflow flag
  uint64_t *wst = nullptr;
  double *wdl = nullptr;
  uint32_t *wlst = nullptr, *wloff = nullptr, *wovf = nullptr;
  int32_t *wnlev = nullptr;
  int32_t *h_wnlev = nullptr;  // pinned: wlanes deepest levels + the overflow flag
  uint32_t *wsrc = nullptr;    // roots of the sweep by sweep lane (batches above B roots)
  unsigned long long *wcnt = nullptr;  // their counters (4 per sweep lane)
  uint32_t wlcap = 0;
  int wlanes = 0;          // sweep lanes: B, or 2B at width 2B
  std::vector<uint32_t> h_wlane;  // sweep lane of the last batch's k-th root (when it was a workgroup sweep)
  // a batch of more than B roots: its halves' lane maps (draw index within
  // the half -> lane), the last half's also in h_lane_of
  bool wg_pair = false;
  int pair_nb[2] = {0, 0};
  std::vector<uint32_t> h_lane_half[2];
  uint32_t wtag = 1;       // tag of level 0 for the next workgroup sweep
  uint32_t wtag_last = 0;  // tag of level 0 of the last batch (when it was a workgroup sweep)
  bool last_wg = false;    // the last batch ran as a workgroup sweep
};

}  // namespace axbc

namespace {

// Flags set by an earlier kernel of the stream (the done flag of the level
// loop): that kernel's completion already orders them, so a relaxed load
// suffices.  An acquire here would invalidate the vector cache in every
// wave, which costs tens of microseconds per launch at these grid sizes.
__device__ __forceinline__ uint32_t ld_flag(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool bit(uint64_t w, int i) { return (w >> i) & 1ull; }

// A launch's algorithmic bytes, added by lane 0 of every wave to one of the
// kind's kKbSlots counters (chosen by wave id; one line each).
__device__ __forceinline__ void kb_add(unsigned long long *kb, unsigned long long bytes) {
  if (!kb || !bytes || lane_id() != 0) return;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  atomicAdd(kb + (size_t)(gw & (kKbSlots - 1)) * kKbStride, bytes);
}

template <int NW>
struct Words {
  uint64_t w[NW];
  __device__ __forceinline__ bool any() const {
    uint64_t a = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) a |= w[j];
    return a != 0;
  }
};

template <int NW>
__device__ __forceinline__ Words<NW> ldw(const uint64_t *p, uint32_t v) {
  Words<NW> r;
#pragma unroll
  for (int j = 0; j < NW; ++j) r.w[j] = p[(size_t)v * NW + j];
  return r;
}

template <int NW>
__device__ __forceinline__ Words<NW> shflw(const Words<NW> &x, int src) {
  Words<NW> r;
#pragma unroll
  for (int j = 0; j < NW; ++j) r.w[j] = __shfl(x.w[j], src);
  return r;
}

// Value of x in lane i for a wave-uniform i (from a ballot, a loop counter or
// a constant): one scalar read of that lane (the result lives in scalar
// registers) instead of a cross-lane permute through the LDS unit.
__device__ __forceinline__ uint32_t rdl(uint32_t x, int i) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, i);
}
__device__ __forceinline__ uint64_t rdl(uint64_t x, int i) {
  return ((uint64_t)rdl((uint32_t)(x >> 32), i) << 32) | rdl((uint32_t)x, i);
}
__device__ __forceinline__ double rdl(double x, int i) {
  return __builtin_bit_cast(double, rdl(__builtin_bit_cast(uint64_t, x), i));
}

template <int NW>
__device__ __forceinline__ Words<NW> rdlw(const Words<NW> &x, int i) {
  Words<NW> r;
#pragma unroll
  for (int j = 0; j < NW; ++j) r.w[j] = rdl(x.w[j], i);
  return r;
}

// Adds this wave's per-root partials v[j] (lane = root) to its own slot of
// counter kind k (0 visits, 1 reach sum, 2 edges, 3 DAG edges).  Only the
// wave with index gw touches slot gw, and kernels of one stream run one
// after another, so plain loads and stores suffice.
template <int NW>
__device__ __forceinline__ void slot_add(unsigned long long *wslot, int k, uint32_t gw, const uint32_t v[NW]) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  bool any = false;
#pragma unroll
  for (int j = 0; j < NW; ++j) any |= v[j] != 0;
  if (!__ballot(any)) return;
  unsigned long long *row = wslot + ((size_t)k * kSlotWaves + gw) * B;
#pragma unroll
  for (int j = 0; j < NW; ++j)
    if (v[j]) row[lane + 64 * j] += v[j];
}

// Adds val to root s of this wave's slot of kind k from any lane: lanes of
// one wave may hit the same root, so the add is atomic, at workgroup scope
// (no other wave touches the slot).
__device__ __forceinline__ void slot_add_root(unsigned long long *wslot, int k, uint32_t gw, int B, int s,
                                              unsigned long long val) {
  if (val)
    __hip_atomic_fetch_add(wslot + ((size_t)k * kSlotWaves + gw) * B + s, val, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Index of the k-th set bit (from 0) of x, which has more than k set bits.
__device__ __forceinline__ int select_bit(uint64_t x, uint32_t k) {
  int pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const uint32_t c = (uint32_t)__popcll(x & ((1ull << width) - 1ull));
    if (k >= c) {
      k -= c;
      x >>= width;
      pos += width;
    }
  }
  return pos;
}

// Root (lane index in the batch) of the k-th set bit of a word set.
template <int NW>
__device__ __forceinline__ int select_root(const uint64_t w[NW], uint32_t k) {
  const uint32_t c0 = (uint32_t)__popcll(w[0]);
  if (NW == 1 || k < c0) return select_bit(w[0], k);
  return 64 + select_bit(w[NW - 1], k - c0);
}

// A wave's buffer of ids to append to the level list, in LDS: ids are
// collected here and written with one tail atomic per kAppendBuf ids (or at
// the end) instead of one atomic per step.  The order of ids in the list
// does not matter to any result.  n is wave-uniform.
struct Appender {
  uint32_t *buf;   // this wave's LDS buffer
  uint32_t n;
  uint32_t *dst;   // list
  uint32_t *tail;  // its length counter
  uint32_t cap;    // its capacity (overflow sets the control block's flag)
};

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void app_flush(Appender &a, axbc::BCtl *ctl) {
  if (a.n == 0) return;
  const int lane = lane_id();
  uint32_t b = 0;
  if (lane == 0) b = atomicAdd(a.tail, a.n);
  b = rdl(b, 0);
  lds_wave_sync();
  for (uint32_t k = (uint32_t)lane; k < a.n; k += kWave) {
    const uint32_t slot = b + k;
    if (slot < a.cap)
      a.dst[slot] = a.buf[k];
    else
      atomicOr(&ctl->overflow, 1u);
  }
  lds_wave_sync();
  a.n = 0;
}

// The last flush of a kernel, by the whole workgroup (every thread calls
// it): one tail atomic for the ids of all its waves.  Atomics on one
// address are served one at a time (about 11 ns each on MI355X), so a
// level whose every wave appends pays for 8192 of them per kernel with
// per-wave flushes and for 2048 with this one.
__device__ __forceinline__ void app_flush_group(Appender &a, axbc::BCtl *ctl) {
  __shared__ uint32_t s_n[kBlock / kWave];
  __shared__ uint32_t s_base;
  const int wv = (int)threadIdx.x / kWave, lane = lane_id();
  if (lane == 0) s_n[wv] = a.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < (int)(blockDim.x / kWave); ++k) t += s_n[k];
    s_base = t ? atomicAdd(a.tail, t) : 0u;
  }
  __syncthreads();
  uint32_t b = s_base;
  for (int k = 0; k < wv; ++k) b += s_n[k];
  for (uint32_t k = (uint32_t)lane; k < a.n; k += kWave) {
    const uint32_t slot = b + k;
    if (slot < a.cap)
      a.dst[slot] = a.buf[k];
    else
      atomicOr(&ctl->overflow, 1u);
  }
  a.n = 0;
}

// Appends v from every lane with pred (all lanes active).
__device__ __forceinline__ void app_push(Appender &a, bool pred, uint32_t v, axbc::BCtl *ctl) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  const uint32_t c = (uint32_t)__popcll(m);
  if (a.n + c > (uint32_t)kAppendBuf) app_flush(a, ctl);
  const int lane = lane_id();
  if (pred) a.buf[a.n + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
  a.n += c;
}

// Split vertices.  An entry with more than `ch` edges is claimed with one
// 64-bit atomic that returns both its slot k and the first of its
// ceil(d/ch) chunk items (split_claim = entries << 40 | items), so sp_base[]
// increases with k, and the item -> slot map after the slots is filled.  The
// push and path-count split kernels give each wave one contiguous range of
// items (a binary search over sp_base for the first, then a forward walk);
// the bottom-up and dependency split kernels hand items out one at a time
// (hand_next) and find their slot in the map.
// The lanes of a wave claim together (pred: this lane's entry i is split,
// with nch items): one atomic for the wave, slots and items in lane order.
// Atomics on one address are served one at a time (about 11 ns each), and a
// level of a large graph can hold 10^5 split entries.
__device__ __forceinline__ void claim_split_wave(axbc::BCtl *ctl, bool pred, uint32_t i, uint32_t nch,
                                                 uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap,
                                                 uint64_t item_cap) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  const int lane = lane_id();
  const uint32_t items = pred ? nch : 0u;
  const uint32_t incl = wave_incl_scan(items);
  const uint32_t total = rdl(incl, kWave - 1);
  unsigned long long old = 0;
  if (lane == 0) old = atomicAdd(&ctl->split_claim, ((unsigned long long)__popcll(m) << 40) + total);
  old = __shfl(old, 0);
  const uint32_t k = (uint32_t)(old >> 40) + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  const uint64_t base = (old & ((1ull << 40) - 1)) + (incl - items);
  const bool fits = k < sp_cap && base + nch <= item_cap;
  if (pred) {
    if (fits) {
      sp_e[k] = i;
      sp_base[k] = base;
    } else {
      atomicOr(&ctl->overflow, 1u);
    }
  }
  // the item -> slot map after the sp_cap slots of sp_e (item_cap words):
  // each claimed entry's items, written by the whole wave
  uint32_t *item_slot = sp_e + sp_cap;
  unsigned long long todo = __ballot(pred && fits);
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const uint32_t ks = rdl(k, src), ns = rdl(nch, src);
    const uint64_t bs = __shfl(base, src);
    for (uint32_t t = (uint32_t)lane; t < ns; t += kWave) item_slot[bs + t] = ks;
  }
}

// Dynamic hand-out of `total` chunk items, one per claim: the items are cut
// into kHandSlices slices with a counter each, a wave starts at its XCD's
// slice (workgroups are dealt to the XCDs round robin) and moves on to the
// next slice when its own is used up.  Atomics on one address are served
// one at a time, so a single counter would bound the launch at ~10 ns per
// item.  Returns false when every slice is used up.  A launch of fewer
// than kHandMin items per wave deals them out in a fixed order instead
// (item w + t waves to wave w): each claim is a round trip to the counter's
// line, and the slices a wave tries once its own is used up cost a round
// trip each, which on small levels cost more than the balance gains.
constexpr uint64_t kHandMin = 4;

struct Hand {
  uint64_t per, next, step;
  int c, left;
  bool fixed;
};

__device__ __forceinline__ Hand hand_begin(uint64_t total) {
  Hand h;
  h.per = (total + kHandSlices - 1) / kHandSlices;
  h.c = (int)(blockIdx.x % kHandSlices);
  h.left = kHandSlices;
  h.step = (uint64_t)gridDim.x * blockDim.x / kWave;
  h.next = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  h.fixed = total < kHandMin * h.step;
  return h;
}

__device__ __forceinline__ bool hand_next(Hand &h, unsigned long long *ctr, uint64_t total, uint64_t &item) {
  if (h.fixed) {
    item = h.next;
    h.next += h.step;
    return item < total;
  }
  while (h.left > 0) {
    const uint64_t lo = (uint64_t)h.c * h.per, hi = min(total, lo + h.per);
    unsigned long long got = ~0ull;
    if (lo < hi && lane_id() == 0) got = atomicAdd(&ctr[h.c * kHandStride], 1ull);
    got = __shfl(got, 0);
    if (lo < hi && got < hi - lo) {
      item = lo + got;
      return true;
    }
    h.c = (h.c + 1) % kHandSlices;
    --h.left;
  }
  return false;
}

// The slot of chunk item `item` and its item range [base, next) (total =
// the items claimed, ns = the slots claimed).
__device__ __forceinline__ uint32_t item_entry(const uint32_t *__restrict__ sp_e,
                                               const uint64_t *__restrict__ sp_base, uint32_t sp_cap, uint32_t ns,
                                               uint64_t total, uint64_t item, uint64_t &base, uint64_t &next) {
  const uint32_t k = min(sp_e[sp_cap + item], ns - 1);
  base = sp_base[k];
  next = k + 1 < ns ? sp_base[k + 1] : total;
  return k;
}

struct ItemWalk {
  uint32_t ns, k;
  uint64_t lo, hi, next;
};

// This wave's item range [lo, hi) and the entry holding lo.
__device__ __forceinline__ ItemWalk walk_begin(const axbc::BCtl *ctl, const uint64_t *__restrict__ sp_base,
                                               uint32_t sp_cap, uint64_t item_cap) {
  const unsigned long long claim = ctl->split_claim;
  ItemWalk r;
  r.ns = min((uint32_t)(claim >> 40), sp_cap);
  const uint64_t total = min((uint64_t)(claim & ((1ull << 40) - 1)), item_cap);
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  const uint64_t per = (total + nwv - 1) / nwv;
  r.lo = min(total, (uint64_t)gw * per);
  r.hi = min(total, r.lo + per);
  r.k = 0;
  r.next = ~0ull;
  if (r.lo < r.hi && r.ns > 0) {
    uint32_t lo = 0, hi = r.ns;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (sp_base[mid] <= r.lo)
        lo = mid;
      else
        hi = mid;
    }
    r.k = lo;
    r.next = lo + 1 < r.ns ? sp_base[lo + 1] : ~0ull;
  } else {
    r.hi = r.lo;
  }
  return r;
}

// Advance to the entry holding `item`; true when the entry changed.
__device__ __forceinline__ bool walk_to(ItemWalk &r, const uint64_t *__restrict__ sp_base, uint64_t item) {
  bool moved = false;
  while (item >= r.next) {
    ++r.k;
    r.next = r.k + 1 < r.ns ? sp_base[r.k + 1] : ~0ull;
    moved = true;
  }
  return moved;
}

// ---------------------------------------------------------------------------
// Batch set-up
// ---------------------------------------------------------------------------

template <int NW>
__global__ void k_bfill(uint32_t n, const uint32_t *__restrict__ reach, uint64_t *vis, uint64_t *nx,
                        uint64_t *fr0, uint64_t *fr1, uint32_t *sig, double *dlt, uint32_t *mark,
                        uint32_t *seen) {
  constexpr int B = 64 * NW;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    for (int j = 0; j < NW; ++j) {
      vis[(size_t)v * NW + j] = 0;
      nx[(size_t)v * NW + j] = 0;
      fr0[(size_t)v * NW + j] = 0;
      fr1[(size_t)v * NW + j] = 0;
    }
    for (int s = 0; s < B; ++s) {
      sig[(size_t)v * B + s] = 0;
      dlt[(size_t)v * B + s] = (double)reach[v];
    }
    if (mark) mark[v] = 0;
    seen[v] = 0;
  }
}

// Undo the previous batch on the vertices it reached (seen == its stamp):
// their VIS and frontier words, and the path-count rows of the vertices
// with more than `chunk` pull edges, whose split items add into the row
// atomically.  Every other row entry is stored before it is read in a batch
// (sigma when its root reaches the vertex, delta by the dependency pass,
// k_bleaf, k_broots or k_bfix), so those rows are left as they are.
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_breset(uint32_t n, const uint32_t *__restrict__ seen, uint32_t stamp,
                                                   const uint32_t *__restrict__ rcol, uint32_t chunk,
                                                   uint64_t *vis, uint64_t *fr0, uint64_t *fr1, uint32_t *sig) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t vb = gw * kWave; vb < n; vb += nwv * kWave) {
    const uint32_t v = vb + (uint32_t)lane;
    const bool hit = v < n && seen[v] == stamp;
    if (hit) {
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        vis[(size_t)v * NW + j] = 0;
        fr0[(size_t)v * NW + j] = 0;
        fr1[(size_t)v * NW + j] = 0;
      }
    }
    unsigned long long todo = __ballot(hit && rcol[v + 1] - rcol[v] > chunk);
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t vv = vb + (uint32_t)i;
#pragma unroll
      for (int j = 0; j < NW; ++j) sig[(size_t)vv * B + lane + 64 * j] = 0;
    }
  }
}

// Zero the path-count rows of every vertex with more than `chunk` pull
// edges (when the split threshold drops below the one the rows were last
// cleared for).
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bsig_clear(uint32_t n, const uint32_t *__restrict__ rcol,
                                                       uint32_t chunk, uint32_t *sig) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t vb = gw * kWave; vb < n; vb += nwv * kWave) {
    const uint32_t v = vb + (uint32_t)lane;
    unsigned long long todo = __ballot(v < n && rcol[v + 1] - rcol[v] > chunk);
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t vv = vb + (uint32_t)i;
#pragma unroll
      for (int j = 0; j < NW; ++j) sig[(size_t)vv * B + lane + 64 * j] = 0;
    }
  }
}

// After the forward sweep, before the dependency pass strips VIS: delta =
// reach for every (vertex, root) the batch did not reach -- the lanes
// outside VIS of the vertices it reached, and whole rows of the vertices
// the previous batch reached (seen == prev) and this one did not.  Vertices
// neither batch reached keep rows equal to reach, so every row entry a
// reader sees is the batch's delta or, unreached, reach (the reference's
// init_delta).
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bfix(uint32_t n, const uint32_t *__restrict__ seen, uint32_t stamp,
                                                 uint32_t prev, int has_prev, const uint64_t *__restrict__ vis,
                                                 const uint32_t *__restrict__ reach, double *dlt) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t vb = gw * kWave; vb < n; vb += nwv * kWave) {
    const uint32_t v = vb + (uint32_t)lane;
    const uint32_t st = v < n ? seen[v] : 0u;
    Words<NW> open;
#pragma unroll
    for (int j = 0; j < NW; ++j) open.w[j] = 0;
    if (v < n && st == stamp) {
      const Words<NW> vw = ldw<NW>(vis, v);
#pragma unroll
      for (int j = 0; j < NW; ++j) open.w[j] = ~vw.w[j];
    } else if (v < n && has_prev && st == prev) {
#pragma unroll
      for (int j = 0; j < NW; ++j) open.w[j] = ~0ull;
    }
    const bool any = open.any();
    const double r = any ? (double)reach[v] : 0.0;
    unsigned long long todo = __ballot(any);
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t vv = vb + (uint32_t)i;
      const double rr = rdl(r, i);
#pragma unroll
      for (int j = 0; j < NW; ++j)
        if (bit(rdl(open.w[j], i), lane)) dlt[(size_t)vv * B + lane + 64 * j] = rr;
    }
  }
}

// Level 0: the roots.  One thread per root.
template <int NW>
__global__ void k_broots(const uint32_t *__restrict__ src, int nb, uint64_t *vis, uint64_t *fr0, uint32_t *sig,
                         uint32_t *q, uint64_t *qw, uint32_t *qoff, uint64_t *act, axbc::BCtl *ctl,
                         unsigned long long *cnt, uint32_t stamp, uint32_t *seen, unsigned long long in_edges,
                         uint32_t mark_base, uint64_t all0, uint64_t all1, const uint32_t *__restrict__ reach,
                         double *dlt, uint8_t *lmode) {
  constexpr int B = 64 * NW;
  const int s = threadIdx.x + blockIdx.x * blockDim.x;
  if (s < nb) {
    const uint32_t v = src[s];
    atomicOr((unsigned long long *)&vis[(size_t)v * NW + s / 64], 1ull << (s % 64));
    atomicOr((unsigned long long *)&fr0[(size_t)v * NW + s / 64], 1ull << (s % 64));
    sig[(size_t)v * B + s] = 1;
    // the root's own delta stays reach unless the depth-0 pass runs
    dlt[(size_t)v * B + s] = (double)reach[v];
    q[s] = v;
    for (int j = 0; j < NW; ++j) qw[(size_t)s * NW + j] = j == s / 64 ? 1ull << (s % 64) : 0ull;
    seen[v] = stamp;
    cnt[4 * s + 0] = 1;  // nvisited (root)
    cnt[4 * s + 1] = 0;  // reach sum of levels > 0
    cnt[4 * s + 2] = 0;  // e_r (root's degree added by the first push)
    cnt[4 * s + 3] = 0;  // e_dag
  }
  if (s == 0) {
    lmode[0] = 1;  // one root per entry
    qoff[0] = 0;
    qoff[1] = (uint32_t)nb;
    for (int j = 0; j < NW; ++j) {
      uint64_t a = 0;
      for (int t = 0; t < nb; ++t)
        if (t / 64 == j) a |= 1ull << (t % 64);
      act[j] = a;
    }
    ctl->tail = (uint32_t)nb;
    ctl->done = 0;
    ctl->empty_lev = 0;
    ctl->overflow = 0;
    ctl->wg_done = 0;
    ctl->split_claim = 0;
    for (int c = 0; c < kHandSlices; ++c) ctl->hand_pull[c * kHandStride] = ctl->hand_vert[c * kHandStride] = 0;
    ctl->bu = 0;
    ctl->cand_edges = in_edges;
    ctl->level = 1;
    ctl->mark_base = mark_base;
    ctl->stamp = stamp;
    ctl->all[0] = all0;
    ctl->all[1] = all1;
    ctl->depth = -1;
    ctl->dmin = 0;
  }
}

// ---------------------------------------------------------------------------
// Forward level L
// ---------------------------------------------------------------------------

// Mark w from a frontier word f; returns true for the one caller that must
// append w to the level's list.
template <int NW>
__device__ __forceinline__ bool push_edge(uint32_t w, const Words<NW> &f, const uint64_t *__restrict__ vis,
                                          uint64_t *nx, uint32_t *mark, uint32_t L) {
  bool first = false;
  // bits already in NX need no atomic (NX only grows during the level, and
  // whoever set them has decided the append)
  if (NW == 1) {
    const uint64_t vw = vis[w], xw = nx[w];
    const uint64_t nw = f.w[0] & ~(vw | xw);
    if (nw) {
      const uint64_t old = atomicOr((unsigned long long *)&nx[w], (unsigned long long)nw);
      first = old == 0ull;
    }
  } else {
    bool any = false;
    uint64_t vw[NW], xw[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      vw[j] = vis[(size_t)w * NW + j];
      xw[j] = nx[(size_t)w * NW + j];
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const uint64_t nw = f.w[j] & ~(vw[j] | xw[j]);
      if (nw) {
        atomicOr((unsigned long long *)&nx[(size_t)w * NW + j], (unsigned long long)nw);
        any = true;
      }
    }
    if (any && mark[w] != L) first = atomicExch(&mark[w], L) != L;
  }
  return first;
}

template <int NW>
__device__ __forceinline__ void bpush_body(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                  const uint64_t *__restrict__ vis, uint64_t *nx, uint32_t *mark,
                                                  uint64_t *fr0, uint64_t *fr1, uint32_t *q,
                                                  const uint64_t *__restrict__ qw, const uint32_t *__restrict__ qoff,
                                                  axbc::BCtl *ctl, uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap,
                                                  uint64_t item_cap, uint32_t qcap, unsigned long long *kb,
                                                  uint32_t kChunk) {
  if (ld_flag(&ctl->done) || ctl->bu) return;
  const int L = ctl->level;
  const uint32_t mark_id = ctl->mark_base + (uint32_t)L;
  uint64_t *fr_clear = (L & 1) ? fr1 : fr0;
  const int lane = lane_id();
  unsigned long long bytes = 0;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  // clear the dense frontier buffer of level L-2 (it receives level L)
  if (L >= 2) {
    const uint32_t cb = qoff[L - 2], ce = qoff[L - 1];
    for (uint32_t i = cb + blockIdx.x * blockDim.x + threadIdx.x; i < ce; i += gridDim.x * blockDim.x) {
      const uint32_t v = q[i];
#pragma unroll
      for (int j = 0; j < NW; ++j) fr_clear[(size_t)v * NW + j] = 0;
    }
  }
  __shared__ uint32_t s_app[kBlock / kWave][kAppendBuf];
  Appender app{s_app[threadIdx.x / kWave], 0, q, &ctl->tail, qcap};
  const uint32_t fb = qoff[L - 1], fe = qoff[L];
  for (uint32_t base = fb + gw * kWave; base < fe; base += nwv * kWave) {
    const uint32_t i = base + lane;
    uint32_t u = 0, st = 0, d = 0;
    Words<NW> f;
#pragma unroll
    for (int j = 0; j < NW; ++j) f.w[j] = 0;
    if (i < fe) {
      u = q[i];
      st = col[u];
      d = col[u + 1] - st;
      f = ldw<NW>(qw, i);
    }
    const bool spl = d > kChunk;  // split: its chunks go to k_bpush_split
    claim_split_wave(ctl, spl, i, spl ? (d + kChunk - 1) / kChunk : 0u, sp_e, sp_base, sp_cap, item_cap);
    if (spl) d = 0;
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t total = rdl(incl, kWave - 1);
    // entries: id 4 + offsets 8 + word 8NW; per edge: id 4 + VIS 8NW
    bytes += (unsigned long long)__popcll(__ballot(i < fe)) * (12 + 8 * NW) +
             (unsigned long long)total * (4 + 8 * NW);
    for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
      const uint32_t e = e0 + lane;
      const int own = wave_owner(incl, e) & (kWave - 1);
      const uint32_t ost = __shfl(st, own);
      const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
      const uint32_t oex = own ? pin : 0u;
      const Words<NW> of = shflw<NW>(f, own);
      bool first = false;
      uint32_t w = 0;
      if (e < total) {
        w = row[ost + (e - oex)];
        first = push_edge<NW>(w, of, vis, nx, mark, mark_id);
      }
      // append to the list (bounded)
      bytes += 4ull * (unsigned long long)__popcll(__ballot(first));  // appended ids
      app_push(app, first, w, ctl);
    }
  }
  app_flush_group(app, ctl);
  kb_add(kb, bytes);
}

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bpush(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                  const uint64_t *__restrict__ vis, uint64_t *nx, uint32_t *mark,
                                                  uint64_t *fr0, uint64_t *fr1, uint32_t *q,
                                                  const uint64_t *__restrict__ qw, const uint32_t *__restrict__ qoff,
                                                  axbc::BCtl *ctl, uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap,
                                                  uint64_t item_cap, uint32_t qcap, unsigned long long *kb,
                                                  uint32_t kChunk) { bpush_body<NW>(col, row, vis, nx, mark, fr0, fr1, q, qw, qoff, ctl, sp_e, sp_base, sp_cap, item_cap, qcap, kb, kChunk); }

// Frontier entries with more than kChunk edges, kChunk edges per item.
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bpush_split(const uint32_t *__restrict__ col,
                                                        const uint32_t *__restrict__ row,
                                                        const uint64_t *__restrict__ vis, uint64_t *nx,
                                                        uint32_t *mark, uint32_t *q, const uint64_t *__restrict__ qw,
                                                        axbc::BCtl *ctl, const uint32_t *__restrict__ sp_e,
                                                        const uint64_t *__restrict__ sp_base, uint32_t sp_cap,
                                                        uint64_t item_cap, uint32_t qcap, unsigned long long *kb,
                                                        uint32_t kChunk) {
  if (ld_flag(&ctl->done) || ctl->bu) return;
  const uint32_t mark_id = ctl->mark_base + (uint32_t)ctl->level;
  const int lane = lane_id();
  ItemWalk r = walk_begin(ctl, sp_base, sp_cap, item_cap);
  __shared__ uint32_t s_app[kBlock / kWave][kAppendBuf];
  Appender app{s_app[threadIdx.x / kWave], 0, q, &ctl->tail, qcap};
  unsigned long long bytes = 0;
  uint32_t i = 0, u = 0, st = 0, d = 0;
  Words<NW> f;
#pragma unroll
  for (int j = 0; j < NW; ++j) f.w[j] = 0;
  uint64_t base = 0;
  if (r.lo < r.hi) {
    i = sp_e[r.k];
    u = q[i];
    st = col[u];
    d = col[u + 1] - st;
    f = ldw<NW>(qw, i);
    base = sp_base[r.k];
  }
  for (uint64_t item = r.lo; item < r.hi; ++item) {
    if (walk_to(r, sp_base, item)) {
      i = sp_e[r.k];
      u = q[i];
      st = col[u];
      d = col[u + 1] - st;
      f = ldw<NW>(qw, i);
      base = sp_base[r.k];
    }
    const uint32_t cb = (uint32_t)(item - base) * kChunk, ce = min(d, cb + kChunk);
    if (cb >= ce) continue;
    bytes += (unsigned long long)(ce - cb) * (4 + 8 * NW);
    for (uint32_t k0 = cb; k0 < ce; k0 += kWave) {
      const uint32_t k = k0 + lane;
      bool first = false;
      uint32_t w = 0;
      if (k < ce) {
        w = row[st + k];
        first = push_edge<NW>(w, f, vis, nx, mark, mark_id);
      }
      bytes += 4ull * (unsigned long long)__popcll(__ballot(first));
      app_push(app, first, w, ctl);
    }
  }
  app_flush_group(app, ctl);
  kb_add(kb, bytes);
}

// sigma pull for the vertices new at level L: q[qoff[L] .. tail).  A wave
// per vertex; lanes = roots.  Vertices with more than kChunk pull edges are
// split into kChunk work items that add their partial counts atomically.
template <int NW>
__device__ __forceinline__ void sigma_chunk(uint32_t st, uint32_t cb, uint32_t ce, const uint32_t *__restrict__ rrow,
                                            const uint64_t *__restrict__ frp, const Words<NW> &nw,
                                            const uint32_t *__restrict__ sig, uint32_t acc[NW], uint32_t dag[NW],
                                            uint32_t *contrib, uint64_t found[NW]) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  for (uint32_t k0 = cb; k0 < ce; k0 += kWave) {
    const uint32_t k = k0 + lane;
    uint32_t u = 0;
    Words<NW> m;
#pragma unroll
    for (int j = 0; j < NW; ++j) m.w[j] = 0;
    if (k < ce) {
      u = rrow[st + k];
      const Words<NW> f = ldw<NW>(frp, u);
#pragma unroll
      for (int j = 0; j < NW; ++j) m.w[j] = f.w[j] & nw.w[j];
    }
    // edges with roots to add, kBurst at a time: all row loads of a burst
    // are issued before any is used
    unsigned long long todo = __ballot(m.any());
    while (todo) {
      uint32_t got[kBurst][NW];
#pragma unroll
      for (int t = 0; t < kBurst; ++t) {
        const int i = todo ? __ffsll((long long)todo) - 1 : 0;
        const bool live = todo != 0;
        todo &= todo - 1;
        const uint32_t uu = rdl(u, i);
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          const uint64_t mm = live ? rdl(m.w[j], i) : 0ull;
          *contrib += (uint32_t)__popcll(mm);
          found[j] |= mm;
          got[t][j] = bit(mm, lane) ? sig[(size_t)uu * B + lane + 64 * j] : 0u;
          dag[j] += bit(mm, lane) ? 1u : 0u;
        }
      }
#pragma unroll
      for (int t = 0; t < kBurst; ++t)
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[j] += got[t][j];
    }
  }
}

// Path counts of up to 64 vertices at once (one per lane: st = first pull
// edge, d = pull degree, 0 to skip, mask = roots to look for).  Their edges
// are dealt 64 per step in vertex order; for an edge u -> owner with
// m = FR[u] & mask(owner) != 0, lanes = roots add sigma_s[u].  Set edges are
// taken in lane order, kBurst row loads in flight, so every owner's sums
// are complete when the owner changes: flush(owner lane, acc, found) then
// receives them (found = the roots that had a predecessor).
template <int NW, class Flush>
__device__ __forceinline__ void group_sigma(uint32_t st, uint32_t d, const Words<NW> &mask,
                                            const uint32_t *__restrict__ rrow, const uint64_t *__restrict__ frp,
                                            const uint32_t *__restrict__ sig, uint32_t dag[NW], uint32_t *contrib,
                                            Flush flush) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  const uint32_t incl = wave_incl_scan(d);
  const uint32_t total = rdl(incl, kWave - 1);
  int cur = -1;
  uint32_t acc[NW];
  uint64_t found[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) acc[j] = 0, found[j] = 0;
  for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
    const uint32_t e = e0 + lane;
    const int own = wave_owner(incl, e) & (kWave - 1);
    const uint32_t ost = __shfl(st, own);
    const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
    const uint32_t oex = own ? pin : 0u;
    const Words<NW> om = shflw<NW>(mask, own);
    uint32_t u = 0;
    Words<NW> m;
#pragma unroll
    for (int j = 0; j < NW; ++j) m.w[j] = 0;
    if (e < total) {
      u = rrow[ost + (e - oex)];
      const Words<NW> f = ldw<NW>(frp, u);
#pragma unroll
      for (int j = 0; j < NW; ++j) m.w[j] = f.w[j] & om.w[j];
    }
    unsigned long long todo = __ballot(m.any());
    while (todo) {
      uint32_t got[kBurst][NW];
      uint64_t mmt[kBurst][NW];
      int ot[kBurst];
#pragma unroll
      for (int t = 0; t < kBurst; ++t) {
        const int i = todo ? __ffsll((long long)todo) - 1 : 0;
        const bool live = todo != 0;
        todo &= todo - 1;
        const uint32_t uu = rdl(u, i);
        ot[t] = live ? (int)rdl((uint32_t)own, i) : -1;
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          const uint64_t mm = live ? rdl(m.w[j], i) : 0ull;
          mmt[t][j] = mm;
          *contrib += (uint32_t)__popcll(mm);
          got[t][j] = bit(mm, lane) ? sig[(size_t)uu * B + lane + 64 * j] : 0u;
          dag[j] += bit(mm, lane) ? 1u : 0u;
        }
      }
#pragma unroll
      for (int t = 0; t < kBurst; ++t) {
        if (ot[t] < 0) continue;
        if (ot[t] != cur) {
          if (cur >= 0) flush(cur, acc, found);
          cur = ot[t];
#pragma unroll
          for (int j = 0; j < NW; ++j) acc[j] = 0, found[j] = 0;
        }
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[j] += got[t][j], found[j] |= mmt[t][j];
      }
    }
  }
  if (cur >= 0) flush(cur, acc, found);
}

// Path counts of the small new vertices of a group of 64 (lane = vertex;
// small: at most kSparseDeg pull edges), one (vertex, root) pair per lane:
// sigma_s(w) = sum of sigma_s(u) over u in radj(w) with s in FR[u] (uint32,
// wrapping; exact in any order).  DAG edges go to the wave's slot.
// Returns the wave's number of terms.
template <int NW>
__device__ __forceinline__ uint32_t sigma_pairs(bool small, uint32_t w, const Words<NW> &nw, uint32_t st, uint32_t d,
                                                const uint32_t *__restrict__ rrow, const uint64_t *__restrict__ frp,
                                                uint32_t *sig, unsigned long long *wslot, uint32_t gw) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  uint32_t np = 0;
  if (small)
#pragma unroll
    for (int j = 0; j < NW; ++j) np += (uint32_t)__popcll(nw.w[j]);
  const uint32_t incl = wave_incl_scan(np);
  const uint32_t total = rdl(incl, kWave - 1);
  uint32_t terms = 0;
  for (uint32_t p0 = 0; p0 < total; p0 += kWave) {
    const uint32_t p = p0 + (uint32_t)lane;
    const int own = wave_owner(incl, p) & (kWave - 1);
    const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
    const uint32_t ex = own ? pin : 0u;
    const uint32_t ow = __shfl(w, own), ost = __shfl(st, own), od = __shfl(d, own);
    const Words<NW> onw = shflw<NW>(nw, own);
    if (p < total) {
      const int s = select_root<NW>(onw.w, p - ex);
      const int sj = s >> 6;
      const uint64_t sb = 1ull << (s & 63);
      uint32_t acc = 0, dag = 0;
      for (uint32_t e = 0; e < od; e += 4) {
        uint32_t u[4];
        bool use[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) u[t] = e + t < od ? rrow[ost + e + t] : 0u;
#pragma unroll
        for (int t = 0; t < 4; ++t) use[t] = e + t < od && (frp[(size_t)u[t] * NW + sj] & sb);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (use[t]) {
            acc += sig[(size_t)u[t] * B + s];
            ++dag;
          }
      }
      sig[(size_t)ow * B + s] = acc;
      slot_add_root(wslot, 3, gw, B, s, dag);
      terms += dag;
    }
  }
  return (uint32_t)wave_sum_u64(terms);
}

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bsigma_split(const uint32_t *__restrict__ rcol,
                                                         const uint32_t *__restrict__ rrow,
                                                         const uint64_t *__restrict__ qw, const uint64_t *fr0,
                                                         const uint64_t *fr1, uint32_t *sig,
                                                         const uint32_t *__restrict__ q, axbc::BCtl *ctl,
                                                         unsigned long long *wslot, const uint32_t *__restrict__ sp_e,
                                                         const uint64_t *__restrict__ sp_base, uint32_t sp_cap,
                                                         uint64_t item_cap, unsigned long long *kb, uint32_t kChunk) {
  constexpr int B = 64 * NW;
  if (ld_flag(&ctl->done) || ctl->bu) return;
  const uint64_t *frp = (ctl->level & 1) ? fr0 : fr1;
  const int lane = lane_id();
  ItemWalk r = walk_begin(ctl, sp_base, sp_cap, item_cap);
  if (r.lo >= r.hi) return;
  unsigned long long bytes = 0;
  uint32_t dag_tot[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) dag_tot[j] = 0;
  uint32_t w = q[sp_e[r.k]];
  uint32_t st = rcol[w], d = rcol[w + 1] - st;
  Words<NW> nw = ldw<NW>(qw, sp_e[r.k]);
  uint64_t base = sp_base[r.k];
  bool first_item = r.lo == base;
  for (uint64_t item = r.lo; item < r.hi; ++item) {
    if (walk_to(r, sp_base, item)) {
      w = q[sp_e[r.k]];
      st = rcol[w];
      d = rcol[w + 1] - st;
      nw = ldw<NW>(qw, sp_e[r.k]);
      base = sp_base[r.k];
      first_item = true;
    }
    const uint32_t c = (uint32_t)(item - base);
    uint32_t acc[NW], dag[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) acc[j] = dag[j] = 0;
    uint32_t contrib = 0;
    const uint32_t cb = c * kChunk, ce = min(d, cb + kChunk);
    if (cb >= ce) continue;
    uint64_t found[NW] = {};
    sigma_chunk<NW>(st, cb, ce, rrow, frp, nw, sig, acc, dag, &contrib, found);
    uint32_t roots = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      if (acc[j]) atomicAdd(&sig[(size_t)w * B + lane + 64 * j], acc[j]);
      dag_tot[j] += dag[j];
      roots += (uint32_t)__popcll(nw.w[j]);
    }
    // per item: edges id 4 + frontier word 8NW, terms 4; per entry (once):
    // word 8NW + sigma row write 4 per root (id and offsets: k_bsigma)
    bytes += (unsigned long long)(ce - cb) * (4 + 8 * NW) + 4ull * contrib;
    if (first_item) bytes += 8 * NW + 4ull * roots;
    first_item = false;
  }
  kb_add(kb, bytes);
  slot_add<NW>(wslot, 3, (blockIdx.x * blockDim.x + threadIdx.x) / kWave, dag_tot);
}

// Bottom-up level L (direction-optimising BFS, one pass for discovery and
// path counts): every vertex w not yet reached by all roots pulls over its
// in-edges, m = FR[u] & ~VIS[w]; the roots found, NX[w] = OR of m, reach w
// at level L with sigma_s[w] = sum of sigma_s[u] over u with s in m.  The
// same sums as push + k_bsigma, so the result does not depend on the
// direction.  A wave takes 64 consecutive vertices, one at a time, and
// appends the new ones with one atomic.  Vertices with more than kChunk
// in-edges go to k_bpull_split.
template <int NW>
__device__ __forceinline__ void bpull_body(uint32_t n, const uint32_t *__restrict__ rcol,
                                                  const uint32_t *__restrict__ rrow, const uint64_t *__restrict__ vis,
                                                  uint64_t *nx, uint64_t *fr0, uint64_t *fr1, uint32_t *sig,
                                                  uint32_t *q, const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
                                                  unsigned long long *wslot, uint32_t *sp_e, uint64_t *sp_base,
                                                  uint32_t sp_cap, uint64_t item_cap, uint32_t qcap,
                                                  unsigned long long *kb, uint32_t kChunk,
                                                  unsigned long long *hand) {
  constexpr int B = 64 * NW;
  if (ld_flag(&ctl->done) || !ctl->bu) return;
  const int L = ctl->level;
  const uint64_t *frp = (L & 1) ? fr0 : fr1;  // level L-1
  uint64_t *fr_clear = (L & 1) ? fr1 : fr0;   // receives level L
  const uint64_t all[2] = {ctl->all[0], ctl->all[1]};
  // clear the dense frontier buffer of level L-2 (it receives level L)
  if (L >= 2) {
    const uint32_t cb = qoff[L - 2], ce = qoff[L - 1];
    for (uint32_t i = cb + blockIdx.x * blockDim.x + threadIdx.x; i < ce; i += gridDim.x * blockDim.x) {
      const uint32_t v = q[i];
#pragma unroll
      for (int j = 0; j < NW; ++j) fr_clear[(size_t)v * NW + j] = 0;
    }
  }
  unsigned long long bytes = 0;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  uint32_t dag_tot[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) dag_tot[j] = 0;
  __shared__ uint32_t s_app[kBlock / kWave][kAppendBuf];
  Appender app{s_app[threadIdx.x / kWave], 0, q, &ctl->tail, qcap};
  // groups of 64 vertices: handed out one at a time (hand), else the groups
  // gw + t nwv
  const uint64_t ng = ((uint64_t)n + kWave - 1) / kWave;
  Hand h = hand_begin(ng);
  uint64_t g = gw;
  for (;;) {
    if (hand) {
      if (!hand_next(h, hand, ng, g)) break;
    } else if (g >= ng) {
      break;
    }
    const uint32_t vb = (uint32_t)(g * kWave);
    if (!hand) g += nwv;
    const uint32_t v = vb + (uint32_t)lane;
    uint32_t st = 0, d = 0;
    Words<NW> open;  // roots that have not reached v
#pragma unroll
    for (int j = 0; j < NW; ++j) open.w[j] = 0;
    if (v < n) {
      st = rcol[v];
      d = rcol[v + 1] - st;
      const Words<NW> vw = ldw<NW>(vis, v);
#pragma unroll
      for (int j = 0; j < NW; ++j) open.w[j] = all[j] & ~vw.w[j];
    }
    bool cand = d > 0 && open.any();
    bytes += v < n ? 8 + 8 * NW : 0;
    const bool spl = cand && d > kChunk;
    claim_split_wave(ctl, spl, v, spl ? (d + kChunk - 1) / kChunk : 0u, sp_e, sp_base, sp_cap, item_cap);
    if (spl) cand = false;
    unsigned long long todo = __ballot(cand);
    unsigned long long fresh = 0;
    while (todo) {
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t w = vb + (uint32_t)i;
      const uint32_t sw = rdl(st, i), dw = rdl(d, i);
      const Words<NW> ow = rdlw<NW>(open, i);
      uint32_t acc[NW], dag[NW];
      uint64_t found[NW] = {};
#pragma unroll
      for (int j = 0; j < NW; ++j) acc[j] = dag[j] = 0;
      uint32_t contrib = 0;
      sigma_chunk<NW>(sw, 0, dw, rrow, frp, ow, sig, acc, dag, &contrib, found);
      uint32_t roots = 0;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        roots += (uint32_t)__popcll(found[j]);
        if (bit(found[j], lane)) sig[(size_t)w * B + lane + 64 * j] = acc[j];
        dag_tot[j] += dag[j];
      }
      if (roots) {
        if (lane < NW) nx[(size_t)w * NW + lane] = found[lane < NW ? lane : 0];
        fresh |= 1ull << i;
      }
      // per edge: id 4 + frontier word 8NW; per term: sigma 4; per root
      // reached: sigma write 4; NX words written
      bytes += (unsigned long long)dw * (4 + 8 * NW) + 4ull * contrib + 4ull * roots + (roots ? 8 * NW : 0);
    }
    bytes += 4ull * (unsigned long long)__popcll(fresh);
    app_push(app, bit(fresh, lane), v, ctl);
  }
  app_flush_group(app, ctl);
  kb_add(kb, bytes);
  slot_add<NW>(wslot, 3, (blockIdx.x * blockDim.x + threadIdx.x) / kWave, dag_tot);
}

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bpull(uint32_t n, const uint32_t *__restrict__ rcol,
                                                  const uint32_t *__restrict__ rrow, const uint64_t *__restrict__ vis,
                                                  uint64_t *nx, uint64_t *fr0, uint64_t *fr1, uint32_t *sig,
                                                  uint32_t *q, const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
                                                  unsigned long long *wslot, uint32_t *sp_e, uint64_t *sp_base,
                                                  uint32_t sp_cap, uint64_t item_cap, uint32_t qcap,
                                                  unsigned long long *kb, uint32_t kChunk) { bpull_body<NW>(n, rcol, rrow, vis, nx, fr0, fr1, sig, q, qoff, ctl, wslot, sp_e, sp_base, sp_cap, item_cap, qcap, kb, kChunk, ctl->hand_vert); }

// Level expansion in one launch: the top-down push or the bottom-up pull,
// as the previous level's close decided (one launch per level instead of
// two, one of which would exit at once).
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bexpand(
    uint32_t n, const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
    const uint32_t *__restrict__ rcol, const uint32_t *__restrict__ rrow, const uint64_t *__restrict__ vis,
    uint64_t *nx, uint32_t *mark, uint64_t *fr0, uint64_t *fr1, uint32_t *sig, uint32_t *q,
    const uint64_t *__restrict__ qw, const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
    unsigned long long *wslot, uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap, uint64_t item_cap,
    uint32_t qcap, uint32_t kChunk) {
  if (ld_flag(&ctl->done)) return;
  if (ctl->bu)
    bpull_body<NW>(n, rcol, rrow, vis, nx, fr0, fr1, sig, q, qoff, ctl, wslot, sp_e, sp_base, sp_cap, item_cap, qcap,
                   nullptr, kChunk, ctl->hand_vert);
  else
    bpush_body<NW>(col, row, vis, nx, mark, fr0, fr1, q, qw, qoff, ctl, sp_e, sp_base, sp_cap, item_cap, qcap,
                   nullptr, kChunk);
}

// Bottom-up chunks of the vertices with more than kChunk in-edges: counts
// and roots found are added atomically; the chunk that first makes NX[w]
// non-empty appends w.
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bpull_split(const uint32_t *__restrict__ rcol,
                                                        const uint32_t *__restrict__ rrow,
                                                        const uint64_t *__restrict__ vis, uint64_t *nx,
                                                        const uint64_t *fr0, const uint64_t *fr1, uint32_t *sig,
                                                        uint32_t *mark, uint32_t *q, axbc::BCtl *ctl,
                                                        unsigned long long *wslot, const uint32_t *__restrict__ sp_e,
                                                        const uint64_t *__restrict__ sp_base, uint32_t sp_cap,
                                                        uint64_t item_cap, uint32_t qcap, unsigned long long *kb,
                                                        uint32_t kChunk) {
  constexpr int B = 64 * NW;
  if (ld_flag(&ctl->done) || !ctl->bu) return;
  const uint64_t *frp = (ctl->level & 1) ? fr0 : fr1;
  const uint32_t mark_id = ctl->mark_base + (uint32_t)ctl->level;
  const uint64_t all[2] = {ctl->all[0], ctl->all[1]};
  const int lane = lane_id();
  const unsigned long long claim = ctl->split_claim;
  const uint32_t ns = min((uint32_t)(claim >> 40), sp_cap);
  const uint64_t total = min((uint64_t)(claim & ((1ull << 40) - 1)), item_cap);
  if (ns == 0 || total == 0) return;
  __shared__ uint32_t s_app[kBlock / kWave][kAppendBuf];
  Appender app{s_app[threadIdx.x / kWave], 0, q, &ctl->tail, qcap};
  bool has_items = false;
  unsigned long long bytes = 0;
  uint32_t dag_tot[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) dag_tot[j] = 0;
  uint32_t w = 0, st = 0, d = 0;
  uint64_t base = 0, next = 0;  // the current entry's items: [base, next)
  Words<NW> open;
  // items handed out one at a time, as in k_bdep_split
  Hand hand = hand_begin(total);
  uint64_t item = 0;
  while (hand_next(hand, ctl->hand_pull, total, item)) {
    has_items = true;
    if (item < base || item >= next) {
      w = sp_e[item_entry(sp_e, sp_base, sp_cap, ns, total, item, base, next)];
      st = rcol[w];
      d = rcol[w + 1] - st;
      const Words<NW> vw = ldw<NW>(vis, w);
#pragma unroll
      for (int j = 0; j < NW; ++j) open.w[j] = all[j] & ~vw.w[j];
    }
    const uint32_t cb = (uint32_t)(item - base) * kChunk, ce = min(d, cb + kChunk);
    if (cb >= ce) continue;
    uint32_t acc[NW], dag[NW];
    uint64_t found[NW] = {};
#pragma unroll
    for (int j = 0; j < NW; ++j) acc[j] = dag[j] = 0;
    uint32_t contrib = 0;
    sigma_chunk<NW>(st, cb, ce, rrow, frp, open, sig, acc, dag, &contrib, found);
    bool any = false, first = false;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      if (acc[j]) atomicAdd(&sig[(size_t)w * B + lane + 64 * j], acc[j]);
      dag_tot[j] += dag[j];
      any |= found[j] != 0;
    }
    if (any) {
      if (lane == 0) {
        if (NW == 1) {
          first = atomicOr((unsigned long long *)&nx[w], (unsigned long long)found[0]) == 0ull;
        } else {
#pragma unroll
          for (int j = 0; j < NW; ++j)
            if (found[j]) atomicOr((unsigned long long *)&nx[(size_t)w * NW + j], (unsigned long long)found[j]);
          if (mark[w] != mark_id) first = atomicExch(&mark[w], mark_id) != mark_id;
        }
      }
    }
    app_push(app, lane == 0 && first, w, ctl);
    bytes += (unsigned long long)(ce - cb) * (4 + 8 * NW) + 4ull * contrib;
  }
  app_flush_group(app, ctl);
  if (!has_items) return;
  kb_add(kb, bytes);
  slot_add<NW>(wslot, 3, (blockIdx.x * blockDim.x + threadIdx.x) / kWave, dag_tot);
}

// Close level L and count its paths, in one pass over the level's new
// entries q[qoff[L] .. tail) (lane = entry, 64 at a time): the entry's roots
// NX[w] become its list word, its frontier word and part of VIS[w] (NX is
// cleared); the per-root counters and the level sums for k_bclose_end are
// taken; and, unless the level was found bottom-up (whose pull counted the
// paths already), sigma_s(w) = sum of sigma_s(u) over u in radj(w) with s
// in FR[u] (level L-1) for the vertices of at most kChunk pull edges (the
// others are claimed for k_bsigma_split, which reads their roots from the
// list word).  The path counts read only level L-1 state, the close writes
// only level L state, so the two need no order between them.
template <int NW>
__device__ __forceinline__ void bclose_body(const uint32_t *__restrict__ col, const uint32_t *__restrict__ reach,
                                            const uint32_t *__restrict__ rcol, const uint32_t *__restrict__ rrow,
                                            uint64_t *nx, uint64_t *vis, uint64_t *fr0, uint64_t *fr1, uint32_t *sig,
                                            const uint32_t *__restrict__ q, uint64_t *qw,
                                            const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
                                            unsigned long long *wslot, unsigned long long *lslot, uint32_t *seen,
                                            uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap, uint64_t item_cap,
                                            uint32_t qcap, unsigned long long *kb, uint32_t kChunk, int strict,
                                            const uint8_t *__restrict__ lmode) {
  constexpr int B = 64 * NW;
  if (ld_flag(&ctl->done)) return;
  const int L = ctl->level;
  const bool counts = ctl->bu == 0u;
  const uint32_t stamp = ctl->stamp;
  uint64_t *frc = (L & 1) ? fr1 : fr0;        // level L
  const uint64_t *frp = (L & 1) ? fr0 : fr1;  // level L-1
  const uint64_t all[2] = {ctl->all[0], ctl->all[1]};
  const bool level1 = strict && L == 1;
  const bool sparse = lmode[L - 1] != 0;
  unsigned long long bytes = 0, fe = 0, full = 0, pairs = 0;
  const uint32_t nb = qoff[L], ne = min(ctl->tail, qcap);
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  uint32_t c_vis[NW], c_all[NW], c_er[NW], dag_tot[NW];
  uint64_t a_or[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) c_vis[j] = c_all[j] = c_er[j] = dag_tot[j] = 0, a_or[j] = 0;
  for (uint32_t base = nb + gw * kWave; base < ne; base += nwv * kWave) {
    const uint32_t i = base + lane;
    const bool valid = i < ne;
    uint32_t w = 0, rw = 0, dw = 0, st = 0, d = 0;
    Words<NW> nw;
#pragma unroll
    for (int j = 0; j < NW; ++j) nw.w[j] = 0;
    if (valid) {
      w = q[i];
      nw = ldw<NW>(nx, w);
      const Words<NW> vo = ldw<NW>(vis, w);
      rw = reach[w];
      dw = col[w + 1] - col[w];
      st = rcol[w];
      d = rcol[w + 1] - st;
      // direction bookkeeping: out-edges of the level, and in-edges of the
      // vertices every root has now reached
      bool was_full = true, is_full = true;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        was_full &= (vo.w[j] & all[j]) == all[j];
        is_full &= ((vo.w[j] | nw.w[j]) & all[j]) == all[j];
      }
      fe += dw;
      if (is_full && !was_full) full += d;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        frc[(size_t)w * NW + j] = nw.w[j];
        vis[(size_t)w * NW + j] = vo.w[j] | nw.w[j];
        qw[(size_t)i * NW + j] = nw.w[j];
        nx[(size_t)w * NW + j] = 0;
        a_or[j] |= nw.w[j];
        pairs += (unsigned long long)__popcll(nw.w[j]);
      }
      seen[w] = stamp;
    }
    const uint32_t cntv = (uint32_t)__popcll(__ballot(valid));
    for (uint32_t k = 0; k < cntv; ++k) {
      const uint32_t rk = rdl(rw, (int)k), dk = rdl(dw, (int)k);
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const uint64_t mk = rdl(nw.w[j], (int)k);
        if (bit(mk, lane)) {
          c_vis[j] += 1;
          c_all[j] += rk;
          c_er[j] += dk;
        }
      }
    }
    // entry: id 4 + NX 8NW + reach 4 + offsets 8 + seen 4; writes FR, VIS
    // (read+write), list word, NX: 5 * 8NW
    bytes += (unsigned long long)cntv * (20 + 48 * NW);
    if (!counts) continue;
    if (level1) {
      // level 1 over strictly increasing lists: the frontier is the roots
      // (sigma 1 each), and a new vertex has exactly one edge from every
      // root in its word, so sigma_s(w) = 1 there without scanning its
      // in-edges (a hub next to the roots would be scanned whole); the same
      // DAG edges count
      unsigned long long todo = __ballot(nw.any());
      while (todo) {
        const int k = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const uint32_t wk = rdl(w, k);
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          const uint64_t mk = rdl(nw.w[j], k);
          bytes += 4ull * (unsigned long long)__popcll(mk);
          if (bit(mk, lane)) {
            sig[(size_t)wk * B + lane + 64 * j] = 1u;
            dag_tot[j] += 1;
          }
        }
      }
      continue;
    }
    const bool split = valid && d > kChunk;
    claim_split_wave(ctl, split, i, split ? (d + kChunk - 1) / kChunk : 0u, sp_e, sp_base, sp_cap, item_cap);
    bytes += 12ull * (unsigned long long)__popcll(__ballot(split));
    Words<NW> snw = nw;  // the roots counted here (none for a split vertex)
    if (split) {
      d = 0;
#pragma unroll
      for (int j = 0; j < NW; ++j) snw.w[j] = 0;
    }
    uint32_t contrib = 0;
    // sparse levels: the small vertices as (vertex, root) pairs
    const bool small = sparse && d > 0 && d <= kSparseDeg;
    if (sparse) contrib += sigma_pairs<NW>(small, w, snw, st, d, rrow, frp, sig, wslot, gw);
    group_sigma<NW>(st, small ? 0u : d, snw, rrow, frp, sig, dag_tot, &contrib,
                    [&](int o, const uint32_t *acc, const uint64_t *found) {
                      const uint32_t wo = rdl(w, o);
#pragma unroll
                      for (int j = 0; j < NW; ++j)
                        if (bit(found[j], lane)) sig[(size_t)wo * B + lane + 64 * j] = acc[j];
                    });
    // per vertex: pull offsets 8 + sigma write 4 per root; per edge: id 4 +
    // frontier word 8NW; per term: sigma 4
    uint32_t roots = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) roots += d ? (uint32_t)__popcll(snw.w[j]) : 0u;
    const uint32_t dsum = (uint32_t)wave_sum_u64(d), rsum = (uint32_t)wave_sum_u64(roots);
    const uint32_t nv = (uint32_t)__popcll(__ballot(valid && d > 0));
    bytes += (unsigned long long)nv * 8 + (unsigned long long)dsum * (4 + 8 * NW) + 4ull * contrib + 4ull * rsum;
  }
  fe = wave_sum_u64(fe);
  full = wave_sum_u64(full);
  pairs = wave_sum_u64(pairs);
#pragma unroll
  for (int j = 0; j < NW; ++j) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) a_or[j] |= __shfl_xor(a_or[j], off);
  }
  kb_add(kb, bytes);
  // per-root counts into this wave's slots
  slot_add<NW>(wslot, 0, gw, c_vis);
  slot_add<NW>(wslot, 1, gw, c_all);
  slot_add<NW>(wslot, 2, gw, c_er);
  if (counts && !level1 && sparse) {
    // the pairs added to this slot atomically: stay atomic
#pragma unroll
    for (int j = 0; j < NW; ++j) slot_add_root(wslot, 3, gw, B, lane + 64 * j, dag_tot[j]);
  } else {
    slot_add<NW>(wslot, 3, gw, dag_tot);
  }
  // level sums: one slot per workgroup (every workgroup writes its slot,
  // zero or not; k_bclose_end adds them)
  __shared__ unsigned long long s_ls[kBlock / kWave][5];
  const int wv = (int)threadIdx.x / kWave;
  if (lane == 0) {
    s_ls[wv][0] = fe;
    s_ls[wv][1] = full;
    s_ls[wv][2] = a_or[0];
    s_ls[wv][3] = NW > 1 ? a_or[NW - 1] : 0ull;
    s_ls[wv][4] = pairs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < (int)(blockDim.x / kWave); ++k) {
      t[0] += s_ls[k][0];
      t[1] += s_ls[k][1];
      t[2] |= s_ls[k][2];
      t[3] |= s_ls[k][3];
      t[4] += s_ls[k][4];
    }
    unsigned long long *ls = lslot + kLevelSlot * (size_t)blockIdx.x;
#pragma unroll
    for (int k = 0; k < 5; ++k) ls[k] = t[k];
  }
}

template <int NW>
__global__ __launch_bounds__(kBlock) void k_bclose(
    const uint32_t *__restrict__ col, const uint32_t *__restrict__ reach, const uint32_t *__restrict__ rcol,
    const uint32_t *__restrict__ rrow, uint64_t *nx, uint64_t *vis, uint64_t *fr0, uint64_t *fr1, uint32_t *sig,
    const uint32_t *__restrict__ q, uint64_t *qw, const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
    unsigned long long *wslot, unsigned long long *lslot, uint32_t *seen, uint32_t *sp_e, uint64_t *sp_base,
    uint32_t sp_cap, uint64_t item_cap, uint32_t qcap, unsigned long long *kb, uint32_t kChunk, int strict,
    const uint8_t *__restrict__ lmode) {
  bclose_body<NW>(col, reach, rcol, rrow, nx, vis, fr0, fr1, sig, q, qw, qoff, ctl, wslot, lslot, seen, sp_e, sp_base,
                  sp_cap, item_cap, qcap, kb, kChunk, strict, lmode);
}

// End of level L, one workgroup, after k_bclose in stream order (so every
// slot of the level is written without fences): the level's sums from the
// k_bclose workgroup slots, the level offsets, the done flag, and the direction
// of the next level.
__device__ __forceinline__ void bclose_end_body(axbc::BCtl *ctl, uint32_t *qoff, uint32_t qcap, uint32_t pull16,
                                                       const unsigned long long *__restrict__ lslot, uint64_t *act,
                                                       int nw, uint8_t *lmode, int nslots) {
  if (ld_flag(&ctl->done)) return;
  __shared__ unsigned long long s_red[kBlock / kWave][5];
  unsigned long long fe = 0, full = 0, a0 = 0, a1 = 0, pairs = 0;
  for (int g = (int)threadIdx.x; g < nslots; g += (int)blockDim.x) {
    const unsigned long long *ls = lslot + kLevelSlot * (size_t)g;
    fe += ls[0];
    full += ls[1];
    a0 |= ls[2];
    a1 |= ls[3];
    pairs += ls[4];
  }
  fe = wave_sum_u64(fe);
  full = wave_sum_u64(full);
  pairs = wave_sum_u64(pairs);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    a0 |= __shfl_xor(a0, off);
    a1 |= __shfl_xor(a1, off);
  }
  const int wv = (int)threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_red[wv][0] = fe;
    s_red[wv][1] = full;
    s_red[wv][2] = a0;
    s_red[wv][3] = a1;
    s_red[wv][4] = pairs;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < (int)(blockDim.x / kWave); ++k) {
    fe += s_red[k][0];
    full += s_red[k][1];
    a0 |= s_red[k][2];
    a1 |= s_red[k][3];
    pairs += s_red[k][4];
  }
  const int L = ctl->level;
  {
    const unsigned long long entries = ctl->tail - qoff[L];
    lmode[L] = pairs <= kSparseRatio * entries && fe <= kSparseAvgDeg * entries ? 1 : 0;
  }
  act[(size_t)L * nw] = a0;
  if (nw > 1) act[(size_t)L * nw + 1] = a1;
  ctl->cand_edges -= full;
  uint32_t tail = ctl->tail;
  if (tail > qcap) {  // list overflow: stop the sweep, the host retries larger
    tail = qcap;
    ctl->overflow = 1;
  }
  qoff[L + 1] = tail;
  if (tail == qoff[L] || ctl->overflow) {
    ctl->empty_lev = (uint32_t)L;
    ctl->done = 1u;
  }
  // next level bottom-up when the vertices still open to some root have
  // at most pull16/16 times the in-edges that a push would scan
  const unsigned long long cand = ctl->cand_edges;
  ctl->bu = pull16 && cand * 16ull <= (unsigned long long)pull16 * fe ? 1u : 0u;
  ctl->split_claim = 0;
  for (int c = 0; c < kHandSlices; ++c) ctl->hand_pull[c * kHandStride] = ctl->hand_vert[c * kHandStride] = 0;
  if (!ctl->done) ctl->level = L + 1;
}

__global__ __launch_bounds__(kBlock) void k_bclose_end(axbc::BCtl *ctl, uint32_t *qoff, uint32_t qcap, uint32_t pull16,
                                                       const unsigned long long *__restrict__ lslot, uint64_t *act,
                                                       int nw, uint8_t *lmode) { bclose_end_body(ctl, qoff, qcap, pull16, lslot, act, nw, lmode, kBGrid); }

// Per-root counters of the batch: cnt[4s + k] += the sum of the wave slots
// of kind k.  Workgroup (k, chunk) adds the slots of its chunk of waves for
// every root (thread = root), so the slot rows are read whole.
template <int NW>
__global__ void k_bcnt_reduce(const unsigned long long *__restrict__ wslot, unsigned long long *cnt) {
  constexpr int B = 64 * NW;
  const int s = (int)threadIdx.x;
  const int k = (int)blockIdx.y;
  const int chunks = (int)gridDim.x;
  const int per = (kSlotWaves + chunks - 1) / chunks;
  const int g0 = (int)blockIdx.x * per, g1 = min(kSlotWaves, g0 + per);
  if (s >= B) return;
  unsigned long long t = 0;
  for (int g = g0; g < g1; ++g) t += wslot[((size_t)k * kSlotWaves + g) * B + s];
  if (t) atomicAdd(&cnt[4 * s + k], t);
}

// Zeroes n 64-bit words.
__global__ void k_bzero(unsigned long long *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// x = (reach + 1) / sigma for the entries of the deepest level L of the
// sweep: they have no dependencies, and the dependency pass (which writes x
// of every shallower entry) never visits them.  A wave per 64 entries, then
// entry by entry with lane = root.
template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bleaf(const uint32_t *__restrict__ q,
                                                  const uint64_t *__restrict__ qw,
                                                  const uint32_t *__restrict__ qoff, int L,
                                                  const uint32_t *__restrict__ reach,
                                                  const uint32_t *__restrict__ sig, double *xr, double *dlt) {
  constexpr int B = 64 * NW;
  const uint32_t nb = qoff[L], ne = qoff[L + 1];
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t base = nb + gw * kWave; base < ne; base += nwv * kWave) {
    const uint32_t i = base + lane;
    uint32_t w = 0, rw = 0;
    Words<NW> f;
#pragma unroll
    for (int j = 0; j < NW; ++j) f.w[j] = 0;
    if (i < ne) {
      w = q[i];
      f = ldw<NW>(qw, i);
      rw = reach[w];
    }
    const uint32_t cntv = (uint32_t)__popcll(__ballot(i < ne));
    for (uint32_t k = 0; k < cntv; ++k) {
      const uint32_t wk = rdl(w, (int)k), rk = rdl(rw, (int)k);
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const uint64_t mk = rdl(f.w[j], (int)k);
        if (bit(mk, lane)) {
          dlt[(size_t)wk * B + lane + 64 * j] = (double)rk;
          xr[(size_t)wk * B + lane + 64 * j] = ((double)rk + 1.0) / (double)sig[(size_t)wk * B + lane + 64 * j];
        }
      }
    }
  }
}

// e_r of the roots (degree of the root itself).
__global__ void k_broot_deg(const uint32_t *__restrict__ col, const uint32_t *__restrict__ src, int nb,
                            unsigned long long *cnt) {
  const int s = threadIdx.x + blockIdx.x * blockDim.x;
  if (s < nb) atomicAdd(&cnt[4 * s + 2], (unsigned long long)(col[src[s] + 1] - col[src[s]]));
}

// ---------------------------------------------------------------------------
// Backward depth d
// ---------------------------------------------------------------------------

template <int NW>
__device__ __forceinline__ void bstrip_body(uint64_t *vis, const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw,
                         const uint32_t *__restrict__ qoff, const axbc::BCtl *ctl, int depth) {
  if (depth < ctl->dmin) return;
  const int lev = depth + 1;
  const uint32_t b = qoff[lev], e = qoff[lev + 1];
  for (uint32_t i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    const uint32_t v = q[i];
#pragma unroll
    for (int j = 0; j < NW; ++j) vis[(size_t)v * NW + j] &= ~qw[(size_t)i * NW + j];
  }
}

template <int NW>
__global__ void k_bstrip(uint64_t *vis, const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw,
                         const uint32_t *__restrict__ qoff, axbc::BCtl *ctl) {
  // a depth step is this kernel, k_bdep and the split kernels: the step's
  // depth comes from k_bdep of the step before (sdepth), and is handed to
  // the kernels after this one (depth) with the split claims cleared -- no
  // kernel reads the field that the kernel before it writes
  const int d = ctl->sdepth;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->depth = d;
    ctl->split_claim = 0;
    for (int c = 0; c < kHandSlices; ++c) ctl->hand_dep[c * kHandStride] = ctl->hand_ent[c * kHandStride] = 0;
  }
  bstrip_body<NW>(vis, q, qw, qoff, ctl, d);
}

// Dependencies of one vertex over adj[cb, ce), per root lane, left fold in
// adjacency order into acc.
template <int NW>
__device__ __forceinline__ uint32_t dep_chunk(uint32_t st, uint32_t cb, uint32_t ce, const uint32_t *__restrict__ row,
                                              const uint64_t *__restrict__ vis, const Words<NW> &f,
                                              const double *__restrict__ xr, const double sv[NW], double acc[NW]) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  uint32_t contrib = 0;  // (edge, root) terms, wave-uniform
  for (uint32_t k0 = cb; k0 < ce; k0 += kWave) {
    const uint32_t k = k0 + lane;
    uint32_t w = 0;
    Words<NW> m;
#pragma unroll
    for (int j = 0; j < NW; ++j) m.w[j] = 0;
    if (k < ce) {
      w = row[st + k];
      const Words<NW> vw = ldw<NW>(vis, w);
#pragma unroll
      for (int j = 0; j < NW; ++j) m.w[j] = f.w[j] & ~vw.w[j];
    }
    // kBurst edges at a time: loads first, then the adds in edge order
    unsigned long long todo = __ballot(m.any());
    while (todo) {
      double got[kBurst][NW];
      bool use[kBurst][NW];
#pragma unroll
      for (int t = 0; t < kBurst; ++t) {
        const int i = todo ? __ffsll((long long)todo) - 1 : 0;
        const bool live = todo != 0;
        todo &= todo - 1;
        const uint32_t ww = rdl(w, i);
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          const uint64_t mm = live ? rdl(m.w[j], i) : 0ull;
          contrib += (uint32_t)__popcll(mm);
          use[t][j] = bit(mm, lane);
          got[t][j] = use[t][j] ? xr[(size_t)ww * B + lane + 64 * j] : 0.0;
        }
      }
#pragma unroll
      for (int t = 0; t < kBurst; ++t)
#pragma unroll
        for (int j = 0; j < NW; ++j)
          if (use[t][j]) acc[j] += got[t][j] * sv[j];
    }
  }
  return contrib;
}

// Dependencies of one entry (v, f) of d <= kChunk edges, lanes = roots:
// delta_s(v) = reach + sum x_s(w) sigma_s(v) in adjacency order (entries of
// more than kChunk edges are claimed for k_bdep_split by the caller).
// Returns the algorithmic bytes.
template <int NW>
__device__ __forceinline__ unsigned long long dep_entry(uint32_t i, uint32_t v, const Words<NW> &f, uint32_t st,
                                                        uint32_t d, const uint32_t *__restrict__ row,
                                                        const uint64_t *__restrict__ vis,
                                                        const uint32_t *__restrict__ sig, double *dlt, double *xr,
                                                        const uint32_t *__restrict__ reach, axbc::BCtl *ctl,
                                                        uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap,
                                                        uint64_t part_rows, uint32_t kChunk) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  double sv[NW], acc[NW];
  const double r = (double)reach[v];
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    // only the entry's roots are used: the other lanes skip the load
    sv[j] = bit(f.w[j], lane) ? (double)sig[(size_t)v * B + lane + 64 * j] : 1.0;
    acc[j] = r;
  }
  const uint32_t contrib = dep_chunk<NW>(st, 0, d, row, vis, f, xr, sv, acc);
  uint32_t roots = 0;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    roots += (uint32_t)__popcll(f.w[j]);
    if (bit(f.w[j], lane)) {
      dlt[(size_t)v * B + lane + 64 * j] = acc[j];
      xr[(size_t)v * B + lane + 64 * j] = (acc[j] + 1.0) / sv[j];
    }
  }
  // entry: id 4 + word 8NW + offsets 8 + reach 4; per edge: id 4 + VIS 8NW;
  // per term: x 8; per root of the entry: sigma 4 + delta 8 + x 8
  return 16 + 8 * NW + (unsigned long long)d * (4 + 8 * NW) + 8ull * contrib + 20ull * roots;
}

// Dependencies of the small entries of a group of 64 (lane = entry; small:
// at most kSparseDeg edges), one (entry, root) pair per lane: the lane folds
// its root's terms over the entry's edges in adjacency order -- the same
// operations in the same order as dep_entry, so the same bits.  Returns the
// algorithmic bytes (the same definition as dep_entry's).
template <int NW>
__device__ __forceinline__ unsigned long long dep_pairs(bool small, uint32_t v, const Words<NW> &f, uint32_t st,
                                                        uint32_t d, const uint32_t *__restrict__ row,
                                                        const uint64_t *__restrict__ vis,
                                                        const uint32_t *__restrict__ sig, double *dlt, double *xr,
                                                        const uint32_t *__restrict__ reach) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  uint32_t np = 0;
  if (small)
#pragma unroll
    for (int j = 0; j < NW; ++j) np += (uint32_t)__popcll(f.w[j]);
  unsigned long long bytes = small ? 16 + 8 * NW + (unsigned long long)d * (4 + 8 * NW) + 20ull * np : 0ull;
  const uint32_t incl = wave_incl_scan(np);
  const uint32_t total = rdl(incl, kWave - 1);
  uint32_t terms = 0;
  for (uint32_t p0 = 0; p0 < total; p0 += kWave) {
    const uint32_t p = p0 + (uint32_t)lane;
    const int own = wave_owner(incl, p) & (kWave - 1);
    const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
    const uint32_t ex = own ? pin : 0u;
    const uint32_t ov = __shfl(v, own), ost = __shfl(st, own), od = __shfl(d, own);
    const Words<NW> of = shflw<NW>(f, own);
    if (p < total) {
      const int s = select_root<NW>(of.w, p - ex);
      const int sj = s >> 6;
      const uint64_t sb = 1ull << (s & 63);
      const double sv = (double)sig[(size_t)ov * B + s];
      double acc = (double)reach[ov];
      for (uint32_t e = 0; e < od; e += 4) {
        uint32_t w[4];
        bool use[4];
        double x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = e + t < od ? row[ost + e + t] : 0u;
#pragma unroll
        for (int t = 0; t < 4; ++t) use[t] = e + t < od && !(vis[(size_t)w[t] * NW + sj] & sb);
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = use[t] ? xr[(size_t)w[t] * B + s] : 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (use[t]) {
            acc += x[t] * sv;
            ++terms;
          }
      }
      dlt[(size_t)ov * B + s] = acc;
      xr[(size_t)ov * B + s] = (acc + 1.0) / sv;
    }
  }
  return bytes + 8ull * terms;
}

template <int NW>
__device__ __forceinline__ void bdep_body(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                 const uint64_t *__restrict__ vis, const uint32_t *__restrict__ sig,
                                                 double *dlt, double *xr, const uint32_t *__restrict__ reach,
                                                 const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw,
                                                 const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
                                                 uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap, uint64_t part_rows,
                                                 unsigned long long *kb, uint32_t kChunk,
                                                 const uint8_t *__restrict__ lmode, int depth,
                                                 unsigned long long *hand) {
  if (depth < ctl->dmin) return;
  const uint64_t root_mask0 = ctl->rootmask[0], root_mask1 = ctl->rootmask[1];
  const uint32_t b = qoff[depth], e = qoff[depth + 1];
  unsigned long long bytes = 0;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  if (!lmode[depth]) {
    // dense level: one entry per wave at a time, lanes = roots; the wave's
    // entries are loaded 64 at a time (lane = entry) so that their split
    // entries are claimed together.  Group (r, w) holds the entries
    // b + w + (64 r + lane) nwv, so the groups that run at one time hold
    // neighbouring list entries (listed together, they share neighbours and
    // cache lines).  Without a hand-out, wave w takes the groups (0, w),
    // (1, w), ...; with one (hand), a wave takes the groups r nwv + w one at
    // a time.
    const uint64_t rounds = ((uint64_t)e - b + (uint64_t)nwv * kWave - 1) / ((uint64_t)nwv * kWave);
    // groups holding entries: every group of the rounds before the last, and
    // in the last round those whose first entry lies below e
    const uint64_t last = rounds ? std::min<uint64_t>(nwv, (uint64_t)e - b - (rounds - 1) * nwv * kWave) : 0;
    const uint64_t ngroups = rounds ? (rounds - 1) * nwv + last : 0;
    Hand h = hand_begin(ngroups);
    uint64_t g = gw;
    for (;;) {
      if (hand) {
        if (!hand_next(h, hand, ngroups, g)) break;
      } else if (g >= rounds * nwv) {
        break;
      }
      const uint64_t il = (uint64_t)b + g % nwv + ((g / nwv) * kWave + (uint64_t)lane) * nwv;
      if (!hand) g += nwv;
      const bool in = il < e;
      const uint32_t i = in ? (uint32_t)il : 0u;
      uint32_t v = 0, st = 0, d = 0;
      Words<NW> f;
#pragma unroll
      for (int j = 0; j < NW; ++j) f.w[j] = 0;
      if (in) {
        v = q[i];
        f = ldw<NW>(qw, i);
        if (depth == 0) {
          f.w[0] &= root_mask0;
          if (NW > 1) f.w[NW - 1] &= root_mask1;
        }
        st = col[v];
        d = col[v + 1] - st;
      }
      const bool live = in && f.any();
      const bool spl = live && d > kChunk;
      claim_split_wave(ctl, spl, i, spl ? (d + kChunk - 1) / kChunk : 0u, sp_e, sp_base, sp_cap, part_rows);
      bytes += (unsigned long long)__popcll(__ballot(spl)) * (16 + 8 * NW);
      unsigned long long todo = __ballot(live && !spl);
      while (todo) {
        const int k = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        bytes += dep_entry<NW>(rdl(i, k), rdl(v, k), rdlw<NW>(f, k), rdl(st, k), rdl(d, k), row, vis, sig, dlt, xr,
                               reach, ctl, sp_e, sp_base, sp_cap, part_rows, kChunk);
      }
    }
  } else {
    // sparse level (few roots and few edges per entry): a wave takes 64
    // entries; the small ones as (entry, root) pairs, the rare others one at
    // a time (those above the chunk go to k_bdep_split)
    for (uint32_t base = b + gw * kWave; base < e; base += nwv * kWave) {
      const uint32_t i = base + lane;
      uint32_t v = 0, st = 0, d = 0;
      Words<NW> f;
#pragma unroll
      for (int j = 0; j < NW; ++j) f.w[j] = 0;
      if (i < e) {
        v = q[i];
        f = ldw<NW>(qw, i);
        if (depth == 0) {
          f.w[0] &= root_mask0;
          if (NW > 1) f.w[NW - 1] &= root_mask1;
        }
        st = col[v];
        d = col[v + 1] - st;
      }
      const bool live = i < e && f.any();
      const bool small = live && d <= kSparseDeg;
      const bool spl = live && d > kChunk;
      bytes += dep_pairs<NW>(small, v, f, st, d, row, vis, sig, dlt, xr, reach);
      claim_split_wave(ctl, spl, i, spl ? (d + kChunk - 1) / kChunk : 0u, sp_e, sp_base, sp_cap, part_rows);
      bytes += (unsigned long long)__popcll(__ballot(spl)) * (16 + 8 * NW);
      unsigned long long todo = __ballot(live && !small && !spl);
      while (todo) {
        const int k = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        bytes += dep_entry<NW>(base + (uint32_t)k, rdl(v, k), rdlw<NW>(f, k), rdl(st, k), rdl(d, k), row, vis, sig, dlt,
                               xr, reach, ctl, sp_e, sp_base, sp_cap, part_rows, kChunk);
      }
    }
  }
  kb_add(kb, bytes);
}

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bdep(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                 const uint64_t *__restrict__ vis, const uint32_t *__restrict__ sig,
                                                 double *dlt, double *xr, const uint32_t *__restrict__ reach,
                                                 const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw,
                                                 const uint32_t *__restrict__ qoff, axbc::BCtl *ctl,
                                                 uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap, uint64_t part_rows,
                                                 unsigned long long *kb, uint32_t kChunk,
                                                 const uint8_t *__restrict__ lmode) {
  const int d = ctl->depth;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->sdepth = d >= ctl->dmin ? d - 1 : d;  // the next step's
  bdep_body<NW>(col, row, vis, sig, dlt, xr, reach, q, qw, qoff, ctl, sp_e, sp_base, sp_cap, part_rows, kb, kChunk,
                lmode, d, ctl->hand_ent);
}

// Chunks of the split entries: per item the partial sums of its kChunk
// edges (left fold from 0.0) go to its row of `part`.
template <int NW>
// Items are handed out one at a time (hand_next), not in fixed ranges per
// wave: their costs differ with the roots that have terms there, and with
// fixed ranges the launch waits for its slowest wave.  An item's partial row
// is the same whichever wave computes it.
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bdep_split(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                                       const uint64_t *__restrict__ vis, const uint32_t *__restrict__ sig,
                                                       const double *__restrict__ xr, const uint32_t *__restrict__ q,
                                                       const uint64_t *__restrict__ qw, axbc::BCtl *ctl,
                                                       const uint32_t *__restrict__ sp_e,
                                                       const uint64_t *__restrict__ sp_base, uint32_t sp_cap,
                                                       uint64_t part_rows, double *part, unsigned long long *kb,
                                                       uint32_t kChunk) {
  constexpr int B = 64 * NW;
  const int depth = ctl->depth;
  if (depth < ctl->dmin) return;
  const uint64_t root_mask0 = ctl->rootmask[0], root_mask1 = ctl->rootmask[1];
  const int lane = lane_id();
  const unsigned long long claim = ctl->split_claim;
  const uint32_t ns = min((uint32_t)(claim >> 40), sp_cap);
  const uint64_t total = min((uint64_t)(claim & ((1ull << 40) - 1)), part_rows);
  if (ns == 0 || total == 0) return;
  unsigned long long bytes = 0;
  uint32_t v = 0, st = 0, d = 0;
  uint64_t base = 0, next = 0;  // the current entry's items: [base, next)
  Words<NW> f;
  double sv[NW];
  Hand hand = hand_begin(total);
  uint64_t item = 0;
  while (hand_next(hand, ctl->hand_dep, total, item)) {
    if (item < base || item >= next) {
      const uint32_t i = sp_e[item_entry(sp_e, sp_base, sp_cap, ns, total, item, base, next)];
      v = q[i];
      st = col[v];
      d = col[v + 1] - st;
      f = ldw<NW>(qw, i);
      if (depth == 0) {
        f.w[0] &= root_mask0;
        if (NW > 1) f.w[NW - 1] &= root_mask1;
      }
#pragma unroll
      for (int j = 0; j < NW; ++j) sv[j] = bit(f.w[j], lane) ? (double)sig[(size_t)v * B + lane + 64 * j] : 1.0;
    }
    const uint32_t cb = (uint32_t)(item - base) * kChunk, ce = min(d, cb + kChunk);
    if (cb >= ce) continue;
    double acc[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) acc[j] = 0.0;
    const uint32_t contrib = dep_chunk<NW>(st, cb, ce, row, vis, f, xr, sv, acc);
#pragma unroll
    for (int j = 0; j < NW; ++j) part[(size_t)item * B + lane + 64 * j] = acc[j];
    // per edge: id 4 + VIS 8NW; per term: x 8 (partials are scratch)
    bytes += (unsigned long long)(ce - cb) * (4 + 8 * NW) + 8ull * contrib;
  }
  kb_add(kb, bytes);
}

// One wave per split entry: delta = reach + partials in chunk order.
constexpr int kCombineRun = 8;  // partial rows loaded ahead of their adds

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bdep_combine(const uint32_t *__restrict__ col,
                                                         const uint32_t *__restrict__ sig, double *dlt, double *xr,
                                                         const uint32_t *__restrict__ reach,
                                                         const uint32_t *__restrict__ q,
                                                         const uint64_t *__restrict__ qw, const axbc::BCtl *ctl,
                                                         const uint32_t *__restrict__ sp_e,
                                                         const uint64_t *__restrict__ sp_base, uint32_t sp_cap,
                                                         const double *__restrict__ part, unsigned long long *kb,
                                                         uint32_t kChunk) {
  constexpr int B = 64 * NW;
  const int depth = ctl->depth;
  if (depth < ctl->dmin) return;
  const uint64_t root_mask0 = ctl->rootmask[0], root_mask1 = ctl->rootmask[1];
  const unsigned long long claim = ctl->split_claim;
  const uint32_t ns = min((uint32_t)(claim >> 40), sp_cap);
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  unsigned long long bytes = 0;
  for (uint32_t k = gw; k < ns; k += nwv) {
    const uint32_t i = sp_e[k];
    const uint32_t v = q[i];
    const uint32_t d = col[v + 1] - col[v];
    const uint32_t nch = (d + kChunk - 1) / kChunk;
    const uint64_t base = sp_base[k];
    Words<NW> f = ldw<NW>(qw, i);
    if (depth == 0) {
      f.w[0] &= root_mask0;
      if (NW > 1) f.w[NW - 1] &= root_mask1;
    }
    const double r = (double)reach[v];
    // t = reach + p_0 + p_1 + ... in item order; the loads of 8 items are
    // issued before their adds
    const double *pp = part + (size_t)base * B + lane;
    double t[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) t[j] = r;
    uint32_t cc = 0;
    for (; cc + kCombineRun <= nch; cc += kCombineRun) {
      double p[kCombineRun][NW];
#pragma unroll
      for (int u = 0; u < kCombineRun; ++u)
#pragma unroll
        for (int j = 0; j < NW; ++j) p[u][j] = pp[(size_t)(cc + u) * B + 64 * j];
#pragma unroll
      for (int u = 0; u < kCombineRun; ++u)
#pragma unroll
        for (int j = 0; j < NW; ++j) t[j] += p[u][j];
    }
    for (; cc < nch; ++cc)
#pragma unroll
      for (int j = 0; j < NW; ++j) t[j] += pp[(size_t)cc * B + 64 * j];
    uint32_t roots = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      roots += (uint32_t)__popcll(f.w[j]);
      if (bit(f.w[j], lane)) {
        dlt[(size_t)v * B + lane + 64 * j] = t[j];
        xr[(size_t)v * B + lane + 64 * j] = (t[j] + 1.0) / (double)sig[(size_t)v * B + lane + 64 * j];
      }
    }
    // per root of the entry: sigma 4 + delta 8 + x 8
    bytes += 20ull * roots;
  }
  kb_add(kb, bytes);
}

// Between the push's split items and the path-count pull's claims of a
// top-down level.  On a bottom-up level the claims are the pull's (made by
// k_bexpand before this kernel) and stay for k_bpull_split.
__global__ void k_bsplit_reset(axbc::BCtl *ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl->bu) {
    ctl->split_claim = 0;
  }
}

// Backward sweep set-up: depths dmax-1 .. dmin.
__global__ void k_bset_depth(axbc::BCtl *ctl, int first, int last, uint64_t rootmask0, uint64_t rootmask1) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->depth = first;
    ctl->sdepth = first;
    ctl->dmin = last;
    ctl->rootmask[0] = rootmask0;
    ctl->rootmask[1] = rootmask1;
    ctl->split_claim = 0;
  }
}

// ---------------------------------------------------------------------------
// Persistent level loop (graphs without split vertices, many levels)
// ---------------------------------------------------------------------------

// Current value of a barrier word, read by a compare-and-swap that never
// matches (the words never hold ~0): a read-modify-write, performed at the
// coherence point.  A plain or relaxed atomic load of ordinary (coarse-
// grained) device memory, or an add of 0 (which compiles to a load), can be
// served by a stale line of the XCD's own L2 and never see another XCD's
// update.
__device__ __forceinline__ uint32_t coherent_read(uint32_t *p) { return atomicCAS(p, ~0u, ~0u); }

// Words of the persistent loop's grid barrier, in an allocation of their
// own that only atomics touch (sharing a cache sector with fields written by
// plain stores lets that sector's write-back from one XCD's L2 restore stale
// words).  Arrivals go to kBarGroups counters, one 128-byte line each, so
// that the workgroups' atomics spread over memory channels instead of
// queueing on one address; the last arrival of a group arrives at the top
// counter, and the last top arrival moves to the next generation.
enum { kBarGen = 0, kBarErr = 1, kBarSeen = 2, kBarTop = 32, kBarGroup0 = 64, kBarGroups = 32,
       kBarStride = 32, kBarWords = kBarGroup0 + kBarGroups * kBarStride };

// Generation and error words polled with system-scope atomic loads, which
// are not served from the XCD's own L2 (a relaxed agent-scope load was
// measured to keep returning a stale generation).
__device__ __forceinline__ uint32_t poll_read(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Grid-wide barrier of a cooperative launch (every workgroup resident).
// Each wait is bounded by `limit` ticks of the 100 MHz real-time counter;
// a workgroup that waits longer, or sees that another one did, sets the
// error word and gets false, and the kernel ends in every workgroup (the
// host then reports the error).  prog counts each workgroup's arrivals
// (diagnostics of a failed wait).
__device__ __forceinline__ bool grid_sync(uint32_t *bar, uint32_t nblocks, unsigned long long limit,
                                          uint32_t *prog) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    prog[blockIdx.x] += 1;
    const uint32_t gen = coherent_read(&bar[kBarGen]);
    __threadfence();  // this workgroup's writes before its arrival
    const uint32_t grp = blockIdx.x % kBarGroups;
    const uint32_t in_grp = nblocks / kBarGroups + (grp < nblocks % kBarGroups ? 1u : 0u);
    const uint32_t ngrp = nblocks < (uint32_t)kBarGroups ? nblocks : (uint32_t)kBarGroups;
    bool last = false;
    if (atomicAdd(&bar[kBarGroup0 + grp * kBarStride], 1u) + 1u == in_grp) {
      atomicExch(&bar[kBarGroup0 + grp * kBarStride], 0u);
      if (atomicAdd(&bar[kBarTop], 1u) + 1u == ngrp) {
        atomicExch(&bar[kBarTop], 0u);
        last = true;
      }
    }
    if (last) {
      __threadfence();
      atomicAdd(&bar[kBarGen], 1u);
    } else {
      const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
      while (poll_read(&bar[kBarGen]) == gen) {
        if (poll_read(&bar[kBarErr]) || __builtin_amdgcn_s_memrealtime() - t0 > limit) {
          if (atomicOr(&bar[kBarErr], 1u) == 0u) atomicExch(&bar[kBarSeen], coherent_read(&bar[kBarTop]));
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();  // the other workgroups' writes after the barrier
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Resets the barrier words before a persistent launch.
__global__ void k_bar_reset(uint32_t *bar) {
  const int t = (int)threadIdx.x;
  if (blockIdx.x == 0 && t < kBarGroups) atomicExch(&bar[kBarGroup0 + t * kBarStride], 0u);
  if (blockIdx.x == 0 && t == 0) {
    atomicExch(&bar[kBarTop], 0u);
    atomicExch(&bar[kBarErr], 0u);
    atomicExch(&bar[kBarSeen], 0u);
  }
}

// The forward sweep in one launch: per level the push (or the bottom-up
// pull), the path counts, the close, and the close end in workgroup 0, each
// followed by a grid barrier; until the done flag.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_bfwd_persist(
    uint32_t n, const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
    const uint32_t *__restrict__ rcol, const uint32_t *__restrict__ rrow, const uint32_t *__restrict__ reach,
    uint64_t *vis, uint64_t *nx, uint32_t *mark, uint64_t *fr0, uint64_t *fr1, uint32_t *sig, uint32_t *q,
    uint64_t *qw, uint32_t *qoff, uint64_t *act, axbc::BCtl *ctl, unsigned long long *wslot,
    unsigned long long *lslot, uint32_t *seen, uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap,
    uint64_t item_cap, uint32_t qcap, uint32_t kChunk, int strict, uint8_t *lmode, uint32_t pull16,
    unsigned long long limit, uint32_t *bar, uint32_t *prog) {
  const uint32_t nblocks = gridDim.x;
  for (;;) {
    if (ld_flag(&ctl->done)) return;
    if (ctl->bu)
      bpull_body<NW>(n, rcol, rrow, vis, nx, fr0, fr1, sig, q, qoff, ctl, wslot, sp_e, sp_base, sp_cap, item_cap,
                     qcap, nullptr, kChunk, nullptr);
    else
      bpush_body<NW>(col, row, vis, nx, mark, fr0, fr1, q, qw, qoff, ctl, sp_e, sp_base, sp_cap, item_cap, qcap,
                     nullptr, kChunk);
    if (!grid_sync(bar, nblocks, limit, prog)) return;
    bclose_body<NW>(col, reach, rcol, rrow, nx, vis, fr0, fr1, sig, q, qw, qoff, ctl, wslot, lslot, seen, sp_e, sp_base,
                    sp_cap, item_cap, qcap, nullptr, kChunk, strict, lmode);
    if (!grid_sync(bar, nblocks, limit, prog)) return;
    if (blockIdx.x == 0) bclose_end_body(ctl, qoff, qcap, pull16, lslot, act, NW, lmode, (int)nblocks);
    if (!grid_sync(bar, nblocks, limit, prog)) return;
  }
}

// The backward sweep in one launch: per depth (first down to last) the VIS
// strip and the dependency pass, each followed by a grid barrier.  The depth is the loop's own; no split vertices here.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_bbwd_persist(
    const uint32_t *__restrict__ col, const uint32_t *__restrict__ row, uint64_t *vis,
    const uint32_t *__restrict__ sig, double *dlt, double *xr, const uint32_t *__restrict__ reach,
    const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw, const uint32_t *__restrict__ qoff,
    axbc::BCtl *ctl, uint32_t *sp_e, uint64_t *sp_base, uint32_t sp_cap, uint64_t part_rows, uint32_t kChunk,
    const uint8_t *__restrict__ lmode, int first, int last,
    unsigned long long limit, uint32_t *bar, uint32_t *prog) {
  const uint32_t nblocks = gridDim.x;
  for (int depth = first; depth >= last; --depth) {
    bstrip_body<NW>(vis, q, qw, qoff, ctl, depth);
    if (!grid_sync(bar, nblocks, limit, prog)) return;
    bdep_body<NW>(col, row, vis, sig, dlt, xr, reach, q, qw, qoff, ctl, sp_e, sp_base, sp_cap, part_rows, nullptr,
                  kChunk, lmode, depth, nullptr);
    if (!grid_sync(bar, nblocks, limit, prog)) return;
  }
}

// ---------------------------------------------------------------------------
// Stop test and bc update
// ---------------------------------------------------------------------------

// Term of root s for vertex v (the sequence of one-root updates).
template <int NW>
__device__ __forceinline__ double bc_term(uint32_t v, int s, const uint32_t *__restrict__ src,
                                          const double *__restrict__ prob, const uint32_t *__restrict__ r0,
                                          const double *__restrict__ dlt, const unsigned long long *__restrict__ cnt,
                                          bool *has) {
  constexpr int B = 64 * NW;
  *has = true;
  if (v == src[s]) {
    const uint32_t rr = r0[s];
    if (rr == 0) {
      *has = false;
      return 0.0;
    }
    const uint32_t all = (uint32_t)cnt[4 * s + 1] + rr;
    const unsigned long long nvt = cnt[4 * s + 0] + (unsigned long long)all;
    return (double)((unsigned long long)rr * (nvt - 2ull));
  }
  const double r1 = (double)(uint32_t)(r0[s] + 1u);
  return dlt[(size_t)v * B + s] * r1 / prob[s];
}

template <int NW>
__global__ void k_bstop(const uint32_t *__restrict__ check, uint32_t x, uint32_t n, const double *__restrict__ bc,
                        double cn, int nb, const uint32_t *__restrict__ lane_of, const uint32_t *__restrict__ src,
                        const double *__restrict__ prob, const uint32_t *__restrict__ r0,
                        const double *__restrict__ dlt, const unsigned long long *__restrict__ cnt,
                        unsigned long long *ok_bits) {
  // ok_bits[j] bit s: every check vertex is satisfied after root s (draw
  // order; root s sits in lane lane_of[s])
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < x; i += gridDim.x * blockDim.x) {
    const uint32_t a = check[i];
    double b = a < n ? bc[a] : 0.0;
    uint64_t sat[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) sat[j] = 0;
    for (int s = 0; s < nb; ++s) {
      if (a < n) {
        bool has;
        const double t = bc_term<NW>(a, (int)lane_of[s], src, prob, r0, dlt, cnt, &has);
        if (has) b += t;
      }
      if (!(b / 2.0 < cn)) sat[s / 64] |= 1ull << (s % 64);
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) atomicAnd(&ok_bits[j], (unsigned long long)sat[j]);
  }
}

// bc += the batch's terms in draw order, roots [0, nr).  A workgroup takes
// kUpdVerts vertices at a time: the terms of all lanes are computed with
// threads over (vertex, lane), so every wave reads whole 512-byte segments
// of the delta rows, and stored in LDS at their draw position (the inverse
// of lane_of); then one thread per vertex folds its row left to right, the
// sequence of one-root updates bit for bit, with independent LDS reads the
// compiler can issue ahead of the dependent adds.
constexpr int kUpdVerts = 32;

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_bupdate(uint32_t n, const uint32_t *__restrict__ seen, uint32_t stamp,
                                                    const uint32_t *__restrict__ reach, int nr,
                                                    const uint32_t *__restrict__ lane_of,
                                                    const uint32_t *__restrict__ src, const double *__restrict__ prob,
                                                    const uint32_t *__restrict__ r0, const double *__restrict__ dlt,
                                                    const unsigned long long *__restrict__ cnt, double *bc,
                                                    unsigned long long *kb) {
  constexpr int B = 64 * NW;
  constexpr int kUpdPer = kUpdVerts * B / kBlock;  // (vertex, lane) terms per thread
  static_assert(kUpdVerts * B % kBlock == 0, "tile must split evenly over the workgroup");
  __shared__ double tile[kUpdVerts][B + 1];
  __shared__ double s_r1[B], s_p[B], s_root[B];
  __shared__ uint32_t s_src[B], s_draw[B];
  __shared__ uint8_t s_use[B];
  __shared__ uint8_t s_want[kUpdVerts];
  for (int l = threadIdx.x; l < B; l += blockDim.x) s_use[l] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < B; k += blockDim.x) {
    if (k < nr) {
      const uint32_t l = lane_of[k];
      const uint32_t rr = r0[l];
      s_draw[l] = (uint32_t)k;
      s_use[l] = 1;
      s_src[l] = src[l];
      s_r1[l] = (double)(uint32_t)(rr + 1u);
      s_p[l] = prob[l];
      // the root's own term (update_bc's v == v0 branch); absent when rr == 0
      const uint32_t all = (uint32_t)cnt[4 * l + 1] + rr;
      const unsigned long long nvt = cnt[4 * l + 0] + (unsigned long long)all;
      s_root[l] = rr ? (double)((unsigned long long)rr * (nvt - 2ull)) : 0.0;
    }
  }
  __syncthreads();
  unsigned long long bytes = 0;
  for (uint32_t vb = blockIdx.x * kUpdVerts; vb < n; vb += gridDim.x * kUpdVerts) {
    if (threadIdx.x < kUpdVerts) {
      const uint32_t v = vb + threadIdx.x;
      const bool want = v < n && (seen[v] == stamp || reach[v] != 0);
      s_want[threadIdx.x] = want;
      bytes += v < n ? 8 + (want ? 16 + 8ull * (unsigned)nr : 0) : 0;
    }
    __syncthreads();
    // all of a thread's row loads first, then the terms
    double got[kUpdPer];
#pragma unroll
    for (int r = 0; r < kUpdPer; ++r) {
      const int flat = threadIdx.x + r * kBlock;
      const int vi = flat / B, l = flat % B;
      const uint32_t vv = vb + (uint32_t)vi;
      got[r] = s_want[vi] && s_use[l] && vv != s_src[l] ? dlt[(size_t)vv * B + l] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kUpdPer; ++r) {
      const int flat = threadIdx.x + r * kBlock;
      const int vi = flat / B, l = flat % B;
      const uint32_t vv = vb + (uint32_t)vi;
      if (s_want[vi] && s_use[l]) tile[vi][s_draw[l]] = vv == s_src[l] ? s_root[l] : got[r] * s_r1[l] / s_p[l];
    }
    __syncthreads();
    if (threadIdx.x < kUpdVerts && s_want[threadIdx.x]) {
      const uint32_t v = vb + threadIdx.x;
      double acc = bc[v];
#pragma unroll 8
      for (int k = 0; k < nr; ++k) acc += tile[threadIdx.x][k];
      bc[v] = acc;
    }
    __syncthreads();
  }
  if (kb) kb_add(kb, wave_sum_u64(bytes));
}

// Utility of the batch's roots for a set of target vertices, in a fixed
// order: wave gw sums weight * delta_s(v) over the targets gw, gw + W, ...
// into its own partial row (lane = root, coalesced rows, no atomics), then
// k_butility_sum adds the W partial rows in wave order.  The same count
// gives the same grid, so the result does not depend on the run.

template <int NW>
__global__ __launch_bounds__(kBlock, kWavesPerSimd) void k_butility(const uint32_t *__restrict__ vert, const double *__restrict__ wt,
                                                     uint32_t nv, const double *__restrict__ dlt, double *part) {
  constexpr int B = 64 * NW;
  const int lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  double acc[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) acc[j] = 0.0;
  for (uint32_t i = gw; i < nv; i += nwv) {
    const uint32_t v = vert[i];
    const double w = wt[i];
#pragma unroll
    for (int j = 0; j < NW; ++j) acc[j] += w * dlt[(size_t)v * B + lane + 64 * j];
  }
#pragma unroll
  for (int j = 0; j < NW; ++j) part[(size_t)gw * B + lane + 64 * j] = acc[j];
}

// out[s] = sum of the partial rows in wave order; one thread per root.
__global__ void k_butility_sum(const double *__restrict__ part, uint32_t rows, int B, double *out) {
  const int s = (int)threadIdx.x;
  if (s >= B) return;
  double t = 0.0;
  for (uint32_t r = 0; r < rows; ++r) t += part[(size_t)r * B + s];
  out[s] = t;
}

// One root's levels and path counts from the batch's level list: lvl = -1
// and sigma = 0 first (k_blevels_init), then every list entry of level L
// whose word holds the root's lane sets lvl = L and copies the root's path
// count from the vertex's row.
__global__ void k_blevels_init(uint32_t n, int32_t *lvl, uint32_t *sigma) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    lvl[v] = -1;
    sigma[v] = 0;
  }
}

template <int NW>
__global__ void k_blevels(const uint32_t *__restrict__ q, const uint64_t *__restrict__ qw,
                          const uint32_t *__restrict__ qoff, int levels, int lane,
                          const uint32_t *__restrict__ sig, int32_t *lvl, uint32_t *sigma) {
  constexpr int B = 64 * NW;
  const uint32_t ne = qoff[levels];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += gridDim.x * blockDim.x) {
    if (!((qw[(size_t)i * NW + lane / 64] >> (lane % 64)) & 1ull)) continue;
    // level of entry i: the last L with qoff[L] <= i
    int lo = 0, hi = levels;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (qoff[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t v = q[i];
    lvl[v] = lo;
    sigma[v] = sig[(size_t)v * B + lane];
  }
}

template <int NW>
__global__ void k_bcopy_delta(uint32_t n, int s, const double *__restrict__ dlt, double *out) {
  constexpr int B = 64 * NW;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    out[v] = dlt[(size_t)v * B + s];
}

// Whether every adjacency list is strictly increasing and every edge has its
// reverse (then the transpose equals the CSC); bit 0 of *bad set otherwise.
// A wave per vertex, the reverse found by binary search.
__global__ __launch_bounds__(kBlock) void k_mirror_check(const uint32_t *__restrict__ col,
                                                         const uint32_t *__restrict__ row, uint32_t n,
                                                         unsigned int *bad) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t i = gw; i < n; i += nwv) {
    const uint32_t b = col[i], d = col[i + 1] - b;
    bool ok = true;
    for (uint32_t k = lane_id(); k < d; k += kWave) {
      const uint32_t j = row[b + k];
      ok &= j < n && (k == 0 || row[b + k - 1] < j);
      if (ok) {
        const uint32_t *a = row + col[j];
        uint32_t lo = 0, hi = col[j + 1] - col[j];
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (a[mid] < i)
            lo = mid + 1;
          else
            hi = mid;
        }
        ok = lo < col[j + 1] - col[j] && a[lo] == i;
      }
    }
    if (__ballot(!ok) && lane_id() == 0) atomicOr(bad, 1u);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace axbc {

int mirror_symmetric(axbc_ctx *c, bool *sym) {
  if (c->mirror_state < 0) {
    unsigned int *bad = nullptr;
    HIP_TRY(hipMalloc((void **)&bad, sizeof(unsigned int)));
    HIP_TRY(hipMemsetAsync(bad, 0, sizeof(unsigned int), c->stream));
    hipLaunchKernelGGL(k_mirror_check, dim3(4096), dim3(kBlock), 0, c->stream, c->col, c->row, c->n, bad);
    unsigned int hb = 1;
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream);
    const hipError_t e3 = hipStreamSynchronize(c->stream);
    (void)hipFree(bad);
    HIP_TRY(e1);
    HIP_TRY(e2);
    HIP_TRY(e3);
    c->mirror_state = hb ? 0 : 1;
  }
  *sym = c->mirror_state == 1;
  return AXBC_OK;
}

void batch_free(axbc_ctx *c) {
  BatchState *b = c->batch;
  if (!b) return;
  if (b->gfwd) (void)hipGraphExecDestroy(b->gfwd);
  if (b->gbwd) (void)hipGraphExecDestroy(b->gbwd);
  void *d[] = {b->vis, b->nx, b->fr[0], b->fr[1], b->sig, b->dlt, b->xr, b->mark, b->seen, b->q, b->qw, b->qoff,
               b->act, b->sp_e, b->sp_base, b->part, b->cnt, b->src, b->prob, b->r0,
               b->lane_of, b->check, b->ctl, b->wslot, b->lslot, b->lmode, b->prog, b->bar, b->row_out, b->util_v, b->util_w, b->util_part, b->util_out,
               b->wst, b->wdl, b->wlst, b->wloff, b->wnlev, b->wsrc, b->wcnt};  // wovf lies inside wnlev
  for (void *p : d)
    if (p) (void)hipFree(p);
  if (b->h_wnlev) (void)hipHostFree(b->h_wnlev);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  if (b->h_cnt) (void)hipHostFree(b->h_cnt);
  if (b->h_act) (void)hipHostFree(b->h_act);
  if (b->charged) c->bytes -= std::min(c->bytes, b->bytes);
  delete b;
  c->batch = nullptr;
}

// The batch functions act on c->batch and c->stream; for the second state
// of a batch of more than 128 bit-packed roots this swaps in batch2 and
// stream2 for its lifetime (one host thread drives both states).
struct SecondSlot {
  axbc_ctx *c;
  explicit SecondSlot(axbc_ctx *cc) : c(cc) {
    std::swap(c->batch, c->batch2);
    std::swap(c->stream, c->stream2);
  }
  ~SecondSlot() {
    std::swap(c->batch, c->batch2);
    std::swap(c->stream, c->stream2);
  }
  SecondSlot(const SecondSlot &) = delete;
  SecondSlot &operator=(const SecondSlot &) = delete;
};

void batch_free_all(axbc_ctx *c) {
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  batch_free(c);
  if (c->batch2) {
    SecondSlot two(c);
    batch_free(c);
  }
  for (auto &e : c->pair_ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  c->stream2 = nullptr;
  c->pair_last = false;
}

}  // namespace axbc

namespace {

template <typename T>
int balloc(BatchState *b, T **p, size_t count) {
  const size_t bytes = sizeof(T) * (count ? count : 1);
  if (b->plan_only) {
    b->bytes += bytes;
    *p = nullptr;
    return AXBC_OK;
  }
  if (hipMalloc((void **)p, bytes) != hipSuccess)
    return fail(AXBC_ENOMEM, "batch allocation of %s failed (line %d)", "device memory", __LINE__);
  b->bytes += bytes;
  return AXBC_OK;
}

// An allocation for the batch after its creation (getters, utility):
// checked against the context's budget and counted in its total.
template <typename T>
int balloc_more(axbc_ctx *c, T **p, size_t count, const char *what) {
  int rc;
  if ((rc = budget_check(c, sizeof(T) * (count ? count : 1), what))) return rc;
  if ((rc = balloc(c->batch, p, count))) return rc;
  c->bytes += sizeof(T) * (count ? count : 1);
  return AXBC_OK;
}

// Pull direction: radj(w) = {u : w in adj(u)}.  Symmetric graphs reuse the
// CSC; otherwise the transpose is built on the host once.
int ensure_transpose(axbc_ctx *c) {
  if (c->rcol) return AXBC_OK;
  // strictly increasing lists in which every edge has its reverse: the
  // transpose is the CSC itself, decided on the device without copying it
  {
    bool sym = false;
    int rc;
    if ((rc = mirror_symmetric(c, &sym))) return rc;
    if (sym) {
      uint32_t mx = 0;
      for (uint32_t v = 0; v < c->n; ++v) mx = std::max(mx, c->h_col[v + 1] - c->h_col[v]);
      c->max_out = c->max_in = mx;
      c->symmetric = true;
      c->pull_strict = true;
      c->rcol = c->col;
      c->rrow = c->row;
      return AXBC_OK;
    }
  }
  std::vector<uint32_t> col(c->n + 1), row(c->ned);
  HIP_TRY(hipMemcpyAsync(col.data(), c->col, sizeof(uint32_t) * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
  if (c->ned)
    HIP_TRY(hipMemcpyAsync(row.data(), c->row, sizeof(uint32_t) * c->ned, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<uint32_t> rcol(c->n + 1, 0), rrow(c->ned);
  for (uint64_t k = 0; k < c->ned; ++k) rcol[row[k] + 1]++;
  for (uint32_t v = 0; v < c->n; ++v) rcol[v + 1] += rcol[v];
  std::vector<uint32_t> fill(rcol.begin(), rcol.end() - 1);
  for (uint32_t u = 0; u < c->n; ++u)
    for (uint32_t k = col[u]; k < col[u + 1]; ++k) rrow[fill[row[k]]++] = u;
  // the lists are filled in increasing u: strict unless an edge repeats
  c->pull_strict = true;
  for (uint32_t v = 0; v < c->n && c->pull_strict; ++v)
    for (uint32_t k = rcol[v] + 1; k < rcol[v + 1]; ++k)
      if (rrow[k] <= rrow[k - 1]) {
        c->pull_strict = false;
        break;
      }
  c->max_out = c->max_in = 0;
  for (uint32_t v = 0; v < c->n; ++v) {
    c->max_out = std::max(c->max_out, col[v + 1] - col[v]);
    c->max_in = std::max(c->max_in, rcol[v + 1] - rcol[v]);
  }
  if (rcol == col && rrow == row) {
    c->symmetric = true;
    c->rcol = c->col;
    c->rrow = c->row;
    return AXBC_OK;
  }
  c->symmetric = false;
  {
    int rc;
    if ((rc = budget_check(c, sizeof(uint32_t) * (c->n + 1 + c->ned), "the transposed graph"))) return rc;
  }
  HIP_TRY(hipMalloc((void **)&c->rcol, sizeof(uint32_t) * (c->n + 1)));
  HIP_TRY(hipMalloc((void **)&c->rrow, sizeof(uint32_t) * (c->ned ? c->ned : 1)));
  HIP_TRY(hipMemcpyAsync(c->rcol, rcol.data(), sizeof(uint32_t) * (c->n + 1), hipMemcpyHostToDevice, c->stream));
  if (c->ned)
    HIP_TRY(hipMemcpyAsync(c->rrow, rrow.data(), sizeof(uint32_t) * c->ned, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->bytes += sizeof(uint32_t) * (c->n + 1 + c->ned);
  return AXBC_OK;
}

// Every device buffer of a batch state, sized from b's n, nw, B, qcap,
// sp_cap and part_rows; with b->plan_only set it only adds up the bytes.
int alloc_buffers(BatchState *b) {
  const size_t n = b->n, nw = b->nw, B = b->B, qcap = b->qcap;
  int r;
#define BA(p, cnt) \
  if ((r = balloc(b, &(p), (cnt))) != 0) return r;
  BA(b->vis, n * nw);
  BA(b->nx, n * nw);
  BA(b->fr[0], n * nw);
  BA(b->fr[1], n * nw);
  BA(b->sig, n * B);
  BA(b->dlt, n * B);
  BA(b->xr, n * B);
  if (nw > 1) BA(b->mark, n);
  BA(b->seen, n);
  BA(b->q, qcap);
  BA(b->qw, qcap * nw);
  BA(b->qoff, n + 2);
  BA(b->act, (n + 2) * nw);
  BA(b->sp_e, b->sp_cap + b->part_rows);  // the slots' entries, then the item -> slot map
  BA(b->sp_base, b->sp_cap);
  BA(b->part, b->part_rows * B);
  BA(b->cnt, 4 * B);
  BA(b->wslot, 4ull * kSlotWaves * B);
  BA(b->lslot, (size_t)kLevelSlot * kSlotWaves);
  BA(b->lmode, n + 2);
  BA(b->prog, kBGrid);
  BA(b->bar, kBarWords);
  BA(b->src, B);
  BA(b->prob, B);
  BA(b->r0, B);
  BA(b->lane_of, B);
  BA(b->ctl, 1);
#undef BA
  return AXBC_OK;
}

// Split entries and chunk rows of a graph: vertices with more than
// kChunkMin edges (sized for the smallest chunk, so that the chunk can be
// tuned later), and at most ned / kChunkMin + nsplit + 1 chunk rows.
void size_splits(BatchState *b, const uint32_t *col, size_t n, uint64_t ned) {
  uint32_t nsplit = 0;
  for (size_t v = 0; v < n; ++v) nsplit += (col[v + 1] - col[v]) > kChunkMin;
  b->sp_cap = nsplit + 1;
  b->part_rows = ned / kChunkMin + nsplit + 1;
}

// First capacity of the level list: a vertex is listed once per distinct
// level among the roots, so at most width times; small-world graphs need a
// few times N, high-diameter ones (roots sharing no levels) up to width
// times N.
uint64_t list_bound(uint64_t n, int width) { return std::min<uint64_t>(n * (uint64_t)width + 1024, 0xFFFFFFF0ull); }
uint32_t list_first(uint64_t n, int width) { return (uint32_t)std::min<uint64_t>(n * 8 + 1024, list_bound(n, width)); }

int batch_create(axbc_ctx *c, int width, uint32_t qcap) {
  int rc;
  if ((rc = ensure_transpose(c))) return rc;
  BatchState *b = new BatchState();
  c->batch = b;
  b->nw = width / 64;
  b->B = width;
  b->n = c->n;
  const size_t n = c->n, nw = b->nw, B = b->B;
  size_splits(b, c->h_col.data(), n, c->ned);
  b->chunk = c->batch_chunk;
  b->pull16 = c->batch_pull16;
  b->qcap = qcap;
  // every buffer, first sized (checked against the context's budget), then
  // allocated
  b->plan_only = true;
  (void)alloc_buffers(b);
  const size_t planned = b->bytes;
  b->plan_only = false;
  b->bytes = 0;
  if ((rc = budget_check(c, planned, "the batch state")) || (rc = alloc_buffers(b))) {
    batch_free(c);
    return rc;
  }

  HIP_TRY(hipHostMalloc((void **)&b->h_ctl, sizeof(BCtl), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void **)&b->h_cnt, sizeof(unsigned long long) * 4 * B, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void **)&b->h_act, sizeof(uint64_t) * (n + 2) * nw, hipHostMallocDefault));
  HIP_TRY(hipMemsetAsync(b->ctl, 0, sizeof(BCtl), c->stream));
  HIP_TRY(hipMemsetAsync(b->act, 0, sizeof(uint64_t) * (n + 2) * nw, c->stream));
  if (nw == 1)
    hipLaunchKernelGGL(k_bfill<1>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (uint32_t)n, c->reach, b->vis,
                       b->nx, b->fr[0], b->fr[1], b->sig, b->dlt, b->mark, b->seen);
  else
    hipLaunchKernelGGL(k_bfill<2>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, (uint32_t)n, c->reach, b->vis,
                       b->nx, b->fr[0], b->fr[1], b->sig, b->dlt, b->mark, b->seen);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->bytes += b->bytes;
  b->charged = true;
  return AXBC_OK;
}

int pull_bar(axbc_ctx *c) {
  HIP_TRY(hipMemcpyAsync(c->batch->h_bar, c->batch->bar, sizeof(c->batch->h_bar), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int pull_bctl(axbc_ctx *c) {
  HIP_TRY(hipMemcpyAsync(c->batch->h_ctl, c->batch->ctl, sizeof(BCtl), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

// Where a level's launches go: the context stream, or kernel nodes appended
// in order to a graph under construction.  The graphs are built node by node
// rather than by stream capture: capture is process-wide state in HIP, and
// other threads' launches and copies fail while one is under way (several
// contexts run concurrently when serving).
struct Sink {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
};

template <typename Tuple, size_t... I>
void param_ptrs(Tuple &t, void **out, std::index_sequence<I...>) {
  ((out[I] = (void *)&std::get<I>(t)), ...);
}

// Launch f<<<g, b>>>(a...) on the sink: on its stream, or as the next node
// of its graph (the arguments are copied when the node is created).
template <typename... P, typename... A>
void launch(Sink &k, void (*f)(P...), dim3 g, dim3 b, A... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  std::tuple<P...> t(a...);
  void *args[sizeof...(P) ? sizeof...(P) : 1];
  param_ptrs(t, args, std::index_sequence_for<P...>{});
  hipError_t e;
  if (!k.graph) {
    e = hipLaunchKernel((const void *)f, g, b, args, 0, k.stream);
  } else {
    hipKernelNodeParams np{};
    np.func = (void *)f;
    np.gridDim = g;
    np.blockDim = b;
    np.sharedMemBytes = 0;
    np.kernelParams = args;
    np.extra = nullptr;
    hipGraphNode_t node = nullptr;
    e = hipGraphAddKernelNode(&node, k.graph, k.last ? &k.last : nullptr, k.last ? 1 : 0, &np);
    if (e == hipSuccess) k.last = node;
  }
  if (e != hipSuccess && k.err == hipSuccess) k.err = e;
}

// Cooperative launch of f<<<g, b>>>(a...) on stream s (all workgroups
// resident at once; the runtime refuses a grid that cannot be).
template <typename... P, typename... A>
hipError_t coop_launch(void (*f)(P...), dim3 g, dim3 b, hipStream_t s, A... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  std::tuple<P...> t(a...);
  void *args[sizeof...(P) ? sizeof...(P) : 1];
  param_ptrs(t, args, std::index_sequence_for<P...>{});
  return hipLaunchCooperativeKernel((const void *)f, g, b, args, 0, s);
}

// Grid of f for a persistent launch: one workgroup per CU fewer than the
// occupancy query reports, capped at kBGrid.  (At the reported 7 workgroups
// per CU the last workgroup of every CU of the backward sweep was not
// resident until the others left the kernel: measured, every barrier then
// waits for 256 workgroups that cannot start.)
template <typename F>
int coop_grid(axbc_ctx *c, F f) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kBlock, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) return 0;
  return std::min(kBGrid, std::min(std::max(1, per_cu - 1), kPersistPerCu) * cus);
}

// One forward level: the same launch sequence for every level (the kernels
// read the level from the device), so it can be built into a graph once and
// replayed.
template <int NW>
void enqueue_level(Sink &k, axbc_ctx *c, BatchState *b, bool split_out, bool split_in) {
  const uint32_t n = c->n;
  const bool merged = !c->ktimer.on && b->pull16;
  if (merged) {
    launch(k, k_bexpand<NW>, dim3(kBGrid), dim3(kBlock), n, c->col, c->row, c->rcol, c->rrow, b->vis, b->nx, b->mark,
           b->fr[0], b->fr[1], b->sig, b->q, b->qw, b->qoff, b->ctl, b->wslot, b->sp_e, b->sp_base, b->sp_cap,
           b->part_rows, b->qcap, b->chunk);
  } else {
    kt_begin(c, KK_PUSH);
    launch(k, k_bpush<NW>, dim3(kBGrid), dim3(kBlock), c->col, c->row, b->vis, b->nx, b->mark, b->fr[0],
                       b->fr[1], b->q, b->qw, b->qoff, b->ctl, b->sp_e, b->sp_base, b->sp_cap, b->part_rows, b->qcap,
                       kt_bytes(c, KK_PUSH), b->chunk);
    kt_end(c, KK_PUSH);
  }
  if (split_out) {
    kt_begin(c, KK_PUSH_SPLIT);
    launch(k, k_bpush_split<NW>, dim3(kBGrid), dim3(kBlock), c->col, c->row, b->vis, b->nx, b->mark,
                       b->q, b->qw, b->ctl, b->sp_e, b->sp_base, b->sp_cap, b->part_rows, b->qcap,
                       kt_bytes(c, KK_PUSH_SPLIT), b->chunk);
    kt_end(c, KK_PUSH_SPLIT);
    launch(k, k_bsplit_reset, dim3(1), dim3(64), b->ctl);
  }
  if (b->pull16) {
    if (!merged) {
      kt_begin(c, KK_PULL);
      launch(k, k_bpull<NW>, dim3(kBGrid), dim3(kBlock), n, c->rcol, c->rrow, b->vis, b->nx, b->fr[0],
                         b->fr[1], b->sig, b->q, b->qoff, b->ctl, b->wslot, b->sp_e, b->sp_base, b->sp_cap, b->part_rows,
                         b->qcap, kt_bytes(c, KK_PULL), b->chunk);
      kt_end(c, KK_PULL);
    }
    if (split_in) {
      kt_begin(c, KK_PULL_SPLIT);
      launch(k, k_bpull_split<NW>, dim3(kBGrid), dim3(kBlock), c->rcol, c->rrow, b->vis, b->nx,
                         b->fr[0], b->fr[1], b->sig, b->mark, b->q, b->ctl, b->wslot, b->sp_e, b->sp_base, b->sp_cap,
                         b->part_rows, b->qcap, kt_bytes(c, KK_PULL_SPLIT), b->chunk);
      kt_end(c, KK_PULL_SPLIT);
    }
  }
  // close and path counts (k_bclose_end adds up its per-workgroup level slots)
  kt_begin(c, KK_CLOSE);
  launch(k, k_bclose<NW>, dim3(kBGrid), dim3(kBlock), c->col, c->reach, c->rcol, c->rrow, b->nx, b->vis, b->fr[0],
         b->fr[1], b->sig, b->q, b->qw, b->qoff, b->ctl, b->wslot, b->lslot, b->seen, b->sp_e, b->sp_base,
         b->sp_cap, b->part_rows, b->qcap, kt_bytes(c, KK_CLOSE), b->chunk, c->pull_strict ? 1 : 0,
         (const uint8_t *)b->lmode);
  kt_end(c, KK_CLOSE);
  if (split_in) {
    kt_begin(c, KK_SIGMA_SPLIT);
    launch(k, k_bsigma_split<NW>, dim3(kBGrid), dim3(kBlock), c->rcol, c->rrow, b->qw, b->fr[0],
                       b->fr[1], b->sig, b->q, b->ctl, b->wslot, b->sp_e, b->sp_base, b->sp_cap, b->part_rows,
                       kt_bytes(c, KK_SIGMA_SPLIT), b->chunk);
    kt_end(c, KK_SIGMA_SPLIT);
  }
  launch(k, k_bclose_end, dim3(1), dim3(kBlock), b->ctl, b->qoff, b->qcap, b->pull16, (const unsigned long long *)b->lslot,
         b->act, NW, b->lmode);
}

// One backward depth step (the depth comes from the device: k_bstrip takes
// it from k_bdep of the step before).  Before depth d the level d+1 bits
// leave VIS.
template <int NW>
void enqueue_depth(Sink &k, axbc_ctx *c, BatchState *b, bool split_out) {
  launch(k, k_bstrip<NW>, dim3(kBGrid / 2), dim3(kBlock), b->vis, b->q, b->qw, b->qoff, b->ctl);
  kt_begin(c, KK_DEP);
  launch(k, k_bdep<NW>, dim3(kBGrid), dim3(kBlock), c->col, c->row, b->vis, b->sig, b->dlt, b->xr,
                     c->reach, b->q, b->qw, b->qoff, b->ctl, b->sp_e, b->sp_base, b->sp_cap, b->part_rows,
                     kt_bytes(c, KK_DEP), b->chunk, (const uint8_t *)b->lmode);
  kt_end(c, KK_DEP);
  if (split_out) {
    kt_begin(c, KK_DEP_SPLIT);
    launch(k, k_bdep_split<NW>, dim3(kBGrid), dim3(kBlock), c->col, c->row, b->vis, b->sig, b->xr,
                       b->q, b->qw, b->ctl, b->sp_e, b->sp_base, b->sp_cap, b->part_rows, b->part,
                       kt_bytes(c, KK_DEP_SPLIT), b->chunk);
    kt_end(c, KK_DEP_SPLIT);
    kt_begin(c, KK_DEP_COMBINE);
    launch(k, k_bdep_combine<NW>, dim3(kBGrid / 4), dim3(kBlock), c->col, b->sig, b->dlt, b->xr,
                       c->reach, b->q, b->qw, b->ctl, b->sp_e, b->sp_base, b->sp_cap, b->part,
                       kt_bytes(c, KK_DEP_COMBINE), b->chunk);
    kt_end(c, KK_DEP_COMBINE);
  }
}

void graphs_free(BatchState *b) {
  if (b->gfwd) (void)hipGraphExecDestroy(b->gfwd);
  if (b->gbwd) (void)hipGraphExecDestroy(b->gbwd);
  b->gfwd = b->gbwd = nullptr;
}

// kGraphSteps levels (depths) per graph, rebuilt when the
// launch sequence changes (split kernels, pull ratio, chunk).
template <int NW>
int ensure_graphs(axbc_ctx *c, bool split_out, bool split_in) {
  BatchState *b = c->batch;
  const uint64_t key = ((uint64_t)b->chunk << 32) | ((uint64_t)b->pull16 << 2) | ((uint64_t)split_out << 1) |
                       (uint64_t)split_in;
  if (b->gfwd && b->gbwd && b->gkey == key) return AXBC_OK;
  graphs_free(b);
  for (int pass = 0; pass < 2; ++pass) {
    hipGraph_t g = nullptr;
    HIP_TRY(hipGraphCreate(&g, 0));
    Sink k;
    k.graph = g;
    for (int step = 0; step < kGraphSteps; ++step) {
      if (pass == 0)
        enqueue_level<NW>(k, c, b, split_out, split_in);
      else
        enqueue_depth<NW>(k, c, b, split_out);
    }
    if (k.err != hipSuccess) {
      (void)hipGraphDestroy(g);
      return fail(AXBC_EHIP, "graph construction failed: %s (line %d)", hipGetErrorString(k.err), __LINE__);
    }
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return fail(AXBC_EHIP, "graph instantiation failed: %s (line %d)", hipGetErrorString(ie), __LINE__);
    (pass == 0 ? b->gfwd : b->gbwd) = ge;
  }
  b->gkey = key;
  return AXBC_OK;
}

// A barrier of the persistent loop waited past its limit: the sweep stopped
// part way, so the batch state is dropped (rebuilt by the next batch) and bc
// is left as it was (the update had not run).
int persist_failed(axbc_ctx *c) {
  {
    // how many workgroups reached how many barriers
    BatchState *b = c->batch;
    std::vector<uint32_t> pr(kBGrid);
    if (hipMemcpy(pr.data(), b->prog, sizeof(uint32_t) * kBGrid, hipMemcpyDeviceToHost) == hipSuccess) {
      std::vector<std::pair<uint32_t, int>> hist;
      for (int g = 0; g < kBGrid; ++g) {
        bool found = false;
        for (auto &h : hist)
          if (h.first == pr[g]) ++h.second, found = true;
        if (!found) hist.push_back({pr[g], 1});
      }
      fprintf(stderr, "persistent loop: barriers reached per workgroup (barriers: workgroups):");
      for (auto &h : hist) fprintf(stderr, " %u: %d", h.first, h.second);
      fprintf(stderr, "\n");
    }
  }
  char what[192];
  snprintf(what, sizeof what, " (grid %d/%d workgroups, %u groups arrived, generation %u)", c->batch->coop_fwd,
           c->batch->coop_bwd, c->batch->h_bar[kBarSeen], c->batch->h_bar[kBarGen]);
  batch_free(c);
  return fail(AXBC_EHIP, "persistent level loop: a barrier wait passed its limit%s (line %d)", what, __LINE__);
}

// Automatic choice of the workgroup sweep: graphs whose previous batch ran
// more than kWgLevels levels (high diameter: the roots' frontiers rarely
// meet, so the bit-packed sweep shares little and pays a grid-wide launch
// per level), and whose lists are short (a thread walks a vertex's list).
constexpr int kWgLevels = 256;
constexpr uint32_t kWgMaxDeg = 64;
constexpr uint32_t kWgFirstLevels = 16384;  // first level-offset capacity per root

bool use_wg(const axbc_ctx *c) {
  if (c->batch_mode == 2) return true;
  if (c->batch_mode == 1) return false;
  return c->batch->lev_guess > kWgLevels && c->max_out <= kWgMaxDeg;
}

// The workgroup sweep's buffers, all or none: released (and their bytes
// taken off the totals) when one of them cannot be had.
void wg_release(axbc_ctx *c) {
  BatchState *b = c->batch;
  const size_t W = (size_t)b->wlanes, n = c->n;
  size_t held = 0;
  if (b->wst) held += W * n * 8;
  if (b->wdl) held += W * n * 8;
  if (b->wlst) held += W * n * 4;
  if (b->wloff) held += W * b->wlcap * 4;
  if (b->wnlev) held += (W + 1) * 4;
  if (b->wsrc) held += W * 4;
  if (b->wcnt) held += W * 4 * 8;
  void *d[] = {b->wst, b->wdl, b->wlst, b->wloff, b->wnlev, b->wsrc, b->wcnt};
  for (void *p : d)
    if (p) (void)hipFree(p);
  if (b->h_wnlev) (void)hipHostFree(b->h_wnlev);
  b->wst = nullptr;
  b->wdl = nullptr;
  b->wlst = b->wloff = b->wovf = b->wsrc = nullptr;
  b->wnlev = b->h_wnlev = nullptr;
  b->wcnt = nullptr;
  b->bytes -= std::min(b->bytes, held);
  c->bytes -= std::min(c->bytes, held);
  b->wlcap = 0;
  b->wlanes = 0;
}

// Sweep lanes: B, or 2B when the context's width is twice the state's
// (width 256 with the workgroup sweep: roots on all CUs).
int wg_alloc(axbc_ctx *c) {
  BatchState *b = c->batch;
  const size_t W = (size_t)(c->batch_width > b->B ? 2 * b->B : b->B), n = c->n;
  const uint32_t lcap = (uint32_t)std::min<uint64_t>((uint64_t)n + 2, kWgFirstLevels);
  int rc;
  if ((rc = budget_check(c, W * n * (8 + 8 + 4) + W * lcap * 4 + (W + 1) * 4 + W * 4 + W * 32,
                         "the workgroup sweep state")))
    return rc;
  b->wlcap = lcap;
  b->wlanes = (int)W;
  const size_t before = b->bytes;
  if ((rc = balloc(b, &b->wst, W * n)) || (rc = balloc(b, &b->wdl, W * n)) || (rc = balloc(b, &b->wlst, W * n)) ||
      (rc = balloc(b, &b->wloff, W * lcap)) || (rc = balloc(b, &b->wnlev, W + 1)) || (rc = balloc(b, &b->wsrc, W)) ||
      (rc = balloc(b, &b->wcnt, 4 * W))) {
    c->bytes += b->bytes - before;
    wg_release(c);
    return rc;
  }
  c->bytes += b->bytes - before;
  b->wovf = (uint32_t *)(b->wnlev + W);
  if (hipHostMalloc((void **)&b->h_wnlev, sizeof(int32_t) * (W + 1), hipHostMallocDefault) != hipSuccess ||
      hipMemsetAsync(b->wst, 0, sizeof(uint64_t) * W * n, c->stream) != hipSuccess) {
    wg_release(c);
    return fail(AXBC_EHIP, "workgroup sweep state set-up failed%s (line %d)", "", __LINE__);
  }
  b->wtag = 1;
  return AXBC_OK;
}

// The per-root rounds of nr roots (src: device array by sweep lane) into the
// workgroup sweep state; per-lane counters to cnt, deepest levels to
// nlev[0 .. nr).  Returns the tag of level 0.  A root deeper than the
// level-offset capacity makes the sweep run again with four times the
// capacity (nothing outside the sweep state has changed).
int wg_run(axbc_ctx *c, int nr, const uint32_t *src, unsigned long long *cnt, std::vector<int> &nlev,
           uint32_t *tag_out) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const size_t W = (size_t)b->wlanes;
  int rc;
  for (;;) {
    // a batch uses the tags wtag .. wtag + its deepest level: start over
    // (every word unvisited) before they could wrap
    if ((uint64_t)b->wtag + n + 2 > 0xFFFFFFFFull) {
      HIP_TRY(hipMemsetAsync(b->wst, 0, sizeof(uint64_t) * W * n, s));
      b->wtag = 1;
    }
    const uint32_t tag0 = b->wtag;
    HIP_TRY(hipMemsetAsync(b->wovf, 0, sizeof(uint32_t), s));
    kt_begin(c, KK_WGSWEEP);
    HIP_TRY(wg_sweep(s, nr, n, c->col, c->row, c->nr ? c->reach : nullptr, src, tag0, b->wst, b->wdl, b->wlst,
                     b->wloff, b->wlcap, cnt, b->wnlev, b->wovf, kt_bytes(c, KK_WGSWEEP), c->ned));
    kt_end(c, KK_WGSWEEP);
    HIP_TRY(hipMemcpyAsync(b->h_wnlev, b->wnlev, sizeof(int32_t) * (W + 1), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int maxl = 0;
    for (int l = 0; l < nr; ++l) maxl = std::max(maxl, (int)b->h_wnlev[l]);
    b->wtag += (uint32_t)maxl + 2u;
    if (b->h_wnlev[W] == 0) {
      nlev.assign(b->h_wnlev, b->h_wnlev + nr);
      b->wtag_last = tag0;
      b->levels = maxl + 1;
      b->lev_guess = maxl + 2;
      *tag_out = tag0;
      return AXBC_OK;
    }
    if (b->wlcap >= n + 2)
      return fail(AXBC_ESTATE, "workgroup sweep: more levels than vertices%s (line %d)", "", __LINE__);
    const size_t old = sizeof(uint32_t) * W * b->wlcap;
    (void)hipFree(b->wloff);
    b->wloff = nullptr;
    b->bytes -= std::min(b->bytes, old);
    c->bytes -= std::min(c->bytes, old);
    b->wlcap = (uint32_t)std::min<uint64_t>((uint64_t)n + 2, (uint64_t)b->wlcap * 4);
    if ((rc = balloc_more(c, &b->wloff, W * b->wlcap, "the workgroup sweep level offsets"))) {
      wg_release(c);  // the next batch starts over
      return rc;
    }
  }
}

// Workgroup sweep of a batch of at most B roots whose roots and lanes are on
// the device: the per-root rounds, then the batch's delta rows and seen
// stamps, which the stop test and the bc update read as after a bit-packed
// sweep.
template <int NW>
int wg_batch(axbc_ctx *c, int nb, std::vector<int> &nlev) {
  BatchState *b = c->batch;
  int rc;
  if (!b->wst && (rc = wg_alloc(c))) return rc;
  b->have = true;
  b->stamp += 1;
  uint32_t tag0 = 0;
  if ((rc = wg_run(c, nb, b->src, b->cnt, nlev, &tag0))) return rc;
  HIP_TRY(wg_out(c->stream, NW, c->n, nb, 0, tag0, b->wst, b->wdl, c->reach, b->dlt, b->seen, b->stamp));
  b->h_wlane.assign(b->h_lane_of.begin(), b->h_lane_of.end());
  b->last_wg = true;
  b->wg_pair = false;
  return AXBC_OK;
}

// The bit-packed sweep of a batch whose roots, probabilities and lanes are
// on the device: forward levels, per-root counters, the delta rows and the
// backward depths.  nlev[l] receives lane l's deepest level.
// A bit-packed sweep in three phases, so that two states can be driven side
// by side from one host thread: sweep_begin enqueues the set-up and the
// first chunk of forward levels, sweep_forward waits for the chunk and
// enqueues the next until the forward sweep is done, sweep_backward takes
// the per-root depths and enqueues the backward sweep.
struct SweepRun {
  int nb = 0;
  bool split_out = false, split_in = false, graphs = false, persist = false, done = false;
  int L = 1, chunk = 0, zeroed = 0;
};

// One chunk of forward levels from level r.L on.
template <int NW>
int forward_chunk(axbc_ctx *c, SweepRun &r) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const int steps = r.graphs ? (r.chunk + kGraphSteps - 1) / kGraphSteps * kGraphSteps : r.chunk;
  // per-level root words of the levels about to run (a BFS has at most n
  // levels: the buffer holds n + 2 rows)
  if (r.L + steps + 2 > r.zeroed && r.zeroed < (int)n + 2) {
    const int upto = std::min(r.L + steps + 2, (int)n + 2);
    HIP_TRY(hipMemsetAsync(b->act + (size_t)r.zeroed * NW, 0, sizeof(uint64_t) * (size_t)(upto - r.zeroed) * NW, s));
    r.zeroed = upto;
  }
  if (r.graphs) {
    for (int j = 0; j < steps; j += kGraphSteps) HIP_TRY(hipGraphLaunch(b->gfwd, s));
  } else {
    Sink k;
    k.stream = s;
    for (int j = 0; j < steps; ++j) enqueue_level<NW>(k, c, b, r.split_out, r.split_in);
    HIP_TRY(k.err);
  }
  r.L += steps;
  return AXBC_OK;
}

template <int NW>
int sweep_begin(axbc_ctx *c, SweepRun &r, bool may_persist) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const int nb = r.nb;
  int rc;
  // rows of vertices above the split threshold start at zero: clear them
  // all when the threshold dropped below the one they were cleared for
  if (b->chunk < b->sig_zero_above)
    hipLaunchKernelGGL(k_bsig_clear<NW>, dim3(kBGrid), dim3(kBlock), 0, s, n, c->rcol, b->chunk, b->sig);
  b->sig_zero_above = b->chunk;
  r.done = false;
  b->had_prev = b->have;
  b->have = true;
  b->stamp += 1;
  // roots of this batch, per word (the "reached by every root" test)
  const uint64_t all0 = nb >= 64 ? ~0ull : (1ull << nb) - 1ull;
  const uint64_t all1 = NW == 1 || nb <= 64 ? 0ull : (nb >= 128 ? ~0ull : (1ull << (nb - 64)) - 1ull);
  r.zeroed = std::min(b->lev_guess + 2, (int)n + 2);
  HIP_TRY(hipMemsetAsync(b->act, 0, sizeof(uint64_t) * (size_t)r.zeroed * NW, s));
  // per-wave counter slots start at zero
  hipLaunchKernelGGL(k_bzero, dim3(1024), dim3(kBlock), 0, s, b->wslot, (size_t)4 * kSlotWaves * NW * 64);
  hipLaunchKernelGGL(k_broots<NW>, dim3((nb + 63) / 64), dim3(64), 0, s, b->src, nb, b->vis, b->fr[0], b->sig, b->q,
                     b->qw, b->qoff, b->act, b->ctl, b->cnt, b->stamp, b->seen, (unsigned long long)c->ned, b->epoch,
                     all0, all1, c->reach, b->dlt, b->lmode);
  hipLaunchKernelGGL(k_broot_deg, dim3((nb + 63) / 64), dim3(64), 0, s, c->col, b->src, nb, b->cnt);
  HIP_TRY(hipGetLastError());
  r.L = 1;
  r.chunk = b->lev_guess;
  // deep graphs without split vertices: the whole sweep in one persistent
  // cooperative launch per direction (no launch per level)
  bool persist = may_persist && c->batch_persist != 0 && !r.split_out && !r.split_in && !c->ktimer.on &&
                 (c->batch_persist == 2 || b->lev_guess > kPersistLevels);
  if (persist && !b->coop_fwd) {
    b->coop_fwd = coop_grid(c, k_bfwd_persist<NW>);
    b->coop_bwd = coop_grid(c, k_bbwd_persist<NW>);
  }
  r.persist = persist && b->coop_fwd > 0 && b->coop_bwd > 0;
  if (r.persist) {
    const unsigned long long limit = c->batch_persist_limit;
    HIP_TRY(hipMemsetAsync(b->prog, 0, sizeof(uint32_t) * kBGrid, s));
    hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(64), 0, s, b->bar);
    HIP_TRY(coop_launch(k_bfwd_persist<NW>, dim3(b->coop_fwd), dim3(kBlock), s, n, c->col, c->row, c->rcol, c->rrow,
                        c->reach, b->vis, b->nx, b->mark, b->fr[0], b->fr[1], b->sig, b->q, b->qw, b->qoff, b->act,
                        b->ctl, b->wslot, b->lslot, b->seen, b->sp_e, b->sp_base, b->sp_cap, b->part_rows, b->qcap,
                        b->chunk, c->pull_strict ? 1 : 0, b->lmode, b->pull16, limit, b->bar, b->prog));
    if ((rc = pull_bctl(c)) || (rc = pull_bar(c))) return rc;
    if (b->h_bar[kBarErr]) return persist_failed(c);
    if (b->h_ctl->overflow) return fail(AXBC_ENOMEM, "batch list capacity exceeded%s (line %d)", "", __LINE__);
    r.L = (int)b->h_ctl->empty_lev + 1;  // levels run (the epoch of the level stamps moves past them)
    r.done = true;
    return AXBC_OK;
  }
  // forward levels, speculatively in chunks; kernels past the last level
  // exit on the device done flag
  return forward_chunk<NW>(c, r);
}

template <int NW>
int sweep_forward(axbc_ctx *c, SweepRun &r) {
  BatchState *b = c->batch;
  int rc;
  if ((rc = pull_bctl(c))) return rc;
  if (b->h_ctl->overflow) return fail(AXBC_ENOMEM, "batch list capacity exceeded%s (line %d)", "", __LINE__);
  if (b->h_ctl->done) {
    r.done = true;
    return AXBC_OK;
  }
  if ((uint64_t)r.L > (uint64_t)c->n + 1)
    return fail(AXBC_ESTATE, "batch forward did not terminate%s (line %d)", "", __LINE__);
  // past the previous batch's depth: grow by an eighth of the levels run
  // (levels launched past the last one still cost a few microseconds per
  // kernel, so doubling overshot deep graphs by thousands of levels)
  r.chunk = std::max(16, r.L / 8);
  return forward_chunk<NW>(c, r);
}

template <int NW>
int sweep_backward(axbc_ctx *c, SweepRun &r, std::vector<int> &nlev) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const int nb = r.nb;
  int rc;
  // per-root counters from the wave slots of the forward sweep
  hipLaunchKernelGGL(k_bcnt_reduce<NW>, dim3(64, 4), dim3(64 * NW), 0, s, (const unsigned long long *)b->wslot, b->cnt);
  const int empty = (int)b->h_ctl->empty_lev;
  const int dmax = empty - 1;  // deepest non-empty level over the batch
  b->levels = empty;
  b->lev_guess = dmax + 2;
  b->epoch += (uint32_t)(r.L + 1);
  // per-root deepest level from the active words: roots whose deepest
  // level is 1 also accumulate their own dependency (bc_func_mono's do-while)
  HIP_TRY(hipMemcpyAsync(b->h_act, b->act, sizeof(uint64_t) * (size_t)(empty + 1) * NW, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  uint64_t rootmask[2] = {0, 0};
  for (int l = 0; l < nb; ++l) {
    for (int lv = 1; lv <= dmax; ++lv)
      if ((b->h_act[(size_t)lv * NW + l / 64] >> (l % 64)) & 1ull) nlev[l] = lv;
    if (nlev[l] == 1) rootmask[l / 64] |= 1ull << (l % 64);
  }
  // delta = reach where the batch did not reach (VIS is still whole)
  hipLaunchKernelGGL(k_bfix<NW>, dim3(kBGrid), dim3(kBlock), 0, s, n, b->seen, b->stamp, b->stamp - 1,
                     b->had_prev ? 1 : 0, b->vis, c->reach, b->dlt);
  // delta and x of the deepest level's entries (the dependency pass sets
  // the others)
  if (dmax >= 1)
    hipLaunchKernelGGL(k_bleaf<NW>, dim3(kBGrid / 4), dim3(kBlock), 0, s, b->q, b->qw, b->qoff, dmax, c->reach, b->sig,
                       b->xr, b->dlt);
  // backward: depths dmax-1 .. dmin, dmin = 0 only for the roots with a
  // single level
  const int dmin = (rootmask[0] | rootmask[1]) ? 0 : 1;
  const int dsteps = dmax - dmin;
  if (dsteps > 0) {
    hipLaunchKernelGGL(k_bset_depth, dim3(1), dim3(64), 0, s, b->ctl, dmax - 1, dmin, rootmask[0],
                       NW > 1 ? rootmask[1] : 0ull);
    if (r.persist) {
      HIP_TRY(hipMemsetAsync(b->prog, 0, sizeof(uint32_t) * kBGrid, s));
      hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(64), 0, s, b->bar);
      HIP_TRY(coop_launch(k_bbwd_persist<NW>, dim3(b->coop_bwd), dim3(kBlock), s, c->col, c->row, b->vis, b->sig,
                          b->dlt, b->xr, c->reach, b->q, b->qw, b->qoff, b->ctl, b->sp_e, b->sp_base, b->sp_cap,
                          b->part_rows, b->chunk, (const uint8_t *)b->lmode, dmax - 1, dmin, c->batch_persist_limit,
                          b->bar, b->prog));
      if ((rc = pull_bar(c))) return rc;
      if (b->h_bar[kBarErr]) return persist_failed(c);
    } else if (r.graphs) {
      for (int j = 0; j < dsteps; j += kGraphSteps) HIP_TRY(hipGraphLaunch(b->gbwd, s));
    } else {
      Sink k;
      k.stream = s;
      for (int j = 0; j < dsteps; ++j) enqueue_depth<NW>(k, c, b, r.split_out);
      HIP_TRY(k.err);
    }
  }
  HIP_TRY(hipGetLastError());
  b->last_wg = false;
  b->wg_pair = false;
  return AXBC_OK;
}

template <int NW>
int packed_sweep(axbc_ctx *c, SweepRun &r, std::vector<int> &nlev) {
  int rc;
  if ((rc = sweep_begin<NW>(c, r, true))) return rc;
  while (!r.done)
    if ((rc = sweep_forward<NW>(c, r))) return rc;
  return sweep_backward<NW>(c, r, nlev);
}

// Lanes of a group of at most B roots: by degree, highest first (ties in
// draw order).  Roots of similar degree have similar level structure, so the
// roots that meet at an edge of the sweep sit in fewer cache lines of the
// root-minor rows.  lane_of[k] = lane of the k-th root (B entries, identity
// past nb); src/p/r0 by lane.
struct Lanes {
  std::vector<uint32_t> lane_of, src, r0;
  std::vector<double> p;
};

Lanes assign_lanes(const axbc_ctx *c, int B, const uint32_t *v0s, const double *ps, int nb) {
  Lanes ln;
  std::vector<uint32_t> root_at(nb);
  for (int k = 0; k < nb; ++k) root_at[k] = (uint32_t)k;
  std::stable_sort(root_at.begin(), root_at.end(), [&](uint32_t a, uint32_t b2) {
    return c->h_col[v0s[a] + 1] - c->h_col[v0s[a]] > c->h_col[v0s[b2] + 1] - c->h_col[v0s[b2]];
  });
  ln.lane_of.resize(B);
  for (int l = 0; l < B; ++l) ln.lane_of[l] = (uint32_t)l;  // lanes past nb are unused
  ln.src.resize(nb);
  ln.r0.resize(nb);
  ln.p.resize(nb);
  for (int l = 0; l < nb; ++l) {
    const uint32_t k = root_at[l];
    ln.lane_of[k] = (uint32_t)l;
    ln.src[l] = v0s[k];
    ln.r0[l] = c->h_reach[v0s[k]];
    ln.p[l] = ps[k];
  }
  return ln;
}

int upload_lanes(axbc_ctx *c, const Lanes &ln, int nb) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(b->src, ln.src.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->prob, ln.p.data(), sizeof(double) * nb, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->r0, ln.r0.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->lane_of, ln.lane_of.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
  return AXBC_OK;
}

// The per-root stats of the last sweep in c->batch (nb entries, by draw
// order; nlev by lane) from its counters.
int fill_stats(axbc_ctx *c, int nb, const Lanes &ln, const std::vector<int> &nlev, axbc_pivot_stats *stats) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(b->h_cnt, b->cnt, sizeof(unsigned long long) * 4 * nb, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int k = 0; k < nb; ++k) {
    const uint32_t l = ln.lane_of[k];
    stats[k].nvisited = b->h_cnt[4 * l + 0];
    stats[k].nlevels = (uint64_t)nlev[l];
    stats[k].e_r = b->h_cnt[4 * l + 2];
    stats[k].e_dag = b->h_cnt[4 * l + 3];
    stats[k].all = (uint32_t)((uint32_t)b->h_cnt[4 * l + 1] * (c->heuristic == 1 || c->heuristic == 3) + ln.r0[l]);
  }
  return AXBC_OK;
}

// After a sweep of nb roots whose lanes, counters (b->cnt) and delta rows
// (b->dlt) are in place: the exact stop test per root, the bc update of the
// applied roots, the context's delta row, and the stats (nb entries, by
// draw order; nlev by lane).  *applied = roots applied.
template <int NW>
int finish_batch(axbc_ctx *c, int nb, uint32_t stop_c, const Lanes &ln, const std::vector<int> &nlev, int *applied,
                 axbc_pivot_stats *stats) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  int nr = nb;
  if (stop_c > 0 && c->ncheck > 0) {
    unsigned long long *okb = (unsigned long long *)b->part;  // scratch, rows are free now
    std::vector<unsigned long long> ones(NW, ~0ull);
    HIP_TRY(hipMemcpyAsync(okb, ones.data(), sizeof(unsigned long long) * NW, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bstop<NW>, dim3(grid_for(c->ncheck, 64)), dim3(kBlock), 0, s, c->check, c->ncheck, n, c->bc,
                       (double)((uint64_t)stop_c * (uint64_t)n), nb, b->lane_of, b->src, b->prob, b->r0, b->dlt,
                       b->cnt, okb);
    unsigned long long okh[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(okh, okb, sizeof(unsigned long long) * NW, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < nb; ++k)
      if ((okh[k / 64] >> (k % 64)) & 1ull) {
        nr = k + 1;
        break;
      }
  }
  kt_begin(c, KK_UPDATE);
  hipLaunchKernelGGL(k_bupdate<NW>, dim3(kBGrid), dim3(kBlock), 0, s, n, b->seen, b->stamp, c->reach, nr,
                     b->lane_of, b->src, b->prob, b->r0, b->dlt, b->cnt, c->bc, kt_bytes(c, KK_UPDATE));
  kt_end(c, KK_UPDATE);
  HIP_TRY(hipGetLastError());
  // the context's delta becomes the last applied root's (the -z 4
  // distribution of the next draw); plain rounds re-initialise it
  if (nr > 0) {
    hipLaunchKernelGGL(k_bcopy_delta<NW>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, (int)ln.lane_of[nr - 1], b->dlt,
                       c->delta);
    c->dirty_delta = true;
  }
  if (applied) *applied = nr;
  return stats ? fill_stats(c, nb, ln, nlev, stats) : AXBC_OK;
}

// Lanes, graphs and the reset of the state's previous bit-packed sweep,
// before a sweep of nb roots in c->batch on c->stream.
template <int NW>
int batch_prep(axbc_ctx *c, const uint32_t *v0s, const double *ps, int nb, bool wg, Lanes &ln, SweepRun &r) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  int rc;
  ln = assign_lanes(c, b->B, v0s, ps, nb);
  b->h_lane_of = ln.lane_of;
  r.nb = nb;
  // split kernels only where some vertex has more edges than a chunk
  r.split_out = c->max_out > b->chunk;
  r.split_in = c->max_in > b->chunk;
  // the level loop runs from prebuilt graphs unless per-kernel timing is on
  r.graphs = !wg && c->use_graphs && !c->ktimer.on;
  if (r.graphs && (rc = ensure_graphs<NW>(c, r.split_out, r.split_in))) return rc;
  if ((rc = upload_lanes(c, ln, nb))) return rc;
  // undo the previous bit-packed sweep (a workgroup sweep leaves the
  // bit-packed state as it found it)
  if (b->have && !b->last_wg)
    hipLaunchKernelGGL(k_breset<NW>, dim3(kBGrid), dim3(kBlock), 0, s, c->n, b->seen, b->stamp, c->rcol, b->chunk,
                       b->vis, b->fr[0], b->fr[1], b->sig);
  return AXBC_OK;
}

template <int NW>
int run_batch(axbc_ctx *c, const uint32_t *v0s, const double *ps, int nb, uint32_t stop_c, int *applied,
              axbc_pivot_stats *stats) {
  int rc;
  const bool wg = use_wg(c);
  Lanes ln;
  SweepRun r;
  if ((rc = batch_prep<NW>(c, v0s, ps, nb, wg, ln, r))) return rc;
  std::vector<int> nlev(nb, 0);  // per lane
  if (wg)
    rc = wg_batch<NW>(c, nb, nlev);
  else
    rc = packed_sweep<NW>(c, r, nlev);
  if (rc) return rc;
  return finish_batch<NW>(c, nb, stop_c, ln, nlev, applied, stats);
}

// A batch of nb > B roots through the bit-packed sweep: roots 0 .. B-1 in
// the first state on the context stream, the others in the second state on
// the second stream, the two sweeps enqueued phase by phase so that they
// run side by side (each state is a complete sweep of its own roots, so
// its rows are those of a batch of those roots alone).  Then the stop test
// and the bc update in draw order: the first half, and the second only
// when the stop test held for none of the first, after the first's update.
// *applied and the stats count both halves.
template <int NW>
int run_packed_pair(axbc_ctx *c, const uint32_t *v0s, const double *ps, int nb, uint32_t stop_c, int *applied,
                    axbc_pivot_stats *stats) {
  int rc;
  const int B = c->batch->B;
  const int nh[2] = {B, nb - B};
  c->pair_last = false;
  if (!c->stream2) HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  for (auto &e : c->pair_ev)
    if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!c->batch2) {
    SecondSlot two(c);
    if ((rc = batch_create(c, B, c->batch2->qcap))) return rc;
  }
  {
    SecondSlot two(c);
    // the second state follows the first's settings
    c->batch->chunk = c->batch2->chunk;
    c->batch->pull16 = c->batch2->pull16;
  }
  // the second stream starts after the work enqueued on the context stream
  HIP_TRY(hipEventRecord(c->pair_ev[0], c->stream));
  HIP_TRY(hipStreamWaitEvent(c->stream2, c->pair_ev[0], 0));
  Lanes ln[2];
  SweepRun r[2];
  std::vector<int> nlev[2] = {std::vector<int>(nh[0], 0), std::vector<int>(nh[1], 0)};
  if ((rc = batch_prep<NW>(c, v0s, ps, nh[0], false, ln[0], r[0]))) return rc;
  {
    SecondSlot two(c);
    if ((rc = batch_prep<NW>(c, v0s + B, ps + B, nh[1], false, ln[1], r[1]))) return rc;
  }
  // no persistent loop here: it waits on the host for its whole sweep
  if ((rc = sweep_begin<NW>(c, r[0], false))) return rc;
  {
    SecondSlot two(c);
    if ((rc = sweep_begin<NW>(c, r[1], false))) return rc;
  }
  while (!r[0].done || !r[1].done) {
    if (!r[0].done && (rc = sweep_forward<NW>(c, r[0]))) return rc;
    if (!r[1].done) {
      SecondSlot two(c);
      if ((rc = sweep_forward<NW>(c, r[1]))) return rc;
    }
  }
  if ((rc = sweep_backward<NW>(c, r[0], nlev[0]))) return rc;
  {
    SecondSlot two(c);
    if ((rc = sweep_backward<NW>(c, r[1], nlev[1]))) return rc;
  }
  int ap[2] = {0, 0};
  if ((rc = finish_batch<NW>(c, nh[0], stop_c, ln[0], nlev[0], &ap[0], stats))) return rc;
  // the second half's stop test reads bc after the first half's update
  HIP_TRY(hipEventRecord(c->pair_ev[0], c->stream));
  {
    SecondSlot two(c);
    HIP_TRY(hipStreamWaitEvent(c->stream, c->pair_ev[0], 0));
    if (ap[0] == nh[0]) {
      if ((rc = finish_batch<NW>(c, nh[1], stop_c, ln[1], nlev[1], &ap[1], stats ? stats + B : nullptr))) return rc;
    } else if (stats) {
      // swept but not applied: the counters all the same
      if ((rc = fill_stats(c, nh[1], ln[1], nlev[1], stats + B))) return rc;
    }
    HIP_TRY(hipEventRecord(c->pair_ev[1], c->stream));
  }
  // later work on the context stream follows the second half
  HIP_TRY(hipStreamWaitEvent(c->stream, c->pair_ev[1], 0));
  c->pair_last = true;
  if (applied) *applied = ap[0] + ap[1];
  return AXBC_OK;
}

// A batch of B < nb <= 2B roots (width 2B, workgroup sweep only): one sweep
// of all of them (a workgroup each), then the stop test and the bc update
// for the first B roots in draw order and, unless the stop test held among
// them, for the rest -- the sequence of one-root rounds, as for any batch.
template <int NW>
int run_wg_pair(axbc_ctx *c, const uint32_t *v0s, const double *ps, int nb, uint32_t stop_c, int *applied,
                axbc_pivot_stats *stats) {
  BatchState *b = c->batch;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const int B = b->B;
  int rc;
  if (!b->wst && (rc = wg_alloc(c))) return rc;
  if (b->wlanes < nb) return fail(AXBC_ESTATE, "workgroup sweep state has too few lanes%s (line %d)", "", __LINE__);
  const int nh[2] = {B, nb - B};
  const Lanes ln[2] = {assign_lanes(c, B, v0s, ps, nh[0]), assign_lanes(c, B, v0s + B, ps + B, nh[1])};
  // sweep lane of half h's lane l: h B + l
  std::vector<uint32_t> src(ln[0].src);
  src.insert(src.end(), ln[1].src.begin(), ln[1].src.end());
  HIP_TRY(hipMemcpyAsync(b->wsrc, src.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
  if (b->have && !b->last_wg)
    hipLaunchKernelGGL(k_breset<NW>, dim3(kBGrid), dim3(kBlock), 0, s, n, b->seen, b->stamp, c->rcol, b->chunk,
                       b->vis, b->fr[0], b->fr[1], b->sig);
  b->have = true;
  b->stamp += 1;
  std::vector<int> nlev_all;
  uint32_t tag0 = 0;
  if ((rc = wg_run(c, nb, b->wsrc, b->wcnt, nlev_all, &tag0))) return rc;
  b->last_wg = true;
  b->wg_pair = true;
  b->h_wlane.resize(nb);
  for (int k = 0; k < nb; ++k) b->h_wlane[k] = (uint32_t)((k / B) * B) + ln[k / B].lane_of[k % B];
  int done = 0;
  for (int h = 0; h < 2; ++h) {
    b->pair_nb[h] = nh[h];
    b->h_lane_half[h] = ln[h].lane_of;
    b->h_lane_of = ln[h].lane_of;
    if (h == 1 && done < B) continue;  // the stop test held in the first half
    if ((rc = upload_lanes(c, ln[h], nh[h]))) return rc;
    HIP_TRY(hipMemcpyAsync(b->cnt, b->wcnt + (size_t)4 * B * h, sizeof(unsigned long long) * 4 * nh[h],
                           hipMemcpyDeviceToDevice, s));
    HIP_TRY(wg_out(s, NW, n, nh[h], h * B, tag0, b->wst, b->wdl, c->reach, b->dlt, b->seen, b->stamp));
    const std::vector<int> nlev(nlev_all.begin() + h * B, nlev_all.begin() + h * B + nh[h]);
    int ap = 0;
    if ((rc = finish_batch<NW>(c, nh[h], stop_c, ln[h], nlev, &ap, stats ? stats + h * B : nullptr))) return rc;
    done += ap;
  }
  if (stats && done < B) {
    // the second half ran but was not applied: its counters all the same
    std::vector<unsigned long long> h(4 * (size_t)nh[1]);
    HIP_TRY(hipMemcpyAsync(h.data(), b->wcnt + (size_t)4 * B, sizeof(unsigned long long) * h.size(),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < nh[1]; ++k) {
      const uint32_t l = ln[1].lane_of[k];
      axbc_pivot_stats &st = stats[B + k];
      st.nvisited = h[4 * l + 0];
      st.nlevels = (uint64_t)nlev_all[B + l];
      st.e_r = h[4 * l + 2];
      st.e_dag = h[4 * l + 3];
      st.all = (uint32_t)((uint32_t)h[4 * l + 1] * (c->heuristic == 1 || c->heuristic == 3) + ln[1].r0[l]);
    }
  }
  if (applied) *applied = done;
  return AXBC_OK;
}

}  // namespace

namespace axbc {
// Utility on the given device buffers: stage the targets, partial rows per
// wave, their sum in wave order, and the per-root values in draw order.
int batch_utility_on(axbc_ctx *c, uint32_t *dv, double *dw, double *part, double *dout, const uint32_t *vertex,
                     const double *weight, uint32_t count, double *out) {
  BatchState *b = c->batch;
  const int B = b->B;
  hipStream_t s = c->stream;
  uint32_t rows = 0;
  if (count) {
    HIP_TRY(hipMemcpyAsync(dv, vertex, sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dw, weight, sizeof(double) * count, hipMemcpyHostToDevice, s));
    const int grid = (int)std::min<uint32_t>(kUtilWaves / 4, (count + 3) / 4);
    rows = (uint32_t)grid * 4;
    if (b->nw == 1)
      hipLaunchKernelGGL(k_butility<1>, dim3(grid), dim3(kBlock), 0, s, dv, dw, count, b->dlt, part);
    else
      hipLaunchKernelGGL(k_butility<2>, dim3(grid), dim3(kBlock), 0, s, dv, dw, count, b->dlt, part);
  }
  // every out entry is written (zero rows give zero), so nothing is cleared
  hipLaunchKernelGGL(k_butility_sum, dim3(1), dim3(128), 0, s, part, rows, B, dout);
  HIP_TRY(hipGetLastError());
  double per_lane[128];
  HIP_TRY(hipMemcpyAsync(per_lane, dout, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int k = 0; k < B; ++k) out[k] = per_lane[b->h_lane_of.empty() ? k : b->h_lane_of[k]];
  return AXBC_OK;
}

// The utility of the last batch's roots, out[k] for k < the context's
// width: after a workgroup sweep of more than B roots, half by half, each
// half's delta rows written again from the sweep state first (the rows and
// lane map left in place are the last half's, as after the batch).
int batch_utility_any(axbc_ctx *c, uint32_t *dv, double *dw, double *part, double *dout, const uint32_t *vertex,
                      const double *weight, uint32_t count, double *out) {
  BatchState *b = c->batch;
  int rc;
  if (c->pair_last) {
    // each state's roots from its own rows (every call ends synchronised)
    const int B = b->B;
    if ((rc = batch_utility_on(c, dv, dw, part, dout, vertex, weight, count, out))) return rc;
    SecondSlot two(c);
    return batch_utility_on(c, dv, dw, part, dout, vertex, weight, count, out + B);
  }
  if (!b->wg_pair) return batch_utility_on(c, dv, dw, part, dout, vertex, weight, count, out);
  for (int h = 0; h < 2; ++h) {
    HIP_TRY(wg_out(c->stream, b->nw, c->n, b->pair_nb[h], h * b->B, b->wtag_last, b->wst, b->wdl, c->reach, b->dlt,
                   b->seen, b->stamp));
    b->h_lane_of = b->h_lane_half[h];
    if ((rc = batch_utility_on(c, dv, dw, part, dout, vertex, weight, count, out + (size_t)h * b->B))) return rc;
  }
  return AXBC_OK;
}
}  // namespace axbc

namespace axbc {
// Root k's dependency row of the last batch gathered into the batch's row
// buffer on the device (stream-ordered); *row points at it.
int batch_delta_row(axbc_ctx *c, int k, const double **row) {
  if (!c || !c->batch || !row || k < 0 || k >= c->batch_width)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  BatchState *b = c->batch;
  int rc;
  if (c->pair_last && k >= b->B) {
    // a root of the second state: its row on the second stream, complete
    // before the caller reads it on the context stream
    const int B = b->B;
    SecondSlot two(c);
    if ((rc = batch_delta_row(c, k - B, row))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AXBC_OK;
  }
  HIP_TRY(hipSetDevice(c->device));
  if (!b->row_out) {
    if ((rc = balloc_more(c, (double **)&b->row_out, c->n, "a per-root row copy"))) return rc;
  }
  double *tmp = (double *)b->row_out;
  if (b->last_wg && k < (int)b->h_wlane.size()) {
    // workgroup sweep: from the root's own state (reach where it did not reach)
    const size_t wl = b->h_wlane[k];
    HIP_TRY(wg_delta(c->stream, c->n, b->wtag_last, b->wst + wl * c->n, b->wdl + wl * c->n, c->reach, tmp));
    *row = tmp;
    return AXBC_OK;
  }
  if (k >= b->B) return fail(AXBC_EINVAL, "no such root in the last batch%s (line %d)", "", __LINE__);
  const int lane = b->h_lane_of.empty() ? k : (int)b->h_lane_of[k];
  if (b->nw == 1)
    hipLaunchKernelGGL(k_bcopy_delta<1>, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, lane, b->dlt, tmp);
  else
    hipLaunchKernelGGL(k_bcopy_delta<2>, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, lane, b->dlt, tmp);
  HIP_TRY(hipGetLastError());
  *row = tmp;
  return AXBC_OK;
}
}  // namespace axbc

extern "C" {

int axbc_set_batch_width(axbc_ctx *c, int width) {
  if (!c || (width != 64 && width != 128 && width != 256))
    return fail(AXBC_EINVAL, "batch width must be 64, 128 or 256%s (line %d)", "", __LINE__);
  // the state's lanes (and the workgroup sweep's) follow the width
  if (c->batch && c->batch_width != width) batch_free_all(c);
  c->batch_width = width;
  return AXBC_OK;
}

int axbc_batch_width(axbc_ctx *c) { return c ? c->batch_width : 0; }

int axbc_set_batch_chunk(axbc_ctx *c, uint32_t chunk) {
  if (!c || chunk < kChunkMin || chunk > (1u << 20))
    return fail(AXBC_EINVAL, "chunk must be in [64, 2^20]%s (line %d)", "", __LINE__);
  c->batch_chunk = chunk;
  if (c->batch) c->batch->chunk = chunk;
  return AXBC_OK;
}

int axbc_set_batch_graphs(axbc_ctx *c, int enable) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  c->use_graphs = enable != 0;
  return AXBC_OK;
}

int axbc_set_batch_persistent(axbc_ctx *c, int mode, double limit_seconds) {
  if (!c || mode < 0 || mode > 2 || !(limit_seconds > 0.0) || limit_seconds > 3600.0)
    return fail(AXBC_EINVAL, "persistent mode 0..2 and a limit in (0, 3600] s%s (line %d)", "", __LINE__);
  c->batch_persist = mode;
  c->batch_persist_limit = (unsigned long long)(limit_seconds * 1e8);  // 100 MHz real-time ticks
  if (c->batch_persist_limit == 0) c->batch_persist_limit = 1;
  return AXBC_OK;
}

int axbc_set_batch_mode(axbc_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 2) return fail(AXBC_EINVAL, "batch mode must be 0, 1 or 2%s (line %d)", "", __LINE__);
  c->batch_mode = mode;
  return AXBC_OK;
}

int axbc_batch_state_bytes(uint64_t n, uint64_t ned, const uint32_t *col, int width, size_t *bytes) {
  if (!col || !bytes || (width != 64 && width != 128 && width != 256) || n == 0 || n > 0xFFFFFFF0ull)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  // width 256: two states of 128 lanes (a batch of more than 128 roots)
  const int lanes = std::min(width, 128);
  BatchState b;
  b.nw = lanes / 64;
  b.B = lanes;
  b.n = (uint32_t)n;
  b.qcap = list_first(n, lanes);
  size_splits(&b, col, n, ned);
  b.plan_only = true;
  (void)alloc_buffers(&b);
  *bytes = b.bytes * (width > 128 ? 2 : 1);
  return AXBC_OK;
}

int axbc_set_plain_pull(axbc_ctx *c, uint32_t ratio16) {
  if (!c || ratio16 > (1u << 16)) return fail(AXBC_EINVAL, "pull ratio must be in [0, 65536]%s (line %d)", "", __LINE__);
  c->plain_pull16 = ratio16;
  return AXBC_OK;
}

int axbc_set_batch_pull(axbc_ctx *c, uint32_t ratio16) {
  if (!c || ratio16 > (1u << 16)) return fail(AXBC_EINVAL, "pull ratio must be in [0, 65536]%s (line %d)", "", __LINE__);
  c->batch_pull16 = ratio16;
  if (c->batch) c->batch->pull16 = ratio16;
  return AXBC_OK;
}

int axbc_batch(axbc_ctx *c, const uint32_t *v0s, const double *ps, int nb, uint32_t stop_c, int *applied,
               axbc_pivot_stats *stats) {
  if (!c || !v0s || !ps) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  if (nb <= 0 || nb > c->batch_width)
    return fail(AXBC_EINVAL, "batch size must be 1..width%s (line %d)", "", __LINE__);
  for (int k = 0; k < nb; ++k)
    if (v0s[k] >= c->n) return fail(AXBC_EINVAL, "root out of range%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(c->device));
  int rc;
  // a state holds at most 128 lanes; width 256 sweeps the roots of a batch
  // above 128 in one workgroup sweep, or in a second bit-packed state beside
  // the first, and applies them in two halves
  const int sw = std::min(c->batch_width, 128);
  const uint64_t bound = list_bound(c->n, sw);
  if (!c->batch) {
    if ((rc = batch_create(c, sw, list_first(c->n, sw)))) return rc;
  }
  const bool pair = nb > c->batch->B;
  c->pair_last = false;
  if (pair && use_wg(c)) return run_wg_pair<2>(c, v0s, ps, nb, stop_c, applied, stats);
  for (int attempt = 0; attempt < 8; ++attempt) {
    if (pair)
      rc = run_packed_pair<2>(c, v0s, ps, nb, stop_c, applied, stats);
    else
      rc = c->batch->nw == 1 ? run_batch<1>(c, v0s, ps, nb, stop_c, applied, stats)
                             : run_batch<2>(c, v0s, ps, nb, stop_c, applied, stats);
    const bool over = (c->batch && c->batch->h_ctl->overflow) || (c->batch2 && c->batch2->h_ctl->overflow);
    if (rc != AXBC_ENOMEM || !c->batch || !over) return rc;
    // grow the level list and retry the batch from scratch (bc untouched:
    // the update runs only after a complete sweep)
    const uint64_t cap = std::min<uint64_t>((uint64_t)c->batch->qcap * 4, bound);
    const int w = c->batch->B;
    batch_free_all(c);
    if ((rc = batch_create(c, w, (uint32_t)cap))) return rc;
  }
  return fail(AXBC_ENOMEM, "batch list capacity%s (line %d)", "", __LINE__);
}

int axbc_batch_utility(axbc_ctx *c, const uint32_t *vertex, const double *weight, uint32_t count, double *out) {
  if (!c || !c->batch || !vertex || !weight || !out) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  for (uint32_t i = 0; i < count; ++i)
    if (vertex[i] >= c->n) return fail(AXBC_EINVAL, "vertex out of range%s (line %d)", "", __LINE__);
  BatchState *b = c->batch;
  HIP_TRY(hipSetDevice(c->device));
  // staging and partial rows kept in the batch state, grown as needed
  if (!b->util_part) {
    int rc;
    if ((rc = balloc_more(c, &b->util_part, (size_t)kUtilWaves * b->B, "utility rows")) ||
        (rc = balloc_more(c, &b->util_out, (size_t)b->B, "utility rows")))
      return rc;
  }
  if (count > b->util_cap) {
    const size_t old = (size_t)b->util_cap * (sizeof(uint32_t) + sizeof(double));
    if (b->util_v) (void)hipFree(b->util_v);
    if (b->util_w) (void)hipFree(b->util_w);
    b->util_v = nullptr;
    b->util_w = nullptr;
    b->util_cap = 0;
    b->bytes -= std::min(b->bytes, old);
    c->bytes -= std::min(c->bytes, old);
    const uint32_t cap = std::max<uint32_t>(count, 4096);
    int rc;
    if ((rc = balloc_more(c, &b->util_v, cap, "utility targets")) || (rc = balloc_more(c, &b->util_w, cap, "utility targets")))
      return rc;
    b->util_cap = cap;
  }
  return batch_utility_any(c, b->util_v, b->util_w, b->util_part, b->util_out, vertex, weight, count, out);
}

int axbc_batch_delta(axbc_ctx *c, int k, double *delta) {
  if (!c || !c->batch || !delta || k < 0 || k >= c->batch_width)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  const double *row = nullptr;
  int rc;
  if ((rc = batch_delta_row(c, k, &row))) return rc;
  HIP_TRY(hipMemcpyAsync(delta, row, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_batch_levels(axbc_ctx *c, int k, int32_t *lvl, uint32_t *sigma) {
  if (!c || !c->batch || !lvl || !sigma || k < 0 || k >= c->batch_width)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  if (c->pair_last && k >= c->batch->B) {
    // a root of the second state (this call ends synchronised)
    const int B = c->batch->B;
    SecondSlot two(c);
    return axbc_batch_levels(c, k - B, lvl, sigma);
  }
  BatchState *b = c->batch;
  if (!b->have || b->levels <= 0) return fail(AXBC_ESTATE, "no batch has run%s (line %d)", "", __LINE__);
  if (!(b->last_wg && k < (int)b->h_wlane.size()) && k >= b->B)
    return fail(AXBC_EINVAL, "no such root in the last batch%s (line %d)", "", __LINE__);
  const int lane = b->last_wg && k < (int)b->h_wlane.size() ? (int)b->h_wlane[k]
                   : b->h_lane_of.empty()                    ? k
                                                             : (int)b->h_lane_of[k];
  int rc;
  HIP_TRY(hipSetDevice(c->device));
  if (!b->row_out) {
    if ((rc = balloc_more(c, (double **)&b->row_out, c->n, "a per-root row copy"))) return rc;
  }
  int32_t *dl = (int32_t *)b->row_out;
  uint32_t *ds = (uint32_t *)(dl + c->n);
  hipStream_t s = c->stream;
  if (b->last_wg) {
    // workgroup sweep: the lane's own state words
    HIP_TRY(wg_levels(s, c->n, b->wtag_last, b->wst + (size_t)lane * c->n, dl, ds));
  } else {
    hipLaunchKernelGGL(k_blevels_init, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->n, dl, ds);
    if (b->nw == 1)
      hipLaunchKernelGGL(k_blevels<1>, dim3(2048), dim3(kBlock), 0, s, b->q, b->qw, b->qoff, b->levels, lane, b->sig,
                         dl, ds);
    else
      hipLaunchKernelGGL(k_blevels<2>, dim3(2048), dim3(kBlock), 0, s, b->q, b->qw, b->qoff, b->levels, lane, b->sig,
                         dl, ds);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(lvl, dl, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(sigma, ds, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return AXBC_OK;
}

}  // extern "C"
