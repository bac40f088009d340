// Sampled-Brandes engine for MI355X (gfx950): device state, the plain
// one-root-per-round kernels and the C ABI of include/axbc.h.
//
// Data layout in HBM (one context = one resident graph):
//   col[N+1], row[ned]      uint32 CSC, adjacency of u = row[col[u]..col[u+1])
//   reach[N]                uint32 pendant counts (-H 1), rlist = {v: reach>0}
//   lvl[N] int32, sigma[N] uint32, delta[N] f64, bc[N] f64
//   queue[N]                uint32, the frontiers of the round back to back;
//                           qoff[L] = start of level L (qoff[0]=0, qoff[1]=1)
//   hubs[], hub_base[]      deferred high-degree vertices of one pass and the
//                           first of their chunk work items; hub_part[] holds
//                           one partial sum per item (backward)
//   ctl                     device counters (queue tail, done flag, ...)
//
// One round = reset of the previous round's visited vertices, root init,
// per level a load-balanced push expansion (wave = 64 frontier vertices,
// their edges dealt out 64 per step) plus a pass over the deferred hubs in
// chunks of kHubChunk edges (each wave one contiguous range of chunks), then
// per depth an atomic-free FP64 pull over the level's vertices (segmented
// wave reduction) plus the hub chunks, whose partial sums are combined in
// chunk order, then the bc update
// over the visited vertices and the reach>0 list.  Levels are enqueued
// speculatively; kernels past the last level exit on the device done flag,
// so the host waits once per round.
//
// Reference semantics restated: src/cuda_kernels.cu:588-658 (forward),
// 719-772 (backward), 269-302 (bc update), 339-373 (reach sum),
// 1521-1554 (reset), 862-876 (root), src/axbc2d.c:1514-1659 (driver).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../../include/axbc.h"
#include "engine_internal.h"
#include "wave.h"

using namespace axbc;

namespace axbc {
thread_local char g_err[512] = "";

int fail(int code, const char *fmt, const char *a, int line) {
  snprintf(g_err, sizeof(g_err), fmt, a, line);
  return code;
}

int budget_check(axbc_ctx *c, size_t more, const char *what) {
  if (!c->budget || c->bytes + more <= c->budget) return AXBC_OK;
  snprintf(g_err, sizeof(g_err), "device memory budget: %s needs %zu bytes more, %zu of %zu in use", what, more,
           c->bytes, c->budget);
  return AXBC_EBUDGET;
}

static hipEvent_t kt_event(axbc_ctx *c) {
  KernelTimer &t = c->ktimer;
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

static bool kt_timed(const axbc_ctx *c, int kind) { return c->ktimer.on && ((c->ktimer.mask >> kind) & 1u); }

void kt_begin(axbc_ctx *c, int kind) {
  if (!kt_timed(c, kind)) return;
  hipEvent_t a = kt_event(c), b = kt_event(c);
  (void)hipEventRecord(a, c->stream);
  c->ktimer.pending.push_back({kind, {a, b}});
}

void kt_end(axbc_ctx *c, int kind) {
  if (!kt_timed(c, kind) || c->ktimer.pending.empty()) return;
  auto &pr = c->ktimer.pending.back();
  if (pr.first != kind) return;
  (void)hipEventRecord(pr.second.second, c->stream);
}

unsigned long long *kt_bytes(axbc_ctx *c, int kind) {
  return kt_timed(c, kind) && c->ktimer.dbytes ? c->ktimer.dbytes + (size_t)kind * kKbSlots * kKbStride : nullptr;
}

// Fold finished event pairs into the per-kind totals.
static int kt_collect(axbc_ctx *c) {
  KernelTimer &t = c->ktimer;
  if (t.pending.empty()) return AXBC_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto &pr : t.pending) {
    float ms = 0, t0 = 0, t1 = 0;
    if (hipEventElapsedTime(&ms, pr.second.first, pr.second.second) == hipSuccess) {
      t.ms[pr.first] += ms;
      t.launches[pr.first] += 1;
      if (t.ref && hipEventElapsedTime(&t0, t.ref, pr.second.first) == hipSuccess &&
          hipEventElapsedTime(&t1, t.ref, pr.second.second) == hipSuccess)
        t.spans[pr.first].push_back({(double)t0, (double)t1});
    }
    t.pool.push_back(pr.second.first);
    t.pool.push_back(pr.second.second);
  }
  t.pending.clear();
  return AXBC_OK;
}

int grid_for(uint64_t n, int cap) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}
}  // namespace axbc

namespace {

// Flags set by an earlier kernel of the stream (the done flag of the level
// loop): that kernel's completion already orders them, so a relaxed load
// suffices.  An acquire here would invalidate the vector cache in every
// wave, which costs tens of microseconds per launch at these grid sizes.
__device__ __forceinline__ uint32_t ld_flag(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Round set-up
// ---------------------------------------------------------------------------

__global__ void k_init_state(uint32_t n, const uint32_t *__restrict__ reach, int *lvl,
                             uint32_t *sigma, double *delta, double *bc) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    lvl[v] = -1;
    sigma[v] = 0;
    delta[v] = (double)reach[v];
    bc[v] = 0.0;
  }
}

// Undo the previous round on its visited vertices only: every other vertex
// still holds lvl -1, sigma 0, delta reach.
__global__ void k_reset_visited(const uint32_t *__restrict__ queue, const DevCtl *ctl,
                                const uint32_t *__restrict__ reach, int *lvl, uint32_t *sigma,
                                double *delta) {
  const uint32_t n = ctl->tail;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = queue[i];
    lvl[v] = -1;
    sigma[v] = 0;
    delta[v] = (double)reach[v];
  }
}

__global__ void k_reset_all(uint32_t n, const uint32_t *__restrict__ reach, int *lvl, uint32_t *sigma,
                            double *delta) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    lvl[v] = -1;
    sigma[v] = 0;
    delta[v] = (double)reach[v];
  }
}

// pull16: bottom-up ratio x16 of the round's levels (0: all top-down; only
// for graphs whose lists are their in-lists).
__global__ void k_set_root(uint32_t v0, int *lvl, uint32_t *sigma, uint32_t *queue, uint32_t *qoff,
                           DevCtl *ctl, const uint32_t *__restrict__ col, unsigned long long ned, uint32_t pull16) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t d0 = col[v0 + 1] - col[v0];
  lvl[v0] = 0;
  sigma[v0] = 1;
  queue[0] = v0;
  qoff[0] = 0;
  qoff[1] = 1;
  ctl->tail = 1;
  ctl->done = 0;
  ctl->empty_lev = 0;
  ctl->hub_claim = 0;
  ctl->wg_done = 0;
  ctl->all = 0;
  ctl->e_r = d0;  // E_r: edges of every vertex when it is found
  ctl->e_dag = 0;
  ctl->ef = 0;
  ctl->eu = ned - d0;
  ctl->bu = 0;  // level 1 (the root's neighbours) top-down
  ctl->pull16 = pull16;
}

// ---------------------------------------------------------------------------
// Forward level L: the frontier queue[qoff[L-1] .. qoff[L]) discovers level L.
// For every edge u->w: an unvisited w is claimed with a CAS on lvl[w] and
// appended; when lvl[w] == L the path count sigma[u] is added to sigma[w]
// (uint32, wrapping, order-free).
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool visit_edge(uint32_t w, uint32_t su, int L, int *lvl, uint32_t *sigma,
                                           unsigned long long &dag) {
  int l = lvl[w];
  bool fresh = false;
  if (l == -1) {
    const int old = atomicCAS(&lvl[w], -1, L);
    fresh = old == -1;
    l = fresh ? L : old;
  }
  if (l == L) {
    atomicAdd(&sigma[w], su);
    ++dag;
  }
  return fresh;
}

// Hub u (more than kHubDeg edges) of the current pass: one 64-bit atomic
// returns its slot k and its first chunk item (hub_claim = hubs << 40 |
// items), so hub_base[] increases with k.
__device__ __forceinline__ void claim_hub(DevCtl *ctl, uint32_t u, uint32_t d, uint32_t *hubs, uint64_t *hub_base,
                                          uint32_t hub_cap, uint64_t item_cap) {
  const uint32_t nch = (d + kHubChunk - 1) / kHubChunk;
  const unsigned long long old = atomicAdd(&ctl->hub_claim, (1ull << 40) + nch);
  const uint32_t k = (uint32_t)(old >> 40);
  const uint64_t base = old & ((1ull << 40) - 1);
  if (k < hub_cap && base + nch <= item_cap) {
    hubs[k] = u;
    hub_base[k] = base;
  } else {
    atomicOr(&ctl->overflow, 1u);
  }
}

// This wave's contiguous range of hub chunk items and the hub holding the
// first one.
struct HubWalk {
  uint32_t nh, k;
  uint64_t lo, hi, next;
};

__device__ __forceinline__ HubWalk hub_walk(const DevCtl *ctl, const uint64_t *__restrict__ hub_base, uint32_t hub_cap,
                                            uint64_t item_cap) {
  const unsigned long long claim = ctl->hub_claim;
  HubWalk r;
  r.nh = min((uint32_t)(claim >> 40), hub_cap);
  const uint64_t total = min((uint64_t)(claim & ((1ull << 40) - 1)), item_cap);
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwv = (gridDim.x * blockDim.x) / kWave;
  const uint64_t per = (total + nwv - 1) / nwv;
  r.lo = min(total, (uint64_t)gw * per);
  r.hi = min(total, r.lo + per);
  r.k = 0;
  r.next = ~0ull;
  if (r.lo < r.hi && r.nh > 0) {
    uint32_t lo = 0, hi = r.nh;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (hub_base[mid] <= r.lo)
        lo = mid;
      else
        hi = mid;
    }
    r.k = lo;
    r.next = lo + 1 < r.nh ? hub_base[lo + 1] : ~0ull;
  } else {
    r.hi = r.lo;
  }
  return r;
}

__device__ __forceinline__ void hub_step(HubWalk &r, const uint64_t *__restrict__ hub_base, uint64_t item) {
  while (item >= r.next) {
    ++r.k;
    r.next = r.k + 1 < r.nh ? hub_base[r.k + 1] : ~0ull;
  }
}

__device__ __forceinline__ void fwd_expand(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                           int *lvl, uint32_t *sigma, uint32_t *queue,
                                           const uint32_t *__restrict__ qoff, DevCtl *ctl, uint32_t *hubs,
                                           uint64_t *hub_base, uint32_t hub_cap, uint64_t item_cap, int L) {
  const uint32_t fb = qoff[L - 1], fe = qoff[L];
  const int lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nw = (gridDim.x * blockDim.x) / kWave;
  unsigned long long ef = 0, dag = 0;
  for (uint32_t base = fb + wid * kWave; base < fe; base += nw * kWave) {
    const uint32_t i = base + lane;
    uint32_t u = 0, st = 0, d = 0, su = 0;
    if (i < fe) {
      u = queue[i];
      st = col[u];
      d = col[u + 1] - st;
      su = sigma[u];
    }
    if (d > kHubDeg) {
      claim_hub(ctl, u, d, hubs, hub_base, hub_cap, item_cap);
      d = 0;
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t total = __shfl(incl, kWave - 1);
    for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
      const uint32_t e = e0 + lane;
      const int own = wave_owner(incl, e) & (kWave - 1);
      const uint32_t ost = __shfl(st, own);
      const uint32_t osu = __shfl(su, own);
      const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
      const uint32_t oex = own ? pin : 0u;
      bool fresh = false;
      uint32_t w = 0;
      if (e < total) {
        w = row[ost + (e - oex)];
        fresh = visit_edge(w, osu, L, lvl, sigma, dag);
        if (fresh) ef += col[w + 1] - col[w];
      }
      wave_append(fresh, w, queue, &ctl->tail);
    }
  }
  ef = wave_sum_u64(ef);
  dag = wave_sum_u64(dag);
  if (lane == 0 && (ef | dag)) {
    atomicAdd(&ctl->ef, ef);
    atomicAdd(&ctl->e_dag, dag);
  }
}

__device__ __forceinline__ void fwd_hubs(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                         int *lvl, uint32_t *sigma, uint32_t *queue, DevCtl *ctl,
                                         const uint32_t *__restrict__ hubs, const uint64_t *__restrict__ hub_base,
                                         uint32_t hub_cap, uint64_t item_cap, int L) {
  const int lane = lane_id();
  unsigned long long dag = 0, ef = 0;
  HubWalk r = hub_walk(ctl, hub_base, hub_cap, item_cap);
  for (uint64_t item = r.lo; item < r.hi; ++item) {
    hub_step(r, hub_base, item);
    const uint32_t u = hubs[r.k];
    const uint32_t st = col[u], d = col[u + 1] - st, su = sigma[u];
    const uint32_t cb = (uint32_t)(item - hub_base[r.k]) * kHubChunk, ce = min(d, cb + kHubChunk);
    for (uint32_t k0 = cb; k0 < ce; k0 += kWave) {
      const uint32_t k = k0 + lane;
      bool fresh = false;
      uint32_t w = 0;
      if (k < ce) {
        w = row[st + k];
        fresh = visit_edge(w, su, L, lvl, sigma, dag);
        if (fresh) ef += col[w + 1] - col[w];
      }
      wave_append(fresh, w, queue, &ctl->tail);
    }
  }
  dag = wave_sum_u64(dag);
  ef = wave_sum_u64(ef);
  if (lane == 0 && (dag | ef)) {
    atomicAdd(&ctl->e_dag, dag);
    atomicAdd(&ctl->ef, ef);
  }
}

// ---------------------------------------------------------------------------
// Bottom-up level L (graphs whose lists are their in-lists): every vertex w
// not yet found sums the path counts of its neighbours at level L-1,
//   sigma[w] = sum_{u in adj(w), lvl[u] = L-1} sigma[u]   (uint32, wrapping),
// the same sum the top-down push adds up edge by edge, and is found at L when
// it has one.  A wave takes 64 consecutive vertices and deals their lists 64
// edges per step with a segmented scan, so each vertex is written by one wave
// with no atomics; lists above the hub degree are claimed as hub chunks
// (fwd_pull_hubs).  Chosen per level by k_fwd_end when the lists of the
// vertices not yet found are at most pull16/16 times those a top-down level
// would scan.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void fwd_pull(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                         int *lvl, uint32_t *sigma, uint32_t *queue, DevCtl *ctl, uint32_t *hubs,
                                         uint64_t *hub_base, uint32_t hub_cap, uint64_t item_cap, uint32_t n,
                                         int L) {
  const int lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nw = (gridDim.x * blockDim.x) / kWave;
  unsigned long long dag = 0, ef = 0;
  for (uint32_t base = wid * kWave; base < n; base += nw * kWave) {
    const uint32_t w = base + lane;
    uint32_t st = 0, d = 0;
    if (w < n && lvl[w] == -1) {
      st = col[w];
      d = col[w + 1] - st;
    }
    const uint32_t dw = d;
    if (d > kHubDeg) {
      claim_hub(ctl, w, d, hubs, hub_base, hub_cap, item_cap);
      d = 0;
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t excl = incl - d;
    const uint32_t total = __shfl(incl, kWave - 1);
    uint32_t acc = 0, cnt = 0;
    for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
      const uint32_t e = e0 + lane;
      const int own = wave_owner(incl, e) & (kWave - 1);
      const uint32_t ost = __shfl(st, own);
      const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
      const uint32_t oex = own ? pin : 0u;
      uint32_t t = 0, tc = 0;
      int seg = kWave + lane;  // inactive lanes form singleton segments
      if (e < total) {
        seg = own;
        const uint32_t u = row[ost + (e - oex)];
        if (lvl[u] == L - 1) {
          t = sigma[u];
          tc = 1;
        }
      }
      // segmented inclusive scan over contiguous owner segments
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(t, off), yc = __shfl_up(tc, off);
        const int sy = __shfl_up(seg, off);
        if (lane >= off && sy == seg) {
          t += y;
          tc += yc;
        }
      }
      const bool has = d > 0 && excl < e0 + kWave && incl > e0;
      const uint32_t last = (incl < e0 + kWave ? incl : e0 + kWave) - 1u - e0;
      const uint32_t got = __shfl(t, has ? (int)last : lane), gotc = __shfl(tc, has ? (int)last : lane);
      if (has) {
        acc += got;
        cnt += gotc;
      }
    }
    const bool fresh = d > 0 && cnt > 0;
    if (fresh) {
      lvl[w] = L;
      sigma[w] = acc;
      dag += cnt;
      ef += dw;
    }
    wave_append(fresh, w, queue, &ctl->tail);
  }
  dag = wave_sum_u64(dag);
  ef = wave_sum_u64(ef);
  if (lane == 0 && (dag | ef)) {
    atomicAdd(&ctl->e_dag, dag);
    atomicAdd(&ctl->ef, ef);
  }
}

// Hub chunks of a bottom-up level: per item the sum and count of its
// neighbours at level L-1; the first item with any claims the hub (CAS on its
// level, then the append), every one adds its sum (uint32, order-free).
__device__ __forceinline__ void fwd_pull_hubs(const uint32_t *__restrict__ col, const uint32_t *__restrict__ row,
                                              int *lvl, uint32_t *sigma, uint32_t *queue, DevCtl *ctl,
                                              const uint32_t *__restrict__ hubs,
                                              const uint64_t *__restrict__ hub_base, uint32_t hub_cap,
                                              uint64_t item_cap, int L) {
  const int lane = lane_id();
  unsigned long long dag = 0, ef = 0;
  HubWalk r = hub_walk(ctl, hub_base, hub_cap, item_cap);
  for (uint64_t item = r.lo; item < r.hi; ++item) {
    hub_step(r, hub_base, item);
    const uint32_t w = hubs[r.k];
    const uint32_t st = col[w], d = col[w + 1] - st;
    const uint32_t cb = (uint32_t)(item - hub_base[r.k]) * kHubChunk, ce = min(d, cb + kHubChunk);
    uint32_t t = 0, tc = 0;
    for (uint32_t k = cb + lane; k < ce; k += kWave) {
      const uint32_t u = row[st + k];
      if (lvl[u] == L - 1) {
        t += sigma[u];
        ++tc;
      }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      t += __shfl_down(t, off);
      tc += __shfl_down(tc, off);
    }
    if (lane == 0 && tc > 0) {
      atomicAdd(&sigma[w], t);
      dag += tc;
      if (atomicCAS(&lvl[w], -1, L) == -1) {
        queue[atomicAdd(&ctl->tail, 1u)] = w;
        ef += d;
      }
    }
  }
  if (lane == 0 && (dag | ef)) {
    atomicAdd(&ctl->e_dag, dag);
    atomicAdd(&ctl->ef, ef);
  }
}

// Forward level L: the direction k_fwd_end chose for it (the flag bu),
// read once per wave, so one launch serves both and no empty kernel is
// launched per level.
__global__ __launch_bounds__(kBlock) void k_fwd_level(const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ row, int *lvl, uint32_t *sigma,
                                                      uint32_t *queue, const uint32_t *__restrict__ qoff,
                                                      DevCtl *ctl, uint32_t *hubs, uint64_t *hub_base,
                                                      uint32_t hub_cap, uint64_t item_cap, uint32_t n, int L) {
  if (ld_flag(&ctl->done)) return;
  if (ld_flag(&ctl->bu))
    fwd_pull(col, row, lvl, sigma, queue, ctl, hubs, hub_base, hub_cap, item_cap, n, L);
  else
    fwd_expand(col, row, lvl, sigma, queue, qoff, ctl, hubs, hub_base, hub_cap, item_cap, L);
}

// The hub chunks of forward level L, in the level's direction.
__global__ __launch_bounds__(kBlock) void k_fwd_hub_level(const uint32_t *__restrict__ col,
                                                          const uint32_t *__restrict__ row, int *lvl,
                                                          uint32_t *sigma, uint32_t *queue, DevCtl *ctl,
                                                          const uint32_t *__restrict__ hubs,
                                                          const uint64_t *__restrict__ hub_base, uint32_t hub_cap,
                                                          uint64_t item_cap, int L) {
  if (ld_flag(&ctl->done)) return;
  if (ld_flag(&ctl->bu))
    fwd_pull_hubs(col, row, lvl, sigma, queue, ctl, hubs, hub_base, hub_cap, item_cap, L);
  else
    fwd_hubs(col, row, lvl, sigma, queue, ctl, hubs, hub_base, hub_cap, item_cap, L);
}

// End of forward level L, one thread, after the level's kernels in stream
// order (their counts are complete without fences): level offset and done.
__global__ void k_fwd_end(uint32_t *qoff, DevCtl *ctl, int L) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || ctl->done) return;
  const uint32_t tail = ctl->tail;
  qoff[L + 1] = tail;
  if (tail == qoff[L]) {
    ctl->empty_lev = (uint32_t)L;
    ctl->done = 1u;
  }
  ctl->hub_claim = 0;
  // the next level's direction: bottom-up scans the lists of the vertices
  // not yet found, top-down those of the ones this level found
  const unsigned long long ef = ctl->ef;
  ctl->e_r += ef;
  ctl->eu -= ef < ctl->eu ? ef : ctl->eu;
  ctl->bu = ctl->pull16 && ctl->eu * 16ull <= (unsigned long long)ctl->pull16 * ef ? 1u : 0u;
  ctl->ef = 0;
}

// Clears the hub claims of a backward depth.
__global__ void k_hub_reset(DevCtl *ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->hub_claim = 0;
}

// ---------------------------------------------------------------------------
// Backward depth d: every v in queue[qoff[d] .. qoff[d+1]) pulls
//   delta[v] = reach[v] + sum_{w in adj(v), lvl[w] = d+1} (delta[w]+1)/sigma[w]*sigma[v]
// with no atomics (each vertex is written by exactly one wave or, for hubs,
// by the last workgroup of the hub pass).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double dep_term(uint32_t w, int dn, double sv, const int *__restrict__ lvl,
                                           const uint32_t *__restrict__ sigma,
                                           const double *__restrict__ delta) {
  return lvl[w] == dn ? ((delta[w] + 1.0) / (double)sigma[w]) * sv : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_bwd_pull(const uint32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ row,
                                                     const int *__restrict__ lvl,
                                                     const uint32_t *__restrict__ sigma, double *delta,
                                                     const uint32_t *__restrict__ reach,
                                                     const uint32_t *__restrict__ queue,
                                                     const uint32_t *__restrict__ qoff, DevCtl *ctl,
                                                     uint32_t *hubs, uint64_t *hub_base, uint32_t hub_cap,
                                                     uint64_t item_cap, int depth) {
  const uint32_t fb = qoff[depth], fe = qoff[depth + 1];
  const int dn = depth + 1;
  const int lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nw = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t base = fb + wid * kWave; base < fe; base += nw * kWave) {
    const uint32_t i = base + lane;
    uint32_t v = 0, st = 0, d = 0;
    double sv = 0.0;
    if (i < fe) {
      v = queue[i];
      st = col[v];
      d = col[v + 1] - st;
      sv = (double)sigma[v];
    }
    bool hub = d > kHubDeg;
    if (hub) {
      claim_hub(ctl, v, d, hubs, hub_base, hub_cap, item_cap);
      d = 0;
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t excl = incl - d;
    const uint32_t total = __shfl(incl, kWave - 1);
    double acc = 0.0;
    for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
      const uint32_t e = e0 + lane;
      const int own = wave_owner(incl, e) & (kWave - 1);
      const uint32_t ost = __shfl(st, own);
      const double osv = __shfl(sv, own);
      const uint32_t pin = __shfl(incl, (own + kWave - 1) & (kWave - 1));
      const uint32_t oex = own ? pin : 0u;
      double t = 0.0;
      int seg = kWave + lane;  // inactive lanes form singleton segments
      if (e < total) {
        seg = own;
        t = dep_term(row[ost + (e - oex)], dn, osv, lvl, sigma, delta);
      }
      // segmented inclusive scan over contiguous owner segments
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const double y = __shfl_up(t, off);
        const int sy = __shfl_up(seg, off);
        if (lane >= off && sy == seg) t += y;
      }
      // every vertex with edges in this step pulls its segment's total from
      // the lane holding its last edge of the step
      const bool has = d > 0 && excl < e0 + kWave && incl > e0;
      const uint32_t last = (incl < e0 + kWave ? incl : e0 + kWave) - 1u - e0;
      const double got = __shfl(t, has ? (int)last : lane);
      if (has) acc += got;
    }
    if (i < fe && !hub) delta[v] = (double)reach[v] + acc;
  }
}

// Hub chunks of depth d: per item the sum of its terms (lane partials in
// edge order, then a fixed wave tree) goes to hub_part[item].
__global__ __launch_bounds__(kBlock) void k_bwd_hubs(const uint32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ row,
                                                     const int *__restrict__ lvl,
                                                     const uint32_t *__restrict__ sigma,
                                                     const double *__restrict__ delta, const DevCtl *ctl,
                                                     const uint32_t *__restrict__ hubs,
                                                     const uint64_t *__restrict__ hub_base, uint32_t hub_cap,
                                                     uint64_t item_cap, double *part, int depth) {
  const int dn = depth + 1;
  const int lane = lane_id();
  HubWalk r = hub_walk(ctl, hub_base, hub_cap, item_cap);
  for (uint64_t item = r.lo; item < r.hi; ++item) {
    hub_step(r, hub_base, item);
    const uint32_t v = hubs[r.k];
    const uint32_t st = col[v], d = col[v + 1] - st;
    const double sv = (double)sigma[v];
    const uint32_t cb = (uint32_t)(item - hub_base[r.k]) * kHubChunk, ce = min(d, cb + kHubChunk);
    double t = 0.0;
    for (uint32_t k = cb + lane; k < ce; k += kWave) t += dep_term(row[st + k], dn, sv, lvl, sigma, delta);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) t += __shfl_down(t, off);
    if (lane == 0) part[item] = t;
  }
}

// One thread per hub: delta = reach + its chunk sums in chunk order.
__global__ __launch_bounds__(kBlock) void k_bwd_hub_combine(const uint32_t *__restrict__ col, double *delta,
                                                            const uint32_t *__restrict__ reach, DevCtl *ctl,
                                                            const uint32_t *__restrict__ hubs,
                                                            const uint64_t *__restrict__ hub_base, uint32_t hub_cap,
                                                            const double *__restrict__ part) {
  const unsigned long long claim = ctl->hub_claim;
  const uint32_t nh = min((uint32_t)(claim >> 40), hub_cap);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += gridDim.x * blockDim.x) {
    const uint32_t v = hubs[k];
    const uint32_t nch = (col[v + 1] - col[v] + kHubChunk - 1) / kHubChunk;
    const double *p = part + hub_base[k];
    double tot = 0.0;
    for (uint32_t c = 0; c < nch; ++c) tot += p[c];
    delta[v] = (double)reach[v] + tot;
  }
}

// ---------------------------------------------------------------------------
// bc update (update_bc, cu:269-286).  Only vertices whose delta is not the
// reset value 0 can change bc: the visited ones, and with -H 1 the
// unvisited ones with reach > 0 (delta = reach).  bc[v] += delta*(r0+1)/p.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_update_visited(const uint32_t *__restrict__ queue,
                                                           const DevCtl *ctl, uint32_t v0, double r1,
                                                           double p, const double *__restrict__ delta,
                                                           double *bc, const int *__restrict__ lvl,
                                                           const uint32_t *__restrict__ reach,
                                                           int sum_reach, uint32_t *all) {
  const uint32_t n = ctl->tail;
  uint32_t s = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = queue[i];
    if (v != v0) bc[v] += delta[v] * r1 / p;
    if (sum_reach && lvl[v] > 0) s += reach[v];
  }
  if (sum_reach) {
    s = (uint32_t)wave_sum_u64(s);
    if (lane_id() == 0 && s) atomicAdd(all, s);
  }
}

__global__ __launch_bounds__(kBlock) void k_update_reach_list(const uint32_t *__restrict__ rlist,
                                                              uint32_t nr, double r1, double p,
                                                              const int *__restrict__ lvl,
                                                              const double *__restrict__ delta,
                                                              double *bc) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
    const uint32_t v = rlist[i];
    if (lvl[v] == -1) bc[v] += delta[v] * r1 / p;
  }
}

__global__ void k_update_root(uint32_t v0, uint32_t r0, const DevCtl *ctl, double *bc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t all = ctl->all + r0;
  const unsigned long long nvt = (unsigned long long)ctl->tail + all;
  if (r0 > 0) bc[v0] += (double)((unsigned long long)r0 * (nvt - 2ull));
}

__global__ void k_stop_test(const uint32_t *__restrict__ check, uint32_t x, uint32_t n,
                            const double *__restrict__ bc, double cn, DevCtl *ctl) {
  uint32_t below = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < x; i += gridDim.x * blockDim.x) {
    const uint32_t a = check[i];
    const double b = a < n ? bc[a] : 0.0;
    if (b / 2.0 < cn) below = 1;
  }
  if (__ballot(below) && lane_id() == 0) atomicOr(&ctl->stop_any, 1u);
}

// Calibration streams: eight 16-byte loads in flight per thread before
// their stores (copy) or their fold (read), grid-strided and coalesced.  The
// copy uses ordinary loads and stores, as the engine's kernels do (it scales
// the hardware byte counters); the read uses nontemporal loads, the fastest
// stream found (sweeping the grid and the loads in flight on MI355X: 6.9 TB/s
// against 6.1 with cached loads), the attainable read rate.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kStreamRun = 8;
constexpr int kStreamGrid = 16384;

__global__ __launch_bounds__(kBlock) void k_copy(const u32x4 *__restrict__ src, u32x4 *dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (kStreamRun - 1) * stride < n16; i += kStreamRun * stride) {
    u32x4 v[kStreamRun];
#pragma unroll
    for (int u = 0; u < kStreamRun; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < kStreamRun; ++u) dst[i + u * stride] = v[u];
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

// The fold's result is stored only when it equals `never` (which the
// caller picks so that it does not occur), so the loads cannot be dropped.
__global__ __launch_bounds__(kBlock) void k_read(const u32x4 *__restrict__ src, uint64_t n16, uint32_t never,
                                                 uint32_t *sink) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (kStreamRun - 1) * stride < n16; i += kStreamRun * stride) {
    u32x4 v[kStreamRun];
#pragma unroll
    for (int u = 0; u < kStreamRun; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kStreamRun; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < n16; i += stride) acc ^= src[i].x ^ src[i].y ^ src[i].z ^ src[i].w;
  if (acc == never) sink[0] = acc;
}

// Random-line gather: every lane reads one double from its own random
// 128-byte line among the first `lines` lines of src (xorshift32 per lane),
// kGatherRun loads in flight per thread, `iters` rounds.  The rate of line
// fetches this reaches over a region of a given size bounds the kernels
// whose rows are read at scattered vertices.
constexpr int kGatherRun = 4;
constexpr int kGatherGrid = 4096;
constexpr uint32_t kGatherIters = 32;

__global__ __launch_bounds__(kBlock) void k_gather(const double *__restrict__ src, uint32_t lines, uint32_t iters,
                                                   uint32_t never, uint32_t *sink) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t s = gid * 2654435761u + 12345u;
  if (s == 0) s = 1;
  double acc = 0.0;
  for (uint32_t it = 0; it < iters; ++it) {
    double v[kGatherRun];
#pragma unroll
    for (int u = 0; u < kGatherRun; ++u) {
      s ^= s << 13;
      s ^= s >> 17;
      s ^= s << 5;
      v[u] = src[(size_t)(s % lines) * 16 + (threadIdx.x & 15)];
    }
#pragma unroll
    for (int u = 0; u < kGatherRun; ++u) acc += v[u];
  }
  if (__double_as_longlong(acc) == (long long)never) sink[0] = 1;
}

__global__ void k_degrees(const uint32_t *__restrict__ col, uint32_t n, uint32_t *deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    deg[v] = col[v + 1] - col[v];
}

}  // namespace

namespace {

template <typename T>
int dalloc(axbc_ctx *c, T **p, size_t count) {
  size_t b = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc((void **)p, b);
  if (e != hipSuccess) return fail(AXBC_ENOMEM, "hipMalloc of %s bytes failed (line %d)", "device", __LINE__);
  c->bytes += b;
  return AXBC_OK;
}

int pull_ctl(axbc_ctx *c) {
  HIP_TRY(hipMemcpyAsync(c->h_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

// A forward level: two launches whatever its direction (the device flag
// selects top-down or bottom-up inside them).
int enqueue_forward(axbc_ctx *c, int L) {
  kt_begin(c, KK_FWD);
  hipLaunchKernelGGL(k_fwd_level, dim3(kExpandGrid), dim3(kBlock), 0, c->stream, c->col, c->row, c->lvl,
                     c->sigma, c->queue, c->qoff, c->ctl, c->hubs, c->hub_base, c->nhub_cap, c->hub_items, c->n, L);
  kt_end(c, KK_FWD);
  kt_begin(c, KK_FWD_HUBS);
  hipLaunchKernelGGL(k_fwd_hub_level, dim3(kHubGrid), dim3(kBlock), 0, c->stream, c->col, c->row, c->lvl,
                     c->sigma, c->queue, c->ctl, c->hubs, c->hub_base, c->nhub_cap, c->hub_items, L);
  kt_end(c, KK_FWD_HUBS);
  hipLaunchKernelGGL(k_fwd_end, dim3(1), dim3(64), 0, c->stream, c->qoff, c->ctl, L);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

int enqueue_backward(axbc_ctx *c, int depth) {
  kt_begin(c, KK_BWD);
  hipLaunchKernelGGL(k_bwd_pull, dim3(kExpandGrid), dim3(kBlock), 0, c->stream, c->col, c->row,
                     c->lvl, c->sigma, c->delta, c->reach, c->queue, c->qoff, c->ctl, c->hubs, c->hub_base,
                     c->nhub_cap, c->hub_items, depth);
  kt_end(c, KK_BWD);
  kt_begin(c, KK_BWD_HUBS);
  hipLaunchKernelGGL(k_bwd_hubs, dim3(kHubGrid), dim3(kBlock), 0, c->stream, c->col, c->row, c->lvl,
                     c->sigma, c->delta, c->ctl, c->hubs, c->hub_base, c->nhub_cap, c->hub_items, c->hub_part,
                     depth);
  kt_end(c, KK_BWD_HUBS);
  hipLaunchKernelGGL(k_bwd_hub_combine, dim3(64), dim3(kBlock), 0, c->stream, c->col, c->delta, c->reach,
                     c->ctl, c->hubs, c->hub_base, c->nhub_cap, c->hub_part);
  hipLaunchKernelGGL(k_hub_reset, dim3(1), dim3(64), 0, c->stream, c->ctl);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

int enqueue_reset_and_root(axbc_ctx *c, uint32_t v0) {
  if (c->dirty_delta) {
    hipLaunchKernelGGL(k_reset_all, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->reach, c->lvl,
                       c->sigma, c->delta);
    c->dirty_delta = false;
  } else if (c->have_round)
    hipLaunchKernelGGL(k_reset_visited, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->queue,
                       c->ctl, c->reach, c->lvl, c->sigma, c->delta);
  // bottom-up levels only where a vertex's list is its in-list
  bool sym = false;
  int rc;
  if (c->plain_pull16 && (rc = mirror_symmetric(c, &sym))) return rc;
  hipLaunchKernelGGL(k_set_root, dim3(1), dim3(64), 0, c->stream, v0, c->lvl, c->sigma, c->queue,
                     c->qoff, c->ctl, c->col, (unsigned long long)c->ned, sym ? c->plain_pull16 : 0u);
  HIP_TRY(hipGetLastError());
  c->have_round = true;
  return AXBC_OK;
}

int enqueue_update(axbc_ctx *c, uint32_t v0, double p) {
  const uint32_t r0 = c->h_reach.empty() ? 0u : c->h_reach[v0];
  const double r1 = (double)(uint32_t)(r0 + 1u);
  const int sum_reach = c->heuristic == 1 || c->heuristic == 3;
  hipLaunchKernelGGL(k_update_visited, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->queue,
                     c->ctl, v0, r1, p, c->delta, c->bc, c->lvl, c->reach, sum_reach, &c->ctl->all);
  if (c->nr)
    hipLaunchKernelGGL(k_update_reach_list, dim3(grid_for(c->nr)), dim3(kBlock), 0, c->stream,
                       c->rlist, c->nr, r1, p, c->lvl, c->delta, c->bc);
  hipLaunchKernelGGL(k_update_root, dim3(1), dim3(64), 0, c->stream, v0, r0, c->ctl, c->bc);
  HIP_TRY(hipGetLastError());
  return AXBC_OK;
}

}  // namespace

extern "C" {

const char *axbc_last_error(void) { return g_err; }

int axbc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int axbc_ctx_create(int device, uint64_t N, uint64_t ned, const uint32_t *col, const uint32_t *row,
                    const uint32_t *reach, int heuristic, axbc_ctx **out, size_t *dev_bytes) {
  if (!out || !col || (ned && !row)) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  if (N == 0 || N > 0xFFFFFFF0ull || ned > 0xFFFFFFFFull)
    return fail(AXBC_EINVAL, "graph size out of range%s (line %d)", "", __LINE__);
  if (col[0] != 0 || col[N] != ned) return fail(AXBC_EINVAL, "col[] does not span row[]%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(device));
  axbc_ctx *c = new axbc_ctx();
  c->device = device;
  c->n = (uint32_t)N;
  c->ned = ned;
  c->heuristic = heuristic;
  int rc;
#define TRY_ALLOC(p, cnt)                  \
  if ((rc = dalloc(c, &(p), (cnt))) != 0) { \
    axbc_ctx_destroy(c);                   \
    return rc;                             \
  }
  // hubs: vertices whose degree exceeds the hub threshold
  uint32_t nhub = 0;
  for (uint64_t v = 0; v < N; ++v) nhub += (col[v + 1] - col[v]) > kHubDeg;
  c->nhub_cap = nhub ? nhub : 1;
  c->h_reach.assign(N, 0u);
  std::vector<uint32_t> rl;
  if (reach) {
    memcpy(c->h_reach.data(), reach, sizeof(uint32_t) * N);
    for (uint64_t v = 0; v < N; ++v)
      if (reach[v]) rl.push_back((uint32_t)v);
  }
  c->nr = (uint32_t)rl.size();
  c->h_col.assign(col, col + N + 1);
  TRY_ALLOC(c->col, N + 1);
  TRY_ALLOC(c->row, ned);
  TRY_ALLOC(c->reach, N);
  TRY_ALLOC(c->rlist, rl.size());
  TRY_ALLOC(c->lvl, N);
  TRY_ALLOC(c->sigma, N);
  TRY_ALLOC(c->delta, N);
  TRY_ALLOC(c->bc, N);
  TRY_ALLOC(c->queue, N);
  TRY_ALLOC(c->qoff, N + 2);
  // chunk items per pass: sum over distinct hubs of ceil(d / kHubChunk)
  c->hub_items = ned / kHubChunk + c->nhub_cap + 1;
  TRY_ALLOC(c->hubs, c->nhub_cap);
  TRY_ALLOC(c->hub_base, c->nhub_cap);
  TRY_ALLOC(c->hub_part, c->hub_items);
  TRY_ALLOC(c->ctl, 1);
#undef TRY_ALLOC
  if (hipHostMalloc((void **)&c->h_ctl, sizeof(DevCtl), hipHostMallocDefault) != hipSuccess) {
    axbc_ctx_destroy(c);
    return fail(AXBC_ENOMEM, "pinned allocation failed%s (line %d)", "", __LINE__);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    axbc_ctx_destroy(c);
    return fail(AXBC_EHIP, "stream creation failed%s (line %d)", "", __LINE__);
  }
  for (auto &e : c->ev) (void)hipEventCreate(&e);
  hipError_t e1 = hipMemcpyAsync(c->col, col, sizeof(uint32_t) * (N + 1), hipMemcpyHostToDevice, c->stream);
  hipError_t e2 = ned ? hipMemcpyAsync(c->row, row, sizeof(uint32_t) * ned, hipMemcpyHostToDevice, c->stream)
                      : hipSuccess;
  hipError_t e3 = hipMemcpyAsync(c->reach, c->h_reach.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice,
                                 c->stream);
  hipError_t e4 = rl.empty() ? hipSuccess
                             : hipMemcpyAsync(c->rlist, rl.data(), sizeof(uint32_t) * rl.size(),
                                              hipMemcpyHostToDevice, c->stream);
  (void)hipMemsetAsync(c->ctl, 0, sizeof(DevCtl), c->stream);
  hipLaunchKernelGGL(k_init_state, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, (uint32_t)N, c->reach,
                     c->lvl, c->sigma, c->delta, c->bc);
  hipError_t e5 = hipStreamSynchronize(c->stream);
  if (e1 || e2 || e3 || e4 || e5) {
    axbc_ctx_destroy(c);
    return fail(AXBC_EHIP, "upload failed: %s (line %d)", hipGetErrorString(e5 ? e5 : e1), __LINE__);
  }
  *out = c;
  if (dev_bytes) *dev_bytes = c->bytes;
  return AXBC_OK;
}

int axbc_ctx_destroy(axbc_ctx *c) {
  if (!c) return AXBC_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  sampler_free(c);
  batch_free_all(c);
  mlp_free(c);
  for (auto &pr : c->ktimer.pending) {
    (void)hipEventDestroy(pr.second.first);
    (void)hipEventDestroy(pr.second.second);
  }
  for (hipEvent_t e : c->ktimer.pool) (void)hipEventDestroy(e);
  if (c->ktimer.ref) (void)hipEventDestroy(c->ktimer.ref);
  if (c->ktimer.dbytes) (void)hipFree(c->ktimer.dbytes);
  if (!c->symmetric) {
    if (c->rcol) (void)hipFree(c->rcol);
    if (c->rrow) (void)hipFree(c->rrow);
  }
  void *ptrs[] = {c->col,   c->row,  c->reach, c->rlist,  c->lvl,      c->sigma, c->delta, c->bc,
                  c->queue, c->qoff, c->hubs,  c->hub_base, c->hub_part, c->ctl,   c->check};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AXBC_OK;
}

int axbc_pivot(axbc_ctx *c, uint32_t v0, double p, axbc_pivot_stats *stats) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  if (v0 >= c->n) return fail(AXBC_EINVAL, "root out of range%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(c->device));
  int rc;
  (void)hipEventRecord(c->ev[0], c->stream);
  if ((rc = enqueue_reset_and_root(c, v0))) return rc;
  // Forward levels, enqueued in chunks; the device done flag makes the
  // kernels past the last level exit immediately.
  int L = 1, chunk = c->lev_guess;
  for (;;) {
    for (int j = 0; j < chunk; ++j)
      if ((rc = enqueue_forward(c, L + j))) return rc;
    L += chunk;
    if ((rc = pull_ctl(c))) return rc;
    if (c->h_ctl->done) break;
    if ((uint64_t)L > (uint64_t)c->n + 1) return fail(AXBC_ESTATE, "forward did not terminate%s (line %d)", "", __LINE__);
    chunk = chunk * 2;
  }
  (void)hipEventRecord(c->ev[1], c->stream);
  const int nlev = (int)c->h_ctl->empty_lev - 1;
  c->last_nlev = nlev;
  c->lev_guess = nlev + 2;
  // Backward depths with bc_func_mono's do-while: nlev-1 .. 1, or 0 alone
  // when the deepest level is 1; nothing when only the root was reached.
  if (nlev >= 1) {
    int depth = nlev - 1;
    do {
      if ((rc = enqueue_backward(c, depth))) return rc;
      --depth;
    } while (depth > 0);
  }
  (void)hipEventRecord(c->ev[2], c->stream);
  if ((rc = enqueue_update(c, v0, p))) return rc;
  (void)hipEventRecord(c->ev[3], c->stream);
  c->last_v0 = v0;
  if (stats) return axbc_last_stats(c, stats);
  return AXBC_OK;
}

int axbc_last_stats(axbc_ctx *c, axbc_pivot_stats *st) {
  if (!c || !st) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  int rc;
  if ((rc = pull_ctl(c))) return rc;
  const DevCtl *h = c->h_ctl;
  const uint32_t r0 = c->h_reach.empty() ? 0u : c->h_reach[c->last_v0];
  st->nvisited = h->tail;
  st->nlevels = (uint64_t)(h->empty_lev ? h->empty_lev - 1 : 0);
  st->e_r = h->e_r;
  st->e_dag = h->e_dag;
  st->all = (uint32_t)(h->all + r0);
  return AXBC_OK;
}

int axbc_last_timing(axbc_ctx *c, float *f, float *b, float *u) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  HIP_TRY(hipEventSynchronize(c->ev[3]));
  float x = 0, y = 0, z = 0;
  HIP_TRY(hipEventElapsedTime(&x, c->ev[0], c->ev[1]));
  HIP_TRY(hipEventElapsedTime(&y, c->ev[1], c->ev[2]));
  HIP_TRY(hipEventElapsedTime(&z, c->ev[2], c->ev[3]));
  if (f) *f = x;
  if (b) *b = y;
  if (u) *u = z;
  return AXBC_OK;
}

static int kt_enable(axbc_ctx *c, bool enable, uint32_t mask) {
  KernelTimer &t = c->ktimer;
  if (enable) {
    const size_t nb = sizeof(unsigned long long) * KK_COUNT * kKbSlots * kKbStride;
    if (!t.dbytes) HIP_TRY(hipMalloc((void **)&t.dbytes, nb));
    HIP_TRY(hipMemsetAsync(t.dbytes, 0, nb, c->stream));
    int rc = kt_collect(c);
    if (rc) return rc;
    for (int k = 0; k < KK_COUNT; ++k) {
      t.ms[k] = 0;
      t.launches[k] = 0;
      t.spans[k].clear();
    }
    if (!t.ref) HIP_TRY(hipEventCreate(&t.ref));
    HIP_TRY(hipEventRecord(t.ref, c->stream));
  } else {
    int rc = kt_collect(c);
    if (rc) return rc;
  }
  t.on = enable;
  t.mask = mask;
  return AXBC_OK;
}

int axbc_kernel_timing(axbc_ctx *c, int enable) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  return kt_enable(c, enable != 0, ~0u);
}

int axbc_kernel_timing_kinds(axbc_ctx *c, uint32_t mask) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  if (mask >> KK_COUNT) return fail(AXBC_EINVAL, "unknown kernel kind in the mask%s (line %d)", "", __LINE__);
  return kt_enable(c, mask != 0, mask ? mask : ~0u);
}

int axbc_kernel_stats(axbc_ctx *c, int kind, double *ms, uint64_t *launches, uint64_t *bytes) {
  if (!c || kind < 0 || kind >= KK_COUNT) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  int rc = kt_collect(c);
  if (rc) return rc;
  unsigned long long b = 0;
  std::vector<unsigned long long> slots((size_t)kKbSlots * kKbStride, 0ull);
  if (c->ktimer.dbytes)
    HIP_TRY(hipMemcpyAsync(slots.data(), c->ktimer.dbytes + (size_t)kind * kKbSlots * kKbStride,
                           sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < kKbSlots; ++k) b += slots[(size_t)k * kKbStride];
  if (ms) *ms = c->ktimer.ms[kind];
  if (launches) *launches = c->ktimer.launches[kind];
  if (bytes) *bytes = b;
  return AXBC_OK;
}

int axbc_kernel_busy(axbc_ctx *c, int kind, double *ms) {
  if (!c || !ms || kind < 0 || kind >= KK_COUNT) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  int rc = kt_collect(c);
  if (rc) return rc;
  std::vector<std::pair<double, double>> s = c->ktimer.spans[kind];
  std::sort(s.begin(), s.end());
  double busy = 0, end = -1e300;
  for (const auto &iv : s) {
    if (iv.first > end) {
      busy += iv.second - iv.first;
      end = iv.second;
    } else if (iv.second > end) {
      busy += iv.second - end;
      end = iv.second;
    }
  }
  *ms = busy;
  return AXBC_OK;
}

int axbc_read_kernel(axbc_ctx *c, uint64_t bytes, int repeats) {
  if (!c || bytes < 16) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  void *a = nullptr;
  uint32_t *sink = nullptr;
  HIP_TRY(hipMalloc(&a, bytes));
  HIP_TRY(hipMalloc((void **)&sink, sizeof(uint32_t)));
  // every word 0x01010101: a fold of an even count of them is 0, of an odd
  // count 0x01010101; `never` is neither
  HIP_TRY(hipMemsetAsync(a, 1, bytes, c->stream));
  for (int r = 0; r < repeats; ++r) {
    kt_begin(c, KK_READ);
    hipLaunchKernelGGL(k_read, dim3(kStreamGrid), dim3(kBlock), 0, c->stream, (const u32x4 *)a, bytes / 16,
                       0xdeadbeefu, sink);
    kt_end(c, KK_READ);
  }
  if (unsigned long long *kb = kt_bytes(c, KK_READ)) {
    const unsigned long long add = (bytes / 16) * 16 * (unsigned long long)repeats;
    unsigned long long cur = 0;
    HIP_TRY(hipMemcpyAsync(&cur, kb, sizeof(cur), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    cur += add;
    HIP_TRY(hipMemcpyAsync(kb, &cur, sizeof(cur), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(a));
  HIP_TRY(hipFree(sink));
  return AXBC_OK;
}

int axbc_gather_kernel(axbc_ctx *c, uint64_t window, int repeats) {
  if (!c || window < 128 || window / 128 > 0xffffffffull)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  const uint64_t lines = window / 128;
  void *a = nullptr;
  uint32_t *sink = nullptr;
  HIP_TRY(hipMalloc(&a, lines * 128));
  HIP_TRY(hipMalloc((void **)&sink, sizeof(uint32_t)));
  // zeros: the sum stays +0.0, whose bits are not `never`
  HIP_TRY(hipMemsetAsync(a, 0, lines * 128, c->stream));
  for (int r = 0; r < repeats; ++r) {
    kt_begin(c, KK_GATHER);
    hipLaunchKernelGGL(k_gather, dim3(kGatherGrid), dim3(kBlock), 0, c->stream, (const double *)a, (uint32_t)lines,
                       kGatherIters, 7u, sink);
    kt_end(c, KK_GATHER);
  }
  if (unsigned long long *kb = kt_bytes(c, KK_GATHER)) {
    // the bytes of the lines fetched: one 128-byte line per load
    const unsigned long long add =
        128ull * kGatherGrid * kBlock * kGatherIters * kGatherRun * (unsigned long long)repeats;
    unsigned long long cur = 0;
    HIP_TRY(hipMemcpyAsync(&cur, kb, sizeof(cur), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    cur += add;
    HIP_TRY(hipMemcpyAsync(kb, &cur, sizeof(cur), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(a));
  HIP_TRY(hipFree(sink));
  return AXBC_OK;
}

int axbc_copy_kernel(axbc_ctx *c, uint64_t bytes, int repeats) {
  if (!c || bytes < 16) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  void *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, bytes));
  HIP_TRY(hipMalloc(&b, bytes));
  HIP_TRY(hipMemsetAsync(a, 1, bytes, c->stream));
  for (int r = 0; r < repeats; ++r) {
    kt_begin(c, KK_COPY);
    hipLaunchKernelGGL(k_copy, dim3(kStreamGrid), dim3(kBlock), 0, c->stream, (const u32x4 *)a, (u32x4 *)b, bytes / 16);
    kt_end(c, KK_COPY);
  }
  if (unsigned long long *kb = kt_bytes(c, KK_COPY)) {
    const unsigned long long add = 2ull * (bytes / 16) * 16 * (unsigned long long)repeats;
    unsigned long long cur = 0;
    HIP_TRY(hipMemcpyAsync(&cur, kb, sizeof(cur), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    cur += add;
    HIP_TRY(hipMemcpyAsync(kb, &cur, sizeof(cur), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(a));
  HIP_TRY(hipFree(b));
  return AXBC_OK;
}

int axbc_sync(axbc_ctx *c) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_set_memory_budget(axbc_ctx *c, size_t bytes) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  c->budget = bytes;
  return AXBC_OK;
}

int axbc_device_bytes(axbc_ctx *c, size_t *bytes) {
  if (!c || !bytes) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  *bytes = c->bytes;
  return AXBC_OK;
}

int axbc_reset_bc(axbc_ctx *c) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemsetAsync(c->bc, 0, sizeof(double) * c->n, c->stream));
  return AXBC_OK;
}

static int copy_out(axbc_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || !dst) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_get_bc(axbc_ctx *c, double *bc) { return copy_out(c, bc, c ? c->bc : nullptr, c ? sizeof(double) * c->n : 0); }
int axbc_get_delta(axbc_ctx *c, double *d) { return copy_out(c, d, c ? c->delta : nullptr, c ? sizeof(double) * c->n : 0); }
int axbc_get_lvl(axbc_ctx *c, int32_t *l) { return copy_out(c, l, c ? c->lvl : nullptr, c ? sizeof(int) * c->n : 0); }
int axbc_get_sigma(axbc_ctx *c, uint32_t *s) { return copy_out(c, s, c ? c->sigma : nullptr, c ? sizeof(uint32_t) * c->n : 0); }

int axbc_get_deg(axbc_ctx *c, uint32_t *deg) {
  if (!c || !deg) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  uint32_t *d = nullptr;
  HIP_TRY(hipMalloc((void **)&d, sizeof(uint32_t) * c->n));
  hipLaunchKernelGGL(k_degrees, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->col, c->n, d);
  HIP_TRY(hipMemcpyAsync(deg, d, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(d));
  return AXBC_OK;
}

int axbc_set_bc(axbc_ctx *c, const double *bc) {
  if (!c || !bc) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemcpyAsync(c->bc, bc, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_set_reach(axbc_ctx *c, const uint32_t *reach) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  std::vector<uint32_t> rl;
  if (reach) {
    memcpy(c->h_reach.data(), reach, sizeof(uint32_t) * c->n);
    for (uint32_t v = 0; v < c->n; ++v)
      if (reach[v]) rl.push_back(v);
  } else {
    std::fill(c->h_reach.begin(), c->h_reach.end(), 0u);
  }
  if (c->rlist) (void)hipFree(c->rlist);
  c->rlist = nullptr;
  c->nr = (uint32_t)rl.size();
  HIP_TRY(hipMalloc((void **)&c->rlist, sizeof(uint32_t) * (rl.empty() ? 1 : rl.size())));
  if (!rl.empty())
    HIP_TRY(hipMemcpyAsync(c->rlist, rl.data(), sizeof(uint32_t) * rl.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->reach, c->h_reach.data(), sizeof(uint32_t) * c->n, hipMemcpyHostToDevice, c->stream));
  // the reset state holds delta = reach
  hipLaunchKernelGGL(k_init_state, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->reach, c->lvl,
                     c->sigma, c->delta, c->bc);
  c->have_round = false;
  batch_free_all(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

void *axbc_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void axbc_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

int axbc_set_checks(axbc_ctx *c, const uint32_t *check, uint32_t x) {
  if (!c || (x && !check)) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  if (c->check) {
    (void)hipFree(c->check);
    c->check = nullptr;
  }
  c->ncheck = x;
  if (!x) return AXBC_OK;
  HIP_TRY(hipMalloc((void **)&c->check, sizeof(uint32_t) * x));
  HIP_TRY(hipMemcpyAsync(c->check, check, sizeof(uint32_t) * x, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AXBC_OK;
}

int axbc_stop_test(axbc_ctx *c, uint32_t cc, int *stop) {
  if (!c || !stop) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  HIP_TRY(hipMemsetAsync(&c->ctl->stop_any, 0, sizeof(uint32_t), c->stream));
  if (c->ncheck)
    hipLaunchKernelGGL(k_stop_test, dim3(grid_for(c->ncheck, 64)), dim3(kBlock), 0, c->stream, c->check,
                       c->ncheck, c->n, c->bc, (double)((uint64_t)cc * (uint64_t)c->n), c->ctl);
  int rc;
  if ((rc = pull_ctl(c))) return rc;
  *stop = c->h_ctl->stop_any ? 0 : 1;
  return AXBC_OK;
}

// ---- fine-grained entry points (one host synchronisation per call) -------

int axbc_begin_round(axbc_ctx *c) {
  if (!c) return fail(AXBC_EINVAL, "null context%s (line %d)", "", __LINE__);
  if (c->dirty_delta) {
    hipLaunchKernelGGL(k_reset_all, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->reach, c->lvl,
                       c->sigma, c->delta);
    c->dirty_delta = false;
  } else if (c->have_round)
    hipLaunchKernelGGL(k_reset_visited, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->queue,
                       c->ctl, c->reach, c->lvl, c->sigma, c->delta);
  HIP_TRY(hipMemsetAsync(c->ctl, 0, sizeof(DevCtl), c->stream));
  c->have_round = false;
  return axbc_sync(c);
}

int axbc_set_root(axbc_ctx *c, uint32_t v0) {
  if (!c || v0 >= c->n) return fail(AXBC_EINVAL, "bad root%s (line %d)", "", __LINE__);
  // the per-level entry points keep every level top-down (the reference's
  // scan_col_mono2 order of work)
  hipLaunchKernelGGL(k_set_root, dim3(1), dim3(64), 0, c->stream, v0, c->lvl, c->sigma, c->queue, c->qoff,
                     c->ctl, c->col, (unsigned long long)c->ned, 0u);
  c->have_round = true;
  c->last_v0 = v0;
  return axbc_sync(c);
}

int axbc_forward_level(axbc_ctx *c, int level, uint32_t *nfrt) {
  if (!c || level < 1) return fail(AXBC_EINVAL, "bad level%s (line %d)", "", __LINE__);
  int rc;
  if ((rc = enqueue_forward(c, level))) return rc;
  if ((rc = pull_ctl(c))) return rc;
  uint32_t q[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(q, c->qoff + level, sizeof(q), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (nfrt) *nfrt = c->h_ctl->done && (int)c->h_ctl->empty_lev <= level ? 0u : q[1] - q[0];
  return AXBC_OK;
}

int axbc_backward_depth(axbc_ctx *c, int depth) {
  if (!c || depth < 0) return AXBC_OK;  // the reference tolerates depth -1
  int rc;
  if ((rc = enqueue_backward(c, depth))) return rc;
  return axbc_sync(c);
}

int axbc_reach_sum(axbc_ctx *c, uint32_t *all) {
  if (!c || !all) return fail(AXBC_EINVAL, "null argument%s (line %d)", "", __LINE__);
  // The reach sum is produced by the visited-vertex pass of the update; a
  // separate pass is only needed for this fine-grained entry point.
  int rc;
  if ((rc = pull_ctl(c))) return rc;
  std::vector<uint32_t> q(c->h_ctl->tail);
  std::vector<int> l(c->n);
  HIP_TRY(hipMemcpyAsync(q.data(), c->queue, sizeof(uint32_t) * q.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(l.data(), c->lvl, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  uint32_t s = 0;
  for (uint32_t v : q)
    if (l[v] > 0) s += c->h_reach[v];
  *all = s;
  return AXBC_OK;
}

int axbc_update_bc(axbc_ctx *c, uint32_t v0, double p, uint64_t nvisited) {
  if (!c || v0 >= c->n) return fail(AXBC_EINVAL, "bad root%s (line %d)", "", __LINE__);
  (void)nvisited;  // recomputed on the device from the queue and reach sum
  int rc;
  if ((rc = enqueue_update(c, v0, p))) return rc;
  return axbc_sync(c);
}

}  // extern "C"
