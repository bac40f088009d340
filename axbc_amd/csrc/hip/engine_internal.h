// Internal definitions shared by the engine's HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../../include/axbc.h"

namespace axbc {

constexpr int kBlock = 256;
constexpr int kExpandGrid = 1024;   // workgroups of the level kernels
constexpr int kHubGrid = 512;       // workgroups of the hub passes
constexpr uint32_t kHubDeg = 1024;    // degree above which a vertex is a hub
constexpr uint32_t kHubChunk = 1024;  // edges per work item of a hub

struct DevCtl {
  uint32_t tail;       // queue length of the current round
  uint32_t done;       // forward finished
  uint32_t empty_lev;  // first empty level (valid when done)
  uint32_t unused0;
  uint32_t wg_done;    // last-workgroup counter
  uint32_t all;        // reach sum of visited non-root vertices
  uint32_t stop_any;   // stop test: a check vertex below threshold
  uint32_t overflow;   // a bounded list ran out of room
  unsigned long long e_r, e_dag;
  unsigned long long hub_claim;  // hubs of the current pass << 40 | chunk items
  // direction of the forward levels: ef = edges of the vertices found in the
  // current level (what a top-down next level scans), eu = edges of the
  // vertices not yet found (what a bottom-up next level scans); bu = the
  // current level runs bottom-up; pull16 = bottom-up when eu <= pull16/16 ef
  // (0: never)
  unsigned long long ef, eu;
  uint32_t bu, pull16;
};

// Device buffers of the samplers (allocated on first use).
struct SamplerBufs {
  void *tmp = nullptr;    // N x 8 bytes: deg^2 (u32) or fixed-point weights (u64)
  void *u32a = nullptr;   // N x 4: wrapped deg^2 prefix
  void *u32b = nullptr;   // N x 4: its running maximum
  void *u32c = nullptr;   // N x 4: neighbour-degree sums
  void *u64a = nullptr;   // N x 8: weight prefix
  void *tiles = nullptr;  // scan tile totals
  unsigned long long *wmax = nullptr;
  uint32_t *pick = nullptr;
  uint32_t *pick_host = nullptr;  // pinned
  uint64_t *host64 = nullptr;     // pinned, 2 values
  uint64_t total = 0;
  int hdn_ready = 0, weights_ready = 0;
  // static LCC distribution (-z 2 / -z 6)
  float *lcc = nullptr;
  uint32_t *lcc_cnt = nullptr;
  uint64_t *lw = nullptr, *lp = nullptr;  // fixed-point weights and their prefix
  uint64_t ltotal = 0;
  int lcc_ready = 0;
  int lcc_sym = -1;  // CSC symmetric, loop-free, strictly sorted: 1 yes, 0 no, -1 not checked
  // draw windows (axbc_draw_window): their own stream, so that a window is
  // searched while a batch runs on the context stream
  hipStream_t wstream = nullptr;
  uint32_t *win = nullptr;  // 2 x win_cap: lrand48 values, picked vertices
  uint32_t win_cap = 0;
};

struct BatchState;  // batch.hip

// Per-kernel instrumentation for the roofline figures: HIP events around
// every launch of the selected kernel kinds, and algorithmic byte counts
// accumulated by the kernels themselves (counted only while enabled).
enum KernelKind {
  KK_PUSH = 0,      // batched forward marking (k_bpush)
  KK_SIGMA = 1,     // batched path-count pull (now inside k_bclose; no launches)
  KK_CLOSE = 2,     // batched level close and path counts (k_bclose)
  KK_DEP = 3,       // batched dependency pull (k_bdep)
  KK_UPDATE = 4,    // batched bc update (k_bupdate)
  KK_FWD = 5,       // plain forward level, either direction (k_fwd_level)
  KK_BWD = 6,       // plain dependency pull (k_bwd_pull)
  KK_COPY = 7,      // calibration copy kernel
  KK_PUSH_SPLIT = 8,   // batched forward marking, split vertices (k_bpush_split)
  KK_SIGMA_SPLIT = 9,  // batched path-count pull, split vertices (k_bsigma_split)
  KK_DEP_SPLIT = 10,   // batched dependency pull, split vertices (k_bdep_split)
  KK_DEP_COMBINE = 11, // batched dependency chunk combine (k_bdep_combine)
  KK_FWD_HUBS = 12,    // plain forward, hub chunks, either direction (k_fwd_hub_level)
  KK_BWD_HUBS = 13,    // plain dependency pull, hub chunks (k_bwd_hubs)
  KK_PULL = 14,        // batched bottom-up level (k_bpull)
  KK_PULL_SPLIT = 15,  // batched bottom-up level, split vertices (k_bpull_split)
  KK_WGSWEEP = 16,     // workgroup sweep, one workgroup per root (k_wgsweep)
  KK_FWD_PULL = 17,    // reserved: the plain level kernels (kinds 5, 12) run bottom-up levels too
  KK_FWD_PULL_HUBS = 18,  // reserved, as 17
  KK_READ = 19,        // calibration read kernel
  KK_GATHER = 20,      // calibration random-line gather kernel
  KK_COUNT = 21
};
constexpr int kKbSlots = 64, kKbStride = 16;

struct KernelTimer {
  bool on = false;
  uint32_t mask = ~0u;  // kinds timed while on (bit k = kind k)
  // device: per kind kKbSlots counters kKbStride words apart (one 128-B line
  // each), summed when read: one counter per kind would serialise the
  // adds of every wave of a launch
  unsigned long long *dbytes = nullptr;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> pool;
  double ms[KK_COUNT] = {0};
  uint64_t launches[KK_COUNT] = {0};
  // each timed launch as [start, end] ms after `ref` (recorded when timing
  // is enabled): launches on two streams overlap, and the union of a kind's
  // spans is the time any of its launches ran
  hipEvent_t ref = nullptr;
  std::vector<std::pair<double, double>> spans[KK_COUNT];
};

thread_local extern char g_err[512];
int fail(int code, const char *fmt, const char *a = "", int line = 0);
// AXBC_EBUDGET (with a message) when `more` device bytes would take the
// context past its budget (none when the budget is 0).
int budget_check(axbc_ctx *c, size_t more, const char *what);

#define HIP_TRY(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return ::axbc::fail(AXBC_EHIP, "HIP error %s (line %d)", hipGetErrorString(e_), __LINE__); \
  } while (0)

}  // namespace axbc

struct axbc_ctx {
  int device = 0;
  uint32_t n = 0;
  uint64_t ned = 0;
  int heuristic = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t bytes = 0;
  // graph
  uint32_t *col = nullptr, *row = nullptr, *reach = nullptr, *rlist = nullptr;
  uint32_t *rcol = nullptr, *rrow = nullptr;  // transpose (pull direction); alias when symmetric
  bool symmetric = true;
  // every pull list strictly increasing (no repeated edge): a level-1 vertex
  // has exactly one edge from each root that reaches it
  bool pull_strict = false;
  uint32_t nr = 0;
  std::vector<uint32_t> h_reach;
  std::vector<uint32_t> h_col;
  // state of the plain rounds
  int *lvl = nullptr;
  uint32_t *sigma = nullptr;
  double *delta = nullptr, *bc = nullptr;
  uint32_t *queue = nullptr, *qoff = nullptr, *hubs = nullptr;
  uint64_t *hub_base = nullptr;
  double *hub_part = nullptr;
  uint32_t nhub_cap = 0;
  uint64_t hub_items = 0;
  axbc::DevCtl *ctl = nullptr;
  axbc::DevCtl *h_ctl = nullptr;  // pinned mirror
  uint32_t *check = nullptr;
  uint32_t ncheck = 0;
  // round bookkeeping
  int lev_guess = 8;
  bool have_round = false;
  bool dirty_delta = false;  // delta holds a batch root's values: full reset before the next round
  uint32_t last_v0 = 0;
  int last_nlev = 0;
  // samplers and batches
  axbc::SamplerBufs *sampler = nullptr;
  axbc::BatchState *batch = nullptr;
  // batches of more than 128 roots through the bit-packed sweep: the roots
  // from 128 on are swept in a second state on a second stream, beside the
  // first 128 (one host thread drives both)
  axbc::BatchState *batch2 = nullptr;
  hipStream_t stream2 = nullptr;
  hipEvent_t pair_ev[2] = {nullptr, nullptr};
  bool pair_last = false;  // the last batch ran in both states
  void *mlp = nullptr;  // learned proposal buffers (mlp.hip)
  int batch_width = 64;
  uint32_t batch_chunk = 256;
  uint32_t batch_pull16 = 32;        // bottom-up ratio x16 for batched levels (0: top-down only)
  uint32_t plain_pull16 = 256;       // the same for the levels of one-root rounds
  int mirror_state = -1;             // CSC strictly sorted and symmetric: 1 yes, 0 no, -1 not checked
  bool use_graphs = true;            // batched level loop from captured HIP graphs
  int batch_persist = 0;             // persistent level loop: 0 never (default), 1 deep graphs, 2 whenever possible
  int batch_mode = 0;                // 0 automatic, 1 bit-packed sweep only, 2 workgroup sweep only
  unsigned long long batch_persist_limit = 1000000000ull;  // barrier wait limit, 100 MHz ticks (10 s)
  uint32_t max_out = 0, max_in = 0;  // largest out- and in-degree (set with the transpose)
  size_t budget = 0;  // device bytes allowed to the context (0: no limit); checked before each allocation after create
  axbc::KernelTimer ktimer;
};

namespace axbc {
int sampler_bufs(axbc_ctx *c, SamplerBufs **out);
int batch_delta_row(axbc_ctx *c, int k, const double **row);
// Whether every list is strictly increasing and every edge has its reverse
// (then the in-edges of a vertex are its list); checked on the device once.
int mirror_symmetric(axbc_ctx *c, bool *sym);
void sampler_free(axbc_ctx *c);
void batch_free(axbc_ctx *c);
// Both batch states, the second stream and its events.
void batch_free_all(axbc_ctx *c);
void mlp_free(axbc_ctx *c);
int grid_for(uint64_t n, int cap = 2048);
// Kronecker generator on the device (kronecker.hip): nedges packed
// (min << 32 | max) edges of 2^scale vertices into edges on stream s.
int kronecker_device_edges(hipStream_t s, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                           const double initiator[4], uint64_t *edges);
// Partial rows of the batch utility (256 workgroups of 4 waves at most).
constexpr int kUtilWaves = 1024;
// Utility of the last batch's roots (axbc_batch_utility) on the given
// device buffers: dv/dw hold `count` staged targets, part kUtilWaves x B
// partial rows, dout B values.
int batch_utility_on(axbc_ctx *c, uint32_t *dv, double *dw, double *part, double *dout, const uint32_t *vertex,
                     const double *weight, uint32_t count, double *out);
// The same for every root of the last batch (out: the context's width),
// also after a workgroup sweep of more than 128 roots.
int batch_utility_any(axbc_ctx *c, uint32_t *dv, double *dw, double *part, double *dout, const uint32_t *vertex,
                      const double *weight, uint32_t count, double *out);
// Workgroup sweep (wgsweep.hip): one workgroup per root of the batch, the
// whole round of each root in one launch, per-root state in the wst / wdl /
// wlst / wloff arrays of n (wloff: lcap) entries per lane.
hipError_t wg_sweep(hipStream_t s, int nb, uint32_t n, const uint32_t *col, const uint32_t *row,
                    const uint32_t *reach, const uint32_t *src, uint32_t tag0, uint64_t *pst, double *pdl,
                    uint32_t *plst, uint32_t *ploff, uint32_t lcap, unsigned long long *cnt, int32_t *nlev,
                    uint32_t *ovf, unsigned long long *kb, uint64_t ned);
// The batch's root-minor delta rows from sweep lanes lane0 .. lane0+nb-1
// (row lane l: delta where that root reached the vertex, reach elsewhere and
// for l >= nb) and seen = stamp where one of them did.
hipError_t wg_out(hipStream_t s, int nw, uint32_t n, int nb, int lane0, uint32_t tag0, const uint64_t *pst,
                  const double *pdl, const uint32_t *reach, double *dlt, uint32_t *seen, uint32_t stamp);
// One sweep lane's delta (reach where its root did not reach).
hipError_t wg_delta(hipStream_t s, uint32_t n, uint32_t tag0, const uint64_t *st, const double *dl,
                    const uint32_t *reach, double *out);
// One lane's levels and path counts from its state words.
hipError_t wg_levels(hipStream_t s, uint32_t n, uint32_t tag0, const uint64_t *st, int32_t *lvl, uint32_t *sigma);
// Event pair around one launch of kind k when instrumentation is on.
void kt_begin(axbc_ctx *c, int kind);
void kt_end(axbc_ctx *c, int kind);
// Device counter for kind k, or nullptr when instrumentation is off.
unsigned long long *kt_bytes(axbc_ctx *c, int kind);
}  // namespace axbc
