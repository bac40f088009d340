/*
 * axbc.h -- C ABI of the MI355X (gfx950) sampled-Brandes engine,
 * libaxbc_hip.so.
 *
 * The engine replaces the device side of AxBC's 1x1 "mono" path, i.e. the
 * entry points src/bc2d.h:136-190 declares and src/cuda_kernels.cu defines.
 * Plain C types only: pointers and sizes, no HIP or torch types.
 *
 * Design differences from the reference boundary, all deliberate:
 *  - Device state lives in a context (one resident graph per context, any
 *    number of contexts per process) instead of file-static globals
 *    (src/cuda_kernels.cu:41-80).  The reference names are provided on top
 *    of a default context by liblegacy (see axbc_legacy.h).
 *  - delta and bc are FP64 (the reference keeps float).
 *  - Functions return a status instead of exiting (MY_CUDA_CHECK,
 *    src/cudamacro.h:1-16); axbc_last_error() gives the message.
 *  - axbc_pivot() runs a whole Brandes round on the device with no host
 *    synchronisation per BFS level.
 *
 * Threading: a context is used by one host thread at a time; all work of a
 * context is ordered on the context's own stream.
 */
#ifndef AXBC_H
#define AXBC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXBC_OK 0
#define AXBC_EINVAL (-1) /* bad argument                      */
#define AXBC_ENOMEM (-2) /* device or host allocation failed  */
#define AXBC_EHIP (-3)   /* HIP runtime error                 */
#define AXBC_ESTATE (-4) /* call out of order                 */
#define AXBC_EBUDGET (-5)/* request exceeds the memory budget */

typedef struct axbc_ctx axbc_ctx;

/* Counters of the last pivot. */
typedef struct {
  uint64_t nvisited; /* vertices reached, root included                */
  uint64_t nlevels;  /* deepest non-empty BFS level                    */
  uint64_t e_r;      /* adjacency entries of the reached vertices      */
  uint64_t e_dag;    /* shortest-path DAG edges (sigma contributions)  */
  uint64_t all;      /* reach sum used by the bc update (-H 1)         */
} axbc_pivot_stats;

/* Message of the last failing call on this thread. */
const char *axbc_last_error(void);

/* Number of visible HIP devices (0 when none). */
int axbc_device_count(void);

/*
 * Upload a CSC graph and allocate the per-vertex state.
 * Replaces initcuda() (src/cuda_kernels.cu:1558-1656) + set_degree +
 * init_bc_1degree_device() (cu:1515-1519).
 *   N            vertices; col[N+1], row[ned] uint32 CSC (column = the
 *                second endpoint, rows sorted, as build_csc makes it,
 *                src/axbc2d.c:883-907)
 *   reach        N pendant counts from the 1-degree reduction, or NULL
 *   heuristic    -H value (0 or 1): 1 adds the reach sum of the visited
 *                vertices to nvisited in the bc update
 * *dev_bytes (may be NULL) receives the device bytes allocated.
 */
int axbc_ctx_create(int device, uint64_t N, uint64_t ned, const uint32_t *col,
                    const uint32_t *row, const uint32_t *reach, int heuristic,
                    axbc_ctx **out, size_t *dev_bytes);
/* Replace the pendant counts (init_bc_1degree_device, cu:1515-1519) and
 * reset the round state; bc is cleared too. */
int axbc_set_reach(axbc_ctx *ctx, const uint32_t *reach);
/* Pinned host memory (CudaMallocHostSet / CudaFreeHost, cu:825-846). */
void *axbc_host_alloc(size_t bytes);
void axbc_host_free(void *p);
/* Device memory of a context: the bytes it holds now, and a budget for
 * what it allocates after creation (batch state and its list growth,
 * sampler and LCC tables, MLP buffers, the transposed graph): an
 * allocation that would take it past the budget fails with AXBC_EBUDGET
 * and leaves the context usable; 0 removes the limit.  No reference
 * counterpart (the serving loop admits jobs by device footprint). */
int axbc_device_bytes(axbc_ctx *ctx, size_t *bytes);
int axbc_set_memory_budget(axbc_ctx *ctx, size_t bytes);
/* Device bytes of the batch state that the first batch of a context on
 * this graph (n vertices, ned CSC entries, column offsets col[n+1]) creates
 * at width 64 or 128 (256: the two states of a batch of more than 128
 * roots), with the level list at its first capacity; a graph
 * that is not symmetric adds its transpose, 4 (n + 1 + ned) bytes.  For
 * admission before a context exists.  No reference counterpart. */
int axbc_batch_state_bytes(uint64_t n, uint64_t ned, const uint32_t *col, int width, size_t *bytes);
/* Release everything (fincuda(), cu:1659-1693). */
int axbc_ctx_destroy(axbc_ctx *ctx);

/*
 * One sampled-Brandes round from root v0 with sampling probability p:
 * reset of the previous round's visited state (setcuda, cu:1532-1554),
 * root (set_mlp_cuda, cu:862-876), the forward levels
 * (scan_col_csc_cuda_mono, cu:1088-1177), the backward depths
 * (scan_frt_csc_cuda_mono, cu:1013-1080), the reach sum
 * (pre_update_bc_cuda, cu:358-373) and the bc update (update_bc_cuda,
 * cu:289-302), as bc_func_mono (src/axbc2d.c:1514-1659) sequences them.
 * stats may be NULL; when given the call waits for the round to finish.
 */
int axbc_pivot(axbc_ctx *ctx, uint32_t v0, double p, axbc_pivot_stats *stats);

/* Wait for all queued work of the context. */
int axbc_sync(axbc_ctx *ctx);

/* bc := 0. */
int axbc_reset_bc(axbc_ctx *ctx);

/* Device -> host copies of the state after the last round.
 * get_bc / get_delta of the reference (cu:1466-1473) at FP64, get_deg
 * (cu:1458), and the per-vertex levels / path counts for parity checks. */
int axbc_get_bc(axbc_ctx *ctx, double *bc);
int axbc_get_delta(axbc_ctx *ctx, double *delta);
int axbc_get_deg(axbc_ctx *ctx, uint32_t *deg);
int axbc_get_lvl(axbc_ctx *ctx, int32_t *lvl);
int axbc_get_sigma(axbc_ctx *ctx, uint32_t *sigma);
int axbc_set_bc(axbc_ctx *ctx, const double *bc);

/*
 * Adaptive stop test of the pivot loop (src/axbc2d.c:2846-2870) on the
 * device: *stop = 1 when bc[a]/2 >= c*N for every check vertex a (ids >= N
 * read as bc 0).  The check list is uploaded once with axbc_set_checks.
 */
int axbc_set_checks(axbc_ctx *ctx, const uint32_t *check, uint32_t x);
int axbc_stop_test(axbc_ctx *ctx, uint32_t c, int *stop);

/* Fine-grained per-level entry points with the reference's call structure
 * (each call synchronises, as the reference's wrappers do).  They drive
 * the same kernels as axbc_pivot; used by the legacy names.
 *   axbc_begin_round   setcuda()          cu:1532
 *   axbc_set_root      set_mlp_cuda()     cu:862
 *   axbc_forward_level scan_col_csc_cuda_mono(ncol, level) -> nfrt
 *   axbc_backward_depth scan_frt_csc_cuda_mono(offset, ncol, depth)
 *   axbc_reach_sum     pre_update_bc_cuda() -> all (lvl > 0 only)
 *   axbc_update_bc     update_bc_cuda(v0, p, ncol, nvisited)           */
int axbc_begin_round(axbc_ctx *ctx);
int axbc_set_root(axbc_ctx *ctx, uint32_t v0);
int axbc_forward_level(axbc_ctx *ctx, int level, uint32_t *nfrt);
int axbc_backward_depth(axbc_ctx *ctx, int depth);
int axbc_reach_sum(axbc_ctx *ctx, uint32_t *all);
int axbc_update_bc(axbc_ctx *ctx, uint32_t v0, double p, uint64_t nvisited);

/*
 * Pivot samplers on the device.
 *
 * High-degree-neighbourhood tables (-z 1; src/axbc2d.c:1758-1845,
 * 2758-2786): builds on the device the wrapping uint32 prefix of
 * (uint32)deg^2, its running maximum and the neighbour-degree sums; *sum is
 * the wrapped total (sum_deg).  pmax_host / sdn_host (N entries each, may
 * be NULL) receive copies.  axbc_hdn_pick returns the first vertex whose
 * wrapped prefix reaches r, as findsample_LOCINT (axbc2d.c:2187-2198) with
 * its "N-1 when none" fallback, and optionally its neighbour-degree sum.
 */
int axbc_hdn_build(axbc_ctx *ctx, uint32_t *pmax_host, uint32_t *sdn_host, uint32_t *sum);
int axbc_hdn_pick(axbc_ctx *ctx, uint32_t r, uint32_t *u, uint32_t *sdn_u);

/*
 * Dynamic distributions (-z 3 bc, 4 delta of the last round, 5 1/(1+bc);
 * src/axbc2d.c:2929-3060) as 64-bit fixed-point weights
 * W[v] = floor(w[v] * 2^s) with one shift s for the array, scanned on the
 * device; *total = sum W.  axbc_weights_pick returns the first v whose
 * prefix reaches r and p = W[v] / total.
 */
int axbc_weights(axbc_ctx *ctx, int kind, uint64_t *total);
int axbc_weights_pick(axbc_ctx *ctx, uint64_t r, uint32_t *v, double *p);
/* W[v] of the last axbc_weights call (mixture probabilities, -z 6);
 * axbc_weights_copy copies all N of them (batch-synchronous -z 6 freezes
 * them at the batch start). */
int axbc_weights_at(axbc_ctx *ctx, uint32_t v, uint64_t *w);
int axbc_weights_copy(axbc_ctx *ctx, uint64_t *w);

/*
 * Local clustering coefficients (lcc_func, src/axbc2d.c:1847-1935) on the
 * device: float lcc[v] = counter / (uint32)(d*(d-1)) with the wrapping
 * uint32 count of (j, x) pairs, j in adj(v), x in adj(j) cap adj(v), x != v;
 * then the static LCC distribution (-z 2) as fixed-point weights and their
 * prefix.  Host copies (N entries each) may be NULL.
 */
int axbc_lcc_build(axbc_ctx *ctx, float *lcc_host, uint64_t *lw_host, uint64_t *lp_host, uint64_t *total);
/*
 * Draws of the static distributions searched on the device, a window at a
 * time (findsample_LOCINT, src/axbc2d.c:2188-2198, over the -z 1 table of
 * 2758-2786; the -z 2 LCC draw of 2872-2894): for the next `count` values
 * of the lrand48 stream with 48-bit state `state` (X; a value advances
 * X = (0x5DEECE66D X + 0xB) mod 2^48 and is X >> 17), lr[p] is the value
 * at position p and v[p] the vertex that a first draw there picks -- table
 * 1: first i with the running maximum of the wrapped (uint32)deg^2 prefix
 * >= lr % (total + 1), total the deg^2 sum (axbc_hdn_build); table 2: first
 * i with the LCC weight prefix >= lr / 2147483647 * total (clamped), total
 * from axbc_lcc_build; N-1 when none.  A host loop consumes its own lrand48
 * values, checks them against lr, and takes v where a first draw falls.
 * Runs on a stream of its own, so a window can be searched while a batch
 * runs.
 */
int axbc_draw_window(axbc_ctx *ctx, int table, uint64_t state, uint64_t total, uint32_t count, uint32_t *lr,
                     uint32_t *v);

/*
 * Multi-source rounds: nb (1..width) roots in one sweep, width 64 or 128
 * (axbc_set_batch_width; default 64), or 256: as a workgroup sweep of all
 * the roots, or as two 128-lane bit-packed sweeps side by side (roots
 * 0..127 on the context stream, the others in a second state on a second
 * stream), the roots applied in two halves of draw order.  Results are those of nb
 * consecutive axbc_pivot calls in the given order: the bc update applies
 * the roots' contributions vertex by vertex in root order.  With stop_c > 0
 * and a check list (axbc_set_checks), the adaptive stop test
 * (src/axbc2d.c:2846-2870) is evaluated after every root of the batch and
 * only the roots up to the first one after which it holds are applied;
 * *applied receives their count.  stats (nb entries) may be NULL.
 * axbc_batch_delta copies root k's delta of the last batch.
 * No reference counterpart: the reference runs one root per round.
 */
int axbc_set_batch_width(axbc_ctx *ctx, int width);
/* Edges per work item for vertices split across waves (64 .. 2^20;
 * default 256): the degree bucket boundary the autotuner sets. */
int axbc_set_batch_chunk(axbc_ctx *ctx, uint32_t chunk);
/* Direction of the batched levels: a level runs bottom-up (every vertex
 * not yet reached by all roots pulls over its in-edges) when those
 * vertices' in-edges are at most ratio16/16 times the out-edges the
 * top-down push would scan; 0 keeps every level top-down.  Default 32.
 * Results do not depend on it. */
int axbc_set_batch_pull(axbc_ctx *ctx, uint32_t ratio16);
/* The same switch for the levels of one-root rounds (axbc_pivot): a level
 * runs bottom-up (every vertex not yet found sums its neighbours' path counts
 * at the level before) when the lists of the vertices not yet found hold at
 * most ratio16/16 times the edges of the vertices the last level found;
 * only on graphs whose every list is strictly increasing and holds each
 * edge's reverse (checked once on the device); 0 keeps every level
 * top-down, as do the per-level entry points.  Default 256 (SCALE 22,
 * high-degree-neighbourhood roots: 155 pivots/s at 32, 173 at 128, 176 at
 * 256, 177 at 1024).  Results do not depend on it. */
int axbc_set_plain_pull(axbc_ctx *ctx, uint32_t ratio16);
/* Level loop of the batched sweep: 1 (default) replays prebuilt HIP graphs
 * of 8 levels (depths) each, the kernels reading the level from device
 * memory, so a level costs no host launches; 0 launches every kernel from
 * the host.  Kernel timing (axbc_kernel_timing) always uses host launches.
 * Results do not depend on it. */
int axbc_set_batch_graphs(axbc_ctx *ctx, int enable);
/* Persistent level loop: on graphs without split vertices (every list
 * below the chunk) whose previous batch ran more than 64 levels (mode 1,
 * default), or on every such graph (mode 2), each direction of the sweep
 * runs in one cooperative launch whose workgroups pass the levels (depths)
 * through grid barriers, instead of several launches per level; mode 0
 * never.  Every barrier wait is bounded by limit_seconds (default 10): a
 * longer wait ends the launch in every workgroup and the batch fails with
 * AXBC_EHIP, bc unchanged and the batch state rebuilt by the next batch.
 * Results do not depend on it. */
int axbc_set_batch_persistent(axbc_ctx *ctx, int mode, double limit_seconds);
/* How a batch sweeps: 1 the bit-packed sweep above; 2 one workgroup per root
 * (the whole round of each root in one launch, levels separated by
 * workgroup barriers, per-root state in HBM: 24 bytes per vertex and lane
 * plus 4 per level and lane, allocated on first use); 0 (default) chooses 2
 * on graphs whose previous batch ran more than 256 levels and whose lists
 * hold at most 64 edges (high-diameter graphs, where the roots' frontiers
 * rarely meet), 1 otherwise.  Results do not depend on it. */
int axbc_set_batch_mode(axbc_ctx *ctx, int mode);
int axbc_batch_width(axbc_ctx *ctx);
int axbc_batch(axbc_ctx *ctx, const uint32_t *v0s, const double *ps, int nb, uint32_t stop_c,
               int *applied, axbc_pivot_stats *stats);
int axbc_batch_delta(axbc_ctx *ctx, int k, double *delta);
/* Root k's BFS levels (-1 unreached) and path counts (uint32, wrapping) of
 * the last batch, from its level list: what axbc_get_lvl / axbc_get_sigma
 * give after a one-root round (the per-pivot state of scan_col_mono2,
 * src/cuda_kernels.cu:588-658).  For parity checks of the batched sweep. */
int axbc_batch_levels(axbc_ctx *ctx, int k, int32_t *lvl, uint32_t *sigma);
/* Utility of the last batch's roots for a set of target vertices:
 * out[k] = sum_i weight[i] * delta_k(vertex[i]) for k < width (FP64; the
 * learned proposal's training target), summed in a fixed order (the same
 * inputs give the same bits).  No reference counterpart. */
int axbc_batch_utility(axbc_ctx *ctx, const uint32_t *vertex, const double *weight, uint32_t count, double *out);
/* axbc_batch_utility without any device allocation, for callers that
 * cannot allocate (code under stream capture, tight serving loops, hosts
 * that forbid hipMalloc).  No counterpart in the reference.  Same result,
 * bit for bit, as axbc_batch_utility on the same inputs.  Its staging and
 * partial rows live in a fixed device array per GPU (about 1.8 MB, held
 * until the process ends), so count is at most AXBC_NOALLOC_UTILITY_MAX
 * (AXBC_EINVAL beyond).  Calls that overlap on one GPU share that array and
 * race: serialise them. */
#define AXBC_NOALLOC_UTILITY_MAX 65536u
int axbc_batch_utility_noalloc(axbc_ctx *ctx, const uint32_t *vertex, const double *weight, uint32_t count,
                               double *out);

/*
 * Per-kernel instrumentation for roofline figures (no reference
 * counterpart).  While enabled, every launch of the main kernels is
 * bracketed by HIP events on the context stream and the kernels add their
 * algorithmic byte counts to device counters.  Kinds: 0 batched push,
 * 1 batched path-count pull, 2 batched level close, 3 batched dependency
 * pull, 4 batched bc update, 5 plain forward level (top-down or
 * bottom-up), 6 plain dependency pull,
 * 7 calibration copy, 8 batched push of split vertices, 9 batched
 * path-count pull of split vertices, 10 batched dependency pull of split
 * vertices, 11 their chunk combine, 12 plain forward hub chunks, 13 plain
 * dependency pull hub chunks, 14 batched bottom-up level, 15 its split
 * vertices, 16 workgroup sweep, 17 and 18 reserved (not launched), 19
 * calibration read, 20 calibration gather.  axbc_kernel_stats
 * returns the summed milliseconds,
 * the launch count and the byte count of one kind since enabling.
 * axbc_kernel_timing_kinds enables the same for the kinds whose bit is set
 * in `mask` only (0 disables), so that a timed run carries the events of
 * the kernel it measures and no others.
 * axbc_copy_kernel runs a plain device copy of `bytes` bytes (kind 7) to
 * calibrate hardware counters against a known traffic; axbc_read_kernel
 * streams `bytes` bytes without writing them back (kind 19): the attainable
 * read rate, the bound of the read-dominated frontier kernels;
 * axbc_gather_kernel reads one double from each of 134 million random
 * 128-byte lines of a `window`-byte buffer per launch (kind 20, counting
 * 128 bytes per line): the attainable rate of scattered line fetches over a
 * region of that size, the bound of the kernels that gather rows at
 * scattered vertices.
 */
int axbc_kernel_timing(axbc_ctx *ctx, int enable);
int axbc_kernel_timing_kinds(axbc_ctx *ctx, uint32_t mask);
int axbc_kernel_stats(axbc_ctx *ctx, int kind, double *ms, uint64_t *launches, uint64_t *bytes);
/* The milliseconds during which at least one timed launch of a kind ran
 * since enabling (launches on the two streams of a width-256 batch overlap,
 * so this can be less than axbc_kernel_stats' sum). */
int axbc_kernel_busy(axbc_ctx *ctx, int kind, double *ms);
int axbc_copy_kernel(axbc_ctx *ctx, uint64_t bytes, int repeats);
int axbc_read_kernel(axbc_ctx *ctx, uint64_t bytes, int repeats);
int axbc_gather_kernel(axbc_ctx *ctx, uint64_t window, int repeats);

/*
 * Graph construction on the device (no reference device counterpart; the
 * host steps are src/axbc2d.c:333-351 mirror, 818-841 norm_graph, 695-807
 * and 866-876 the 1-degree reduction, 883-907 build_csc).  edges: ned
 * packed keys (u << 32 | v), endpoints < N.  undirected: add the reversed
 * non-loop edges; heuristic 1: normalise + 1-degree reduction (reach[]
 * counts), 0: none; then normalise and build the CSC.  col: N + 1
 * entries, row: row_cap entries (at least the final edge count; 2 * ned
 * always suffices), reach: N entries.  Results equal the host pipeline's
 * bit for bit.
 */
int axbc_ingest(int device, const uint64_t *edges, uint64_t ned, uint64_t N, int undirected, int heuristic,
                uint32_t *col, uint32_t *row, uint64_t row_cap, uint32_t *reach, uint64_t *ned_out,
                uint32_t *rem_ed);

/*
 * Kronecker generation on the device (no reference device counterpart; the
 * host generator is generator/make_graph.c:271-353 as called from
 * src/axbc2d.c:518-527).  axbc_kronecker_generate writes the nedges edges
 * of 2^scale vertices (scale 1..30) for the seed pair and the initiator
 * (A, B, C, D) to edges[] as packed keys (min << 32 | max), in the same
 * positions as the host generator's output.  axbc_kronecker_ingest
 * generates on the device and runs axbc_ingest's preprocessing on the
 * result without a host round trip (same outputs as axbc_ingest).
 */
int axbc_kronecker_generate(int device, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                            const double initiator[4], uint64_t *edges);
int axbc_kronecker_ingest(int device, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                          const double initiator[4], int undirected, int heuristic, uint32_t *col, uint32_t *row,
                          uint64_t row_cap, uint32_t *reach, uint64_t *ned_out, uint32_t *rem_ed);

/*
 * Learned pivot proposal (no reference counterpart; stands in for the
 * bc/delta heuristics -z 3/4): a 4 -> 64 -> 64 -> 1 MLP with ReLU, BF16
 * weights and FP32 accumulation, over per-vertex features (log(1+deg), LCC
 * or target adjacency, log(1+bc*bc_scale), log(1+mean neighbour
 * bc*bc_scale)).
 *   axbc_mlp_init        FP32 master parameters (axbc_mlp_param_count()
 *                        floats: W1[64][4], b1, W2[64][64], b2, w3, b3)
 *   axbc_mlp_features    features from the device state (LCC when built)
 *   axbc_mlp_features_spec  the inputs log(1+deg), LCC (built if needed),
 *                        log(1+bc*bc_scale) and the last dependency values
 *                        log(1+delta_k(v)*delta_scale), delta_k being root
 *                        k's row of the last batch
 *   axbc_mlp_targets     target vertices and weights: while set (count > 0),
 *                        feature 1 is log(1 + t(v) / mean positive t) with
 *                        t(v) the sum of the weights of v's neighbours
 *                        instead of the LCC; count 0 clears them
 *   axbc_mlp_set_data    host features / targets (either may be NULL)
 *   axbc_mlp_gradient    MSE gradient over the sample ids (accumulated on the
 *                        device; copied out when grad_host is given)
 *   axbc_mlp_adam        one Adam step with the accumulated gradient
 *   axbc_mlp_forward     scores of all vertices
 *   axbc_mlp_proposal    proposal weights expm1(max(score, 0)) + eps * mean
 *                        (vertices of positive degree; 0 for the others) as the
 *                        fixed-point distribution of axbc_weights_pick
 * axbc_weights_from      fixed-point distribution from any device array of N
 *                        non-negative doubles.
 */
int axbc_mlp_param_count(void);
int axbc_mlp_init(axbc_ctx *ctx, const float *params);
int axbc_mlp_features(axbc_ctx *ctx, double bc_scale);
int axbc_mlp_features_spec(axbc_ctx *ctx, double bc_scale, int k, double delta_scale);
int axbc_mlp_targets(axbc_ctx *ctx, const uint32_t *vertex, const double *weight, uint32_t count);
int axbc_mlp_set_data(axbc_ctx *ctx, const float *x, const float *y);
int axbc_mlp_get_data(axbc_ctx *ctx, float *x, float *score);
int axbc_mlp_gradient(axbc_ctx *ctx, const uint32_t *idx, uint32_t m, float *grad_host, float *loss_host);
int axbc_mlp_adam(axbc_ctx *ctx, float lr);
int axbc_mlp_params(axbc_ctx *ctx, float *master_host);
int axbc_mlp_forward(axbc_ctx *ctx);
int axbc_mlp_proposal(axbc_ctx *ctx, double eps, uint64_t *total);
int axbc_weights_from(axbc_ctx *ctx, const double *w_device, uint64_t *total);

/* Last round's counters (waits for the context). */
int axbc_last_stats(axbc_ctx *ctx, axbc_pivot_stats *stats);

/* Time (ms) of the last axbc_pivot measured with HIP events on the context
 * stream, split into forward, backward and update phases. */
int axbc_last_timing(axbc_ctx *ctx, float *fwd_ms, float *bwd_ms, float *upd_ms);

#ifdef __cplusplus
}
#endif

#endif /* AXBC_H */
