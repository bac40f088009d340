/*
 * CPU oracle for the sampled-Brandes hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of AxBC's 1x1 "mono" path, used by tests/, by the
 * smoke check and by the CPU-baseline leg of bench.py as the checker.  The
 * engine never links, loads or calls it.
 *
 * Types follow the reference where they are exact (vertex ids, degrees,
 * sigma and reach as uint32 with wrap-around, levels as int32 with -1 for
 * unvisited) and widen delta/bc from float to double.
 */
#ifndef AXBC_ORACLE_H
#define AXBC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-pivot counters. */
typedef struct {
  uint64_t nvisited; /* vertices reached, root included */
  uint64_t nlevels;  /* deepest non-empty level (0 when only the root) */
  uint64_t e_r;      /* adjacency entries of reached vertices */
  uint64_t e_dag;    /* (u,w) with w in adj(u), lvl(w) = lvl(u)+1 */
  uint64_t all;      /* reach sum handed to the bc update (incl. reach[v0]) */
} oracle_pivot_stats;

/* One Brandes round from v0 with sampling probability p; bc is updated in
 * place (accumulated), lvl/sigma/delta are overwritten for all N vertices.
 * Restates setcuda + set_mlp_cuda + bc_func_mono + scan_col_mono2 +
 * scan_frt_mono + pre_update_bc_cuda + update_bc. */
void oracle_pivot(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
                  int heuristic, uint32_t v0, double p, int32_t *lvl, uint32_t *sigma, double *delta,
                  double *bc, oracle_pivot_stats *st);

/* The rounds of `count` roots over `nthreads` OpenMP threads (independent
 * rounds of a static distribution, no stop test); bc += their sum.
 * Returns the summed E_r.  Throughput baseline only. */
uint64_t oracle_pivots_omp(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
                           int heuristic, const uint32_t *roots, const double *probs, uint32_t count,
                           int nthreads, double *bc);

/* Pivot-loop configuration (the CLI flags of the 1x1 path). */
typedef struct {
  uint64_t nbfs;        /* -N */
  uint32_t bader;       /* -c */
  uint32_t approx_n;    /* -x */
  int distribution;     /* -z: 0..6 */
  int heuristic;        /* -H */
  int random_seed_mode; /* -R: 0 fixed seed */
  int64_t seed;         /* seed for srand48 when random_seed_mode == 0 */
  uint32_t sync_batch;  /* > 0: dynamic distributions (-z 3/4/5) frozen for
                           every group of sync_batch rounds (batch-synchronous
                           sampling); 0: the reference's per-round update */
} oracle_cfg;

typedef struct {
  uint32_t nrounds, skipped;
  int stopped;          /* left through the adaptive stop test */
  uint64_t attempts;    /* loop iterations consumed */
} oracle_run_stats;

/* Full pivot loop of main() (src/axbc2d.c:2700-3156) for the supported
 * distributions.  Writes the pivot of each round to pivots[] and its
 * probability to probs[] (arrays of length nbfs) and the final bc (length
 * N, accumulated, NOT divided).  Returns 0, or -1 for an unsupported cfg. */
int oracle_run(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
               const oracle_cfg *cfg, uint32_t *pivots, double *probs, double *bc,
               oracle_run_stats *rs);

/* The draws of oracle_run without the rounds: the pivot sequence, the
 * probabilities and the skip count of a static distribution (-z 0, 1, 2)
 * with the stop test off (cfg->bader == 0), which do not depend on bc.
 * Returns -1 for any other cfg. */
int oracle_draws(uint32_t N, const uint32_t *col, const uint32_t *row, const oracle_cfg *cfg,
                 uint32_t *pivots, double *probs, oracle_run_stats *rs);

/* Local clustering coefficient of every vertex as the reference's
 * lcc_func computes it (float). */
void oracle_lcc(uint32_t N, const uint32_t *col, const uint32_t *row, float *lcc);

/* Sampling helpers shared by the loop (exposed for the tests). */
uint32_t oracle_lim_x_sample(uint64_t N);
uint32_t oracle_hdn_sample(uint32_t N, const uint32_t *col, const uint32_t *row, uint32_t r,
                           double *p_out);
uint64_t oracle_fixed_weights(uint32_t N, const double *w, uint64_t *W);
uint32_t oracle_weighted_sample(uint32_t N, const uint64_t *W, uint64_t total, long draw,
                                double *p_out);
/* srand48 / lrand48 on the calling thread's own generator state (the
 * loops seed it with cfg->seed). */
void oracle_srand48(long seed);
long oracle_lrand48(void);

#ifdef __cplusplus
}
#endif

#endif
