/*
 * CPU ORACLE -- test infrastructure only (tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it as the checker; the engine never
 * does).  A sequential plain-C restatement of AxBC's sampled-Brandes path at
 * -p 1x1 ("mono" mode), written from the reference sources:
 *
 *   per pivot  src/axbc2d.c:1514-1659 (bc_func_mono)
 *              src/cuda_kernels.cu:1532-1554 (setcuda), 1521-1530
 *              (init_delta), 862-876 (set_mlp_cuda), 1088-1177 +
 *              588-658 (forward level: scan_col_csc_cuda_mono,
 *              scan_col_mono2), 1013-1080 + 719-772 (backward depth:
 *              scan_frt_csc_cuda_mono, scan_frt_mono), 339-373
 *              (pre_update_bc_cuda), 269-302 (update_bc)
 *   pivot loop src/axbc2d.c:2700-3156, samplers 1758-1845, 1936-2211
 *
 * Arithmetic: sigma is uint32 and wraps modulo 2^32 as the reference's
 * atomicAdd does; levels are int32 (-1 unvisited, 0 root); delta and bc are
 * FP64 (the reference's float, widened).  Every delta sum is a left fold in
 * adjacency order starting from reach[v].
 *
 * Defined behaviour where the reference is undefined (documented in the
 * design notes): a drawn vertex id equal to N (uniform_int(N,0) can return
 * N) counts as a degree-0 vertex and is never marked; a check vertex equal
 * to N reads bc = 0; uniform_int over an empty modulus (sum+1 == 0) uses
 * the raw lrand48 value; no level cap.
 *
 * Dynamic distributions (-z 3/4/5) sample from 64-bit fixed-point weights
 * W[v] = floor(w[v] * 2^s) with one shift s for the whole array, so that the
 * prefix sums are exact in any order; the engine's GPU sampler uses the
 * same definition.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

void oracle_pivot(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
                  int heuristic, uint32_t v0, double p, int32_t *lvl, uint32_t *sigma, double *delta,
                  double *bc, oracle_pivot_stats *st) {
  /* setcuda: lvl = -1, sigma = 0, delta = reach (cu:1532-1554, 1521) */
  for (uint32_t v = 0; v < N; ++v) {
    lvl[v] = -1;
    sigma[v] = 0;
    delta[v] = (double)reach[v];
  }
  uint32_t *queue = (uint32_t *)xcalloc(N, sizeof(uint32_t));
  uint64_t *off = (uint64_t *)xcalloc((size_t)N + 2, sizeof(uint64_t));
  /* set_mlp_cuda(v0, 0, 1): root at level 0 with one shortest path */
  lvl[v0] = 0;
  sigma[v0] = 1;
  queue[0] = v0;
  off[0] = 0;
  off[1] = 1;
  uint64_t tail = 1, e_dag = 0;
  int level;
  /* Forward: the frontier of level L-1 discovers level L; a neighbour w
   * that is unvisited or already discovered at L gains sigma[u]
   * (scan_col_mono2, cu:640-655). */
  for (level = 1;; ++level) {
    uint64_t b = off[level - 1], e = off[level];
    for (uint64_t i = b; i < e; ++i) {
      uint32_t u = queue[i];
      for (uint32_t k = col[u]; k < col[u + 1]; ++k) {
        uint32_t w = row[k];
        if (lvl[w] == -1) {
          lvl[w] = level;
          queue[tail++] = w;
        }
        if (lvl[w] == level) {
          sigma[w] += sigma[u];
          ++e_dag;
        }
      }
    }
    off[level + 1] = tail;
    if (tail == e) break;
  }
  const int nlev = level - 1; /* deepest non-empty level */
  /* Backward (bc_func_mono's do-while, axbc2d.c:1613-1625): depths
   * nlev-1 .. 1, or depth 0 alone when nlev == 1; nothing when nlev == 0. */
  if (nlev >= 1) {
    int depth = nlev - 1;
    do {
      for (uint64_t i = off[depth]; i < off[depth + 1]; ++i) {
        uint32_t v = queue[i];
        double acc = delta[v];
        const double sv = (double)sigma[v];
        for (uint32_t k = col[v]; k < col[v + 1]; ++k) {
          uint32_t w = row[k];
          if (lvl[w] == depth + 1) acc += ((delta[w] + 1.0) / (double)sigma[w]) * sv;
        }
        delta[v] = acc;
      }
      --depth;
    } while (depth > 0);
  }
  /* pre_update_bc_cuda only for -H 1 (axbc2d.c:1636-1639) */
  uint32_t all = 0;
  if (heuristic == 1 || heuristic == 3)
    for (uint32_t v = 0; v < N; ++v)
      if (lvl[v] > 0) all += reach[v];
  all += reach[v0];
  const uint64_t nvt = tail + (uint64_t)all;
  /* update_bc (cu:269-286) over all N vertices */
  const uint32_t r0 = reach[v0];
  const double r1 = (double)(uint32_t)(r0 + 1u);
  for (uint32_t v = 0; v < N; ++v) {
    if (v == v0) continue;
    bc[v] += delta[v] * r1 / p;
  }
  if (r0 > 0) bc[v0] += (double)((uint64_t)r0 * (nvt - 2));
  if (st) {
    uint64_t er = 0;
    for (uint64_t i = 0; i < tail; ++i) er += col[queue[i] + 1] - col[queue[i]];
    st->nvisited = tail;
    st->nlevels = (uint64_t)nlev;
    st->e_r = er;
    st->e_dag = e_dag;
    st->all = all;
  }
  free(queue);
  free(off);
}

/* Throughput form for the CPU baseline: the rounds of `count` given roots
 * (a static distribution, no stop test, so the rounds are independent)
 * spread over `nthreads` OpenMP threads, each with its own round state and
 * bc accumulator, added into bc at the end.  Returns the summed E_r. */
uint64_t oracle_pivots_omp(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
                           int heuristic, const uint32_t *roots, const double *probs, uint32_t count,
                           int nthreads, double *bc) {
  uint64_t er = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : er)
  {
    int32_t *lvl = (int32_t *)xcalloc(N, sizeof(int32_t));
    uint32_t *sigma = (uint32_t *)xcalloc(N, sizeof(uint32_t));
    double *delta = (double *)xcalloc(N, sizeof(double));
    double *mine = (double *)xcalloc(N, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
      oracle_pivot_stats st;
      oracle_pivot(N, col, row, reach, heuristic, roots[i], probs[i], lvl, sigma, delta, mine, &st);
      er += st.e_r;
    }
#pragma omp critical
    for (uint32_t v = 0; v < N; ++v) bc[v] += mine[v];
    free(lvl);
    free(sigma);
    free(delta);
    free(mine);
  }
  return er;
}

/* ---- samplers ------------------------------------------------------------ */

/* int LIMxSAMPLE = N*N (axbc2d.c:2835), compared against the uint32 BSM
 * counter, i.e. as an unsigned 32-bit value. */
uint32_t oracle_lim_x_sample(uint64_t N) { return (uint32_t)(N * N); }

/* srand48 / lrand48 (POSIX: X' = (0x5DEECE66D X + 0xB) mod 2^48, seeded
 * X = seed << 16 | 0x330E, lrand48 = X >> 17) on a per-thread state, so
 * that loops run on several threads at once each draw their own sequence;
 * nrand48 on a caller's state is the same generator. */
static __thread unsigned short rng_x[3];
void oracle_srand48(long seed) {
  rng_x[0] = 0x330E;
  rng_x[1] = (unsigned short)(seed & 0xFFFF);
  rng_x[2] = (unsigned short)((seed >> 16) & 0xFFFF);
}
long oracle_lrand48(void) { return nrand48(rng_x); }

static uint32_t uniform_int(uint32_t max, uint32_t min) {
  /* lrand48() % (max - min + 1) + min in uint32 (axbc2d.c:2156-2164) */
  uint32_t m = max - min + 1u;
  long r = oracle_lrand48();
  return m ? (uint32_t)((uint32_t)r % m) + min : (uint32_t)r + min;
}

static uint32_t select_root1(uint64_t N) { return uniform_int((uint32_t)N, 0); }

/* High-degree-neighbourhood sampling (axbc2d.c:2895-2928) with the static
 * tables of 2758-2786.  r is the drawn value in [0, sum]; the neighbour
 * draw happens inside.  Returns v0 and sets *p_out. */
typedef struct {
  uint32_t *prefix; /* inclusive uint32 prefix of (uint32)deg^2, wrapping */
  uint32_t *sdn;    /* sum of neighbour degrees, wrapping */
  uint32_t sum;
} hdn_tables;

static void hdn_build(uint32_t N, const uint32_t *col, const uint32_t *row, hdn_tables *t) {
  t->prefix = (uint32_t *)xcalloc(N, sizeof(uint32_t));
  t->sdn = (uint32_t *)xcalloc(N, sizeof(uint32_t));
  uint32_t run = 0;
  for (uint32_t v = 0; v < N; ++v) {
    uint64_t d = col[v + 1] - col[v];
    run += (uint32_t)(d * d); /* (LOCINT)pow(deg, 2) */
    t->prefix[v] = run;
    uint32_t s = 0;
    for (uint32_t k = col[v]; k < col[v + 1]; ++k) s += col[row[k] + 1] - col[row[k]];
    t->sdn[v] = s;
  }
  t->sum = N ? t->prefix[N - 1] : 0;
}

static uint32_t hdn_pick(uint32_t N, const uint32_t *col, const uint32_t *row, const hdn_tables *t,
                         uint32_t r, double *p_out) {
  uint32_t i;
  for (i = 0; i < N; ++i)
    if (t->prefix[i] >= r) break;
  uint32_t v = i < N ? i : N - 1;
  uint32_t d = col[v + 1] - col[v];
  if (d > 0) {
    uint32_t nb = d == 1 ? 0 : uniform_int(d - 1, 0);
    v = row[col[v] + nb];
  }
  *p_out = (double)t->sdn[v] / (double)t->sum;
  return v;
}

uint32_t oracle_hdn_sample(uint32_t N, const uint32_t *col, const uint32_t *row, uint32_t r,
                           double *p_out) {
  hdn_tables t;
  hdn_build(N, col, row, &t);
  uint32_t v = hdn_pick(N, col, row, &t, r, p_out);
  free(t.prefix);
  free(t.sdn);
  return v;
}

/* Fixed-point weights: s = 61 - e where wmax * N = m * 2^e, m in [0.5, 1);
 * W[v] = floor(w[v] * 2^s).  Returns the exact total.  Non-finite or
 * negative weights count as 0. */
uint64_t oracle_fixed_weights(uint32_t N, const double *w, uint64_t *W) {
  double wmax = 0;
  for (uint32_t v = 0; v < N; ++v)
    if (isfinite(w[v]) && w[v] > wmax) wmax = w[v];
  uint64_t total = 0;
  if (wmax <= 0) {
    memset(W, 0, sizeof(uint64_t) * N);
    return 0;
  }
  int e;
  frexp(wmax * (double)N, &e);
  const int s = 61 - e;
  for (uint32_t v = 0; v < N; ++v) {
    double x = (isfinite(w[v]) && w[v] > 0) ? ldexp(w[v], s) : 0.0;
    W[v] = (uint64_t)x;
    total += W[v];
  }
  return total;
}

/* First index whose inclusive prefix reaches r, r = lrand48()/RAND_MAX *
 * total (uniform_double, axbc2d.c:2167-2172), clamped to total; N-1 when
 * none (findsample_float + the "myexit == 0" fallback). */
uint32_t oracle_weighted_sample(uint32_t N, const uint64_t *W, uint64_t total, long draw,
                                double *p_out) {
  double rd = (double)draw / 2147483647.0 * (double)total;
  uint64_t r = rd >= (double)total ? total : (uint64_t)rd;
  uint64_t run = 0;
  uint32_t i;
  for (i = 0; i < N; ++i) {
    run += W[i];
    if (run >= r) break;
  }
  uint32_t v = i < N ? i : N - 1;
  *p_out = (double)W[v] / (double)total;
  return v;
}

/* lcc_func (axbc2d.c:1847-1935) at 1x1: for a vertex i of degree d >= 2,
 * counter = number of pairs (j, x) with j in adj(i), x in adj(j), x != i,
 * x in adj(i) (uint32, wrapping); lcc = (float)counter /
 * (float)(uint32)(d*(d-1)); 0 for degree < 2.  Adjacency lists are sorted
 * and duplicate-free, so each membership test is a merge step. */
void oracle_lcc(uint32_t N, const uint32_t *col, const uint32_t *row, float *lcc) {
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t d = col[i + 1] - col[i];
    if (d < 2) {
      lcc[i] = 0.0f;
      continue;
    }
    uint32_t counter = 0;
    for (uint32_t k = col[i]; k < col[i + 1]; ++k) {
      const uint32_t j = row[k];
      uint32_t a = col[j], b = col[i];
      while (a < col[j + 1] && b < col[i + 1]) {
        if (row[a] < row[b])
          ++a;
        else if (row[a] > row[b])
          ++b;
        else {
          if (row[a] != i) ++counter;
          ++a;
          ++b;
        }
      }
    }
    lcc[i] = (float)counter / (float)(uint32_t)(d * (d - 1u));
  }
}

static int run_loop(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
                    const oracle_cfg *cfg, uint32_t *pivots, double *probs, double *bc,
                    oracle_run_stats *rs, int rounds) {
  const int z = cfg->distribution;
  if (z < 0 || z > 6) return -1;
  /* without rounds only the static distributions draw the same sequence,
   * and only without the stop test (which reads bc) */
  if (!rounds && (z > 2 || cfg->bader > 0)) return -1;
  memset(rs, 0, sizeof(*rs));
  memset(bc, 0, sizeof(double) * N);
  uint64_t nbfs = cfg->nbfs < N ? cfg->nbfs : N;
  /* init_rand (axbc2d.c:2143-2153) */
  oracle_srand48(cfg->seed);
  uint32_t *approx = (uint32_t *)xcalloc(cfg->approx_n, sizeof(uint32_t));
  for (uint32_t i = 0; i < cfg->approx_n; ++i) approx[i] = select_root1(N);
  hdn_tables hdn = {0};
  uint32_t mix = (uint32_t)z;
  float *lcc = NULL;
  uint64_t *Wl = NULL, Tl = 0;
  if (z == 0) (void)select_root1(N); /* discarded initial draw (2740) */
  if (z == 6) mix = uniform_int(4, 0); /* initial mixture component (2798) */
  if (z == 1 || z == 6) hdn_build(N, col, row, &hdn);
  if (z == 2 || z == 6) {
    lcc = (float *)xcalloc(N, sizeof(float));
    oracle_lcc(N, col, row, lcc);
    double *w = (double *)xcalloc(N, sizeof(double));
    for (uint32_t v = 0; v < N; ++v) w[v] = (double)lcc[v];
    Wl = (uint64_t *)xcalloc(N, sizeof(uint64_t));
    Tl = oracle_fixed_weights(N, w, Wl);
    free(w);
  }
  int32_t *lvl = (int32_t *)xcalloc(N, sizeof(int32_t));
  uint32_t *sigma = (uint32_t *)xcalloc(N, sizeof(uint32_t));
  double *delta = (double *)xcalloc(N, sizeof(double));
  double *wt = (double *)xcalloc(N, sizeof(double));
  uint64_t *W = (uint64_t *)xcalloc(N, sizeof(uint64_t));
  uint8_t *vs = (uint8_t *)xcalloc(N, 1);
  const uint32_t lim = oracle_lim_x_sample(N);
  uint64_t frozen_total = 0, frozen_at = 0;
  int frozen_ok = 0, frozen_uniform = 0;
  uint64_t *Wb6 = NULL, *Wd6 = NULL, tb6 = 0, td6 = 0, frozen6_at = 0; /* z = 6, batch-synchronous */
  (void)frozen_uniform;
  const double cN = (double)((uint64_t)cfg->bader * (uint64_t)N);
  for (uint64_t ui = 0; ui < nbfs; ++ui) {
    rs->attempts = ui + 1;
    if (cfg->bader > 0) { /* adaptive stop (2846-2870) */
      int stop = 1;
      for (uint32_t i = 0; i < cfg->approx_n; ++i) {
        double b = approx[i] < N ? bc[approx[i]] : 0.0;
        if (b / 2.0 < cN) stop = 0;
      }
      if (stop) {
        rs->stopped = 1;
        rs->attempts = ui;
        break;
      }
    }
    uint32_t v0 = 0, bsm = 0;
    double p = 0;
    int good = 0;
    while (!good) {
      if (bsm > lim) break;
      /* component: the distribution itself, or the mixture's current one */
      const uint32_t k = z == 6 ? mix : (uint32_t)z;
      const int first_group = cfg->sync_batch ? rs->nrounds < cfg->sync_batch : rs->nrounds == 0;
      if (k == 0 || ((k == 3 || k == 4 || k == 5) && first_group)) {
        v0 = select_root1(N);
        p = 1.0 / (double)N;
      } else if (k == 1) {
        uint32_t r = uniform_int(hdn.sum, 0);
        v0 = hdn_pick(N, col, row, &hdn, r, &p);
      } else if (k == 2) {
        v0 = oracle_weighted_sample(N, Wl, Tl, oracle_lrand48(), &p);
      } else if (cfg->sync_batch && (k == 3 || k == 4 || k == 5)) {
        /* batch-synchronous: the distribution of the group's first round */
        if (frozen_at != rs->nrounds / cfg->sync_batch || !frozen_ok) {
          frozen_at = rs->nrounds / cfg->sync_batch;
          frozen_uniform = (rs->nrounds % cfg->sync_batch == 0 && rs->nrounds == 0);
          for (uint32_t v = 0; v < N; ++v)
            wt[v] = k == 3 ? bc[v] : k == 4 ? delta[v] : 1.0 / (1.0 + bc[v]);
          frozen_total = oracle_fixed_weights(N, wt, W);
          frozen_ok = 1;
        }
        v0 = oracle_weighted_sample(N, W, frozen_total, oracle_lrand48(), &p);
      } else {
        for (uint32_t v = 0; v < N; ++v)
          wt[v] = k == 3 ? bc[v] : k == 4 ? delta[v] : 1.0 / (1.0 + bc[v]);
        uint64_t total = oracle_fixed_weights(N, wt, W);
        v0 = oracle_weighted_sample(N, W, total, oracle_lrand48(), &p);
      }
      if (z == 6) { /* mixture probability and next component (3030-3051) */
        const int in = v0 < N;
        p = 1.0 / (double)N;
        if (first_group) {
          p = p + 1.0 / (double)N;
          p = p + 1.0 / (double)N;
        } else if (cfg->sync_batch) {
          /* batch-synchronous: bc and delta of the group's first round */
          if (!Wb6) {
            Wb6 = (uint64_t *)xcalloc(N, sizeof(uint64_t));
            Wd6 = (uint64_t *)xcalloc(N, sizeof(uint64_t));
          }
          if (frozen6_at != rs->nrounds / cfg->sync_batch + 1) {
            frozen6_at = rs->nrounds / cfg->sync_batch + 1;
            for (uint32_t v = 0; v < N; ++v) wt[v] = bc[v];
            tb6 = oracle_fixed_weights(N, wt, Wb6);
            for (uint32_t v = 0; v < N; ++v) wt[v] = delta[v];
            td6 = oracle_fixed_weights(N, wt, Wd6);
          }
          p = p + (in ? (double)Wb6[v0] : 0.0) / (double)tb6;
          p = p + (in ? (double)Wd6[v0] : 0.0) / (double)td6;
        } else {
          for (uint32_t v = 0; v < N; ++v) wt[v] = bc[v];
          uint64_t tb = oracle_fixed_weights(N, wt, W);
          p = p + (in ? (double)W[v0] : 0.0) / (double)tb;
          for (uint32_t v = 0; v < N; ++v) wt[v] = delta[v];
          uint64_t td = oracle_fixed_weights(N, wt, W);
          p = p + (in ? (double)W[v0] : 0.0) / (double)td;
        }
        p = p + (in ? (double)hdn.sdn[v0] : 0.0) / (double)hdn.sum;
        p = p + (in ? (double)Wl[v0] : 0.0) / (double)Tl;
        p = p / 5.0;
        mix = uniform_int(2, 0);
      }
      if (v0 < N && vs[v0])
        ++bsm;
      else
        good = 1;
    }
    if (bsm > lim) continue;
    if (v0 >= N || col[v0 + 1] == col[v0]) { /* degree-0 root (3091-3111) */
      rs->skipped++;
      if (v0 < N) vs[v0] = 1;
      continue;
    }
    if (rounds) oracle_pivot(N, col, row, reach, cfg->heuristic, v0, p, lvl, sigma, delta, bc, NULL);
    vs[v0] = 1;
    pivots[rs->nrounds] = v0;
    probs[rs->nrounds] = p;
    rs->nrounds++;
  }
  free(approx);
  free(lcc);
  free(Wl);
  free(hdn.prefix);
  free(hdn.sdn);
  free(lvl);
  free(sigma);
  free(delta);
  free(wt);
  free(W);
  free(Wb6);
  free(Wd6);
  free(vs);
  return 0;
}

int oracle_run(uint32_t N, const uint32_t *col, const uint32_t *row, const uint32_t *reach,
               const oracle_cfg *cfg, uint32_t *pivots, double *probs, double *bc,
               oracle_run_stats *rs) {
  return run_loop(N, col, row, reach, cfg, pivots, probs, bc, rs, 1);
}

int oracle_draws(uint32_t N, const uint32_t *col, const uint32_t *row, const oracle_cfg *cfg,
                 uint32_t *pivots, double *probs, oracle_run_stats *rs) {
  double *bc = (double *)xcalloc(N, sizeof(double));
  const int rc = run_loop(N, col, row, NULL, cfg, pivots, probs, bc, rs, 0);
  free(bc);
  return rc;
}
