/*
 * Graph500-compatible Kronecker (R-MAT) edge generator.
 *
 * AxBC generates its input with the Graph500 reference generator
 * (src/axbc2d.c:518-527 -> generator/make_graph.c:271-353).  The hot path
 * only sees the graph through the sorted, de-duplicated CSC, so what must be
 * reproduced is the exact MULTISET of generated edges for a given
 * (scale, edge count, seed pair, initiator).  That multiset is fixed by the
 * published algorithm, which this file restates in its own code:
 *
 *  - the L'Ecuyer-Blouin-Couture MRG (m = 2^31-1, k = 5, multipliers
 *    107374182 and 104480), with random access by "skipping" the state
 *    forward by a 192-bit exponent, i.e. multiplying by A^e mod m
 *    (generator/splittable_mrg.c:22-60, 217-231).  The skip table here is a
 *    plain 5x5 matrix table A^(v*256^b) built at start-up by repeated
 *    squaring; any exact evaluation of A^e gives the same state.
 *  - the recursive quadrant split: each node of the 2x2 recursion re-derives
 *    its state from the root state skipped by ((src+N)/n, (tgt+N)/n), draws
 *    the four quadrant counts with sequential binomials (BTRD, Hoermann
 *    1993) or, for <= 20 edges, one categorical draw per edge, and folds the
 *    lower-left quadrant into the upper-right one on the diagonal
 *    ("clip-and-flip", generator/graph_generator.c:95-260).
 *  - the rand_sort vertex permutation of Cong and Bader (bucket by a random
 *    key in [0, 2n+128), sort each bucket ascending, Fisher-Yates shuffle it
 *    with its own skipped stream; generator/permutation_gen.c:148-224).
 *  - (min, max) orientation of every permuted edge
 *    (generator/make_graph.c:63-75).
 *
 * The final edge scramble of the reference (generator/scramble_edges.c)
 * only permutes the order of the edges, so it is not reproduced.
 * The node recursion is independent per node (every node re-derives its
 * state from the root), so the subtrees are generated in parallel with
 * OpenMP; the output is identical for any thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axbc_host.h"

#define MRG_M 0x7FFFFFFFull
#define MRG_MUL1 107374182ull
#define MRG_MUL5 104480ull

static inline uint32_t mrg_fold(uint64_t t) {
  /* t < 2^63: reduce modulo 2^31 - 1 with two folds and a final subtract. */
  t = (t & MRG_M) + (t >> 31);
  t = (t & MRG_M) + (t >> 31);
  return (uint32_t)(t >= MRG_M ? t - MRG_M : t);
}

/* 5x5 matrices over GF(2^31-1), acting on the column state (z1..z5). */
typedef struct { uint32_t a[5][5]; } mat5;

static void mat5_mul(const mat5 *x, const mat5 *y, mat5 *out) {
  mat5 r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 5; ++k) acc = mrg_fold(acc + (uint64_t)x->a[i][k] * y->a[k][j]);
      r.a[i][j] = (uint32_t)acc;
    }
  *out = r;
}

static void mat5_identity(mat5 *m) {
  memset(m, 0, sizeof(*m));
  for (int i = 0; i < 5; ++i) m->a[i][i] = 1;
}

/* skip_tab[b][v] = A^(v * 256^b), b = 0..23 covers a 192-bit exponent. */
static mat5 (*skip_tab)[256] = NULL;
static pthread_once_t skip_once = PTHREAD_ONCE_INIT;

static void build_skip_table(void) {
  mat5 (*t)[256] = (mat5(*)[256])malloc(sizeof(mat5) * 24 * 256);
  if (!t) {
    fprintf(stderr, "axbc: cannot allocate MRG skip table\n");
    abort();
  }
  mat5 step; /* A^(256^b) */
  memset(&step, 0, sizeof(step));
  step.a[0][0] = (uint32_t)MRG_MUL1;
  step.a[0][4] = (uint32_t)MRG_MUL5;
  for (int i = 1; i < 5; ++i) step.a[i][i - 1] = 1;
  for (int b = 0; b < 24; ++b) {
    mat5_identity(&t[b][0]);
    for (int v = 1; v < 256; ++v) mat5_mul(&t[b][v - 1], &step, &t[b][v]);
    /* next step = step^256 = t[b][255] * step */
    mat5 next;
    mat5_mul(&t[b][255], &step, &next);
    step = next;
  }
  skip_tab = t;
}

static inline void mrg_apply(const mat5 *m, axbc_mrg *st) {
  uint32_t r[5];
  for (int i = 0; i < 5; ++i) {
    uint64_t acc = 0;
    for (int k = 0; k < 5; ++k) acc = mrg_fold(acc + (uint64_t)m->a[i][k] * st->z[k]);
    r[i] = (uint32_t)acc;
  }
  memcpy(st->z, r, sizeof(r));
}

void axbc_mrg_skip(axbc_mrg *st, uint64_t e_high, uint64_t e_mid, uint64_t e_low) {
  pthread_once(&skip_once, build_skip_table);
  uint64_t words[3] = {e_low, e_mid, e_high};
  for (int w = 0; w < 3; ++w) {
    uint64_t e = words[w];
    for (int byte = 0; e; ++byte, e >>= 8) {
      unsigned v = (unsigned)(e & 0xFF);
      if (v) mrg_apply(&skip_tab[w * 8 + byte][v], st);
    }
  }
}

void axbc_mrg_seed_from_user(uint64_t seed1, uint64_t seed2, axbc_mrg *st) {
  /* Spread two 64-bit seeds over five non-zero residues
   * (generator/utils.c:84-90). */
  st->z[0] = (uint32_t)((seed1 & 0x3FFFFFFF) + 1);
  st->z[1] = (uint32_t)(((seed1 >> 30) & 0x3FFFFFFF) + 1);
  st->z[2] = (uint32_t)((seed2 & 0x3FFFFFFF) + 1);
  st->z[3] = (uint32_t)(((seed2 >> 30) & 0x3FFFFFFF) + 1);
  st->z[4] = (uint32_t)(((seed2 >> 60) << 4) + (seed1 >> 60) + 1);
}

uint32_t axbc_mrg_next_uint(axbc_mrg *st) {
  uint32_t fresh = mrg_fold(MRG_MUL1 * st->z[0] + MRG_MUL5 * (uint64_t)st->z[4]);
  st->z[4] = st->z[3];
  st->z[3] = st->z[2];
  st->z[2] = st->z[1];
  st->z[1] = st->z[0];
  st->z[0] = fresh;
  return fresh;
}

/* 1/(2^31-1) and 1/(2^31-1)^2, each correctly rounded to double (the
 * square of the rounded reciprocal rounds to the same double). */
static const double kInvM = 1.0 / 2147483647.0;
static const double kInvM2 = (1.0 / 2147483647.0) * (1.0 / 2147483647.0);

double axbc_mrg_next_double(axbc_mrg *st) {
  /* Two draws: the first carries the high part, the second the low part. */
  double hi = (double)axbc_mrg_next_uint(st);
  double lo = (double)axbc_mrg_next_uint(st);
  return hi * kInvM + lo * kInvM2;
}

uint64_t axbc_mrg_random_up_to(axbc_mrg *st, uint64_t n) {
  /* Unbiased draw in [0, n) by rejection (generator/utils.c:52-79). */
  if (n <= 1) return 0;
  if (n <= 0x7FFFFFFFull) {
    uint64_t limit = (0x7FFFFFFFull / n) * n;
    for (;;) {
      uint64_t x = axbc_mrg_next_uint(st);
      if (x < limit) return x % n;
    }
  }
  const uint64_t span = 0x3FFFFFFF00000001ull; /* (2^31-1)^2 */
  uint64_t limit = (span / n) * n;
  for (;;) {
    uint64_t x = (uint64_t)axbc_mrg_next_uint(st) * 0x7FFFFFFFull;
    x += axbc_mrg_next_uint(st);
    if (x < limit) return x % n;
  }
}

/* ---- binomial variates ------------------------------------------------ */

/* Stirling-series correction fc(k) = ln k! - [(k+1/2) ln(k+1) - (k+1) +
 * ln(2 pi)/2]; k <= 9 from the table published with BTRD (Hoermann 1993,
 * "The generation of binomial random variates", Table 1), larger k from the
 * asymptotic series. */
static double stirling_fc(int k) {
  static const double table[10] = {
      0.08106146679532726, 0.04134069595540929, 0.02767792568499834, 0.02079067210376509,
      0.01664469118982119, 0.01387612882307075, 0.01189670994589177, 0.01041126526197209,
      0.009255462182712733, 0.008330563433362871};
  if (k <= 9) return table[k];
  const double c12 = 1. / 12, c360 = 1. / 360, c1260 = 1. / 1260;
  double q = 1. / (k + 1);
  return (c12 - (c360 - c1260 * q * q) * q * q) * q;
}

/*
 * Binomial(n, p) variate.  For n*p < 10 the geometric waiting-time method
 * (Devroye, Non-Uniform Random Variate Generation, ch. X, p. 525) with the
 * series ln(1-p) = 2 r^-1 (1 + r^-2/3 + r^-4/5 + r^-6/7), r = 1 - 2/p
 * (ibid. p. 500); otherwise Hoermann's BTRD transformed-rejection algorithm
 * (steps 0-3.4 of the paper).  The arithmetic follows the published steps
 * term by term so that the variates are bit-identical to the reference
 * implementation (generator/btrd_binomial_distribution.c:37-153).
 */
int64_t axbc_btrd_binomial(int64_t n_in, double p, axbc_mrg *st) {
  if (p == 0.) return 0;
  if (p > .5) return n_in - axbc_btrd_binomial(n_in, 1. - p, st);
  size_t n_orig = (size_t)n_in;
  if ((double)n_orig * p < 10) {
    double r = 1. - 2. / p;
    double ir = 1. / r;
    double ir2 = ir * ir;
    double ir4 = ir2 * ir2;
    double ir6 = ir4 * ir2;
    double ln1mp = 2. * ir * (1 + ir2 / 3. + ir4 / 5. + ir6 / 7.);
    double inv_ln1mp = 1. / ln1mp;
    int count = 0, total = 0;
    do {
      total += (int)ceil(log(axbc_mrg_next_double(st)) * inv_ln1mp);
      ++count;
    } while ((size_t)total <= n_orig);
    return count - 1;
  }
  if (n_orig > 1000000000) {
    return axbc_btrd_binomial(1000000000, p, st) + axbc_btrd_binomial((int64_t)n_orig - 1000000000, p, st);
  }
  int n = (int)n_orig;
  /* step 0: set-up */
  int m = (int)floor((n + 1) * p);
  double r = p / (1. - p);
  double nr = (n + 1) * r;
  double npq = n * p * (1. - p);
  double sq = sqrt(npq);
  double b = 1.15 + 2.53 * sq;
  double a = -.0873 + .0248 * b + .01 * p;
  double c = n * p + .5;
  double alpha = (2.83 + 5.1 / b) * sq;
  double vr = .92 - 4.2 / b;
  double urvr = .86 * vr;
  for (;;) {
    /* step 1 */
    double v = axbc_mrg_next_double(st);
    if (v <= urvr) {
      double u = v / vr - .43;
      /* The paper's step 1 divides by (0.5 - |u|); the reference generator
       * divides by (0.5 + |u|), which narrows the variates (variance about
       * 0.8 n p q).  Kept, since the graphs must equal the reference's. */
      return (int)floor((2 * a / (.5 + fabs(u)) + b) * u + c);
    }
    /* step 2 */
    double w = axbc_mrg_next_double(st);
    double u;
    if (v >= vr) {
      u = w - .5;
    } else {
      u = v / vr - .93;
      u = (u > 0. ? 1. : u < 0. ? -1. : 0.) * .5 - u;
      v = vr * w;
    }
    /* step 3.0 */
    double us = .5 - fabs(u);
    int k = (int)floor((2 * a / us + b) * u + c);
    if (k < 0 || k > n) continue;
    v *= alpha / (a / (us * us) + b);
    int km = k >= m ? k - m : m - k;
    if (km > 15) {
      /* step 3.2: squeeze on log(v) */
      double rho = (km / npq) * (((km / 3 + .625) * km + 1. / 6) / npq + .5);
      double t = -km * km / (2 * npq);
      v = log(v);
      if (v < t - rho) return k;
      if (v > t + rho) continue;
      /* step 3.3 */
      int nm = n - m + 1;
      double h = (m + .5) * log((m + 1) / (r * nm)) + stirling_fc(m) + stirling_fc(n - m);
      /* step 3.4: final acceptance */
      int nk = n - k + 1;
      double bound = h + (n + 1) * log((double)nm / nk) + (k + .5) * log(nk * r / (k + 1)) -
                     stirling_fc(k) - stirling_fc(n - k);
      if (v <= bound) return k;
    } else {
      /* step 3.1: evaluate the ratio f(k)/f(m) recursively */
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

/* ---- rand_sort permutation --------------------------------------------- */

static void shuffle_bucket(int64_t *a, int64_t cnt, axbc_mrg *st) {
  for (int64_t i = cnt - 1; i > 0; --i) {
    int64_t j = (int64_t)axbc_mrg_random_up_to(st, (uint64_t)(i + 1));
    if (i != j) {
      int64_t t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
}

void axbc_rand_permutation(const axbc_mrg *root, int64_t n, int64_t *perm, int nthreads) {
  const int64_t nkeys = 2 * n + 128;
  uint32_t *key = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  int64_t *start = (int64_t *)calloc((size_t)nkeys + 1, sizeof(int64_t));
  if (!key || !start) {
    fprintf(stderr, "axbc: out of memory in rand_permutation\n");
    abort();
  }
  (void)nthreads;
  /* Every element draws its key from its own skipped stream (1, i, 0). */
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < n; ++i) {
    axbc_mrg s = *root;
    axbc_mrg_skip(&s, 1, (uint64_t)i, 0);
    key[i] = (uint32_t)axbc_mrg_random_up_to(&s, (uint64_t)nkeys);
  }
  for (int64_t i = 0; i < n; ++i) start[key[i] + 1]++;
  for (int64_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  /* Counting sort by key, stable in i: each bucket comes out ascending. */
  {
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nkeys);
    memcpy(fill, start, sizeof(int64_t) * (size_t)nkeys);
    for (int64_t i = 0; i < n; ++i) perm[fill[key[i]]++] = i;
    free(fill);
  }
  /* Shuffle every bucket with more than one element using stream
   * (1, key, 100). */
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads)
  for (int64_t k = 0; k < nkeys; ++k) {
    int64_t cnt = start[k + 1] - start[k];
    if (cnt > 1) {
      axbc_mrg s = *root;
      axbc_mrg_skip(&s, 1, (uint64_t)k, 100);
      shuffle_bucket(perm + start[k], cnt, &s);
    }
  }
  free(key);
  free(start);
}

/* ---- Kronecker recursion ------------------------------------------------ */

typedef struct {
  double ini[4];
  int64_t total_nverts;
  axbc_mrg root;
  int64_t *out; /* 2 * nedges */
} kron_ctx;

typedef struct {
  int64_t first, count, nverts, src, tgt;
} kron_node;

static int categorical4(const kron_ctx *k, axbc_mrg *st) {
  double r = axbc_mrg_next_double(st);
  for (int j = 0; j < 3; ++j) {
    if (r < k->ini[j]) return j;
    r -= k->ini[j];
  }
  return 3;
}

static void quadrant_counts(const kron_ctx *k, int64_t nedges, axbc_mrg *st, int64_t cnt[4]) {
  if (nedges <= 20) {
    cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
    for (int64_t e = 0; e < nedges; ++e) cnt[categorical4(k, st)]++;
  } else {
    int64_t left = nedges;
    double rest = 1.;
    for (int i = 0; i < 3; ++i) {
      cnt[i] = axbc_btrd_binomial(left, k->ini[i] / rest, st);
      left -= cnt[i];
      rest -= k->ini[i];
    }
    cnt[3] = left;
  }
}

static inline axbc_mrg node_state(const kron_ctx *k, const kron_node *nd) {
  axbc_mrg s = k->root;
  axbc_mrg_skip(&s, 0, (uint64_t)((nd->src + k->total_nverts) / nd->nverts),
                (uint64_t)((nd->tgt + k->total_nverts) / nd->nverts));
  return s;
}

/* Expand a node: returns the number of children written to `kids` (0 when
 * the node is a leaf and its edges were written). */
static int kron_expand(const kron_ctx *k, const kron_node *nd, kron_node kids[4]) {
  axbc_mrg st = node_state(k, nd);
  if (nd->nverts == 1) {
    for (int64_t i = 0; i < nd->count; ++i) {
      k->out[2 * (nd->first + i)] = nd->src;
      k->out[2 * (nd->first + i) + 1] = nd->tgt;
    }
    return 0;
  }
  if (nd->count == 1) {
    int64_t n = nd->nverts, s = nd->src, t = nd->tgt;
    while (n > 1) {
      int q = categorical4(k, &st);
      int so = q / 2, to = q % 2;
      if (s == t && so > to) {
        int tmp = so;
        so = to;
        to = tmp;
      }
      n /= 2;
      s += n * so;
      t += n * to;
    }
    k->out[2 * nd->first] = s;
    k->out[2 * nd->first + 1] = t;
    return 0;
  }
  int64_t cnt[4];
  quadrant_counts(k, nd->count, &st, cnt);
  if (nd->src == nd->tgt) { /* fold below-diagonal quadrant onto the upper one */
    cnt[1] += cnt[2];
    cnt[2] = 0;
  }
  int64_t half = nd->nverts / 2, fei = nd->first;
  int nk = 0;
  for (int q = 0; q < 4; ++q) {
    if (!cnt[q]) continue;
    kids[nk].first = fei;
    kids[nk].count = cnt[q];
    kids[nk].nverts = half;
    kids[nk].src = nd->src + half * (q / 2);
    kids[nk].tgt = nd->tgt + half * (q % 2);
    ++nk;
    fei += cnt[q];
  }
  return nk;
}

static void kron_subtree(const kron_ctx *k, const kron_node *nd) {
  kron_node kids[4];
  int nk = kron_expand(k, nd, kids);
  for (int i = 0; i < nk; ++i) kron_subtree(k, &kids[i]);
}

uint64_t *axbc_kronecker_edges(int scale, int64_t nedges, uint64_t seed1, uint64_t seed2,
                               const double initiator[4], int nthreads) {
  if (nthreads <= 0) nthreads = 1;
  kron_ctx k;
  for (int i = 0; i < 4; ++i) k.ini[i] = initiator[i];
  k.total_nverts = (int64_t)1 << scale;
  axbc_mrg_seed_from_user(seed1, seed2, &k.root);
  k.out = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(nedges > 0 ? nedges : 1));
  if (!k.out) return NULL;
  if (nedges > 0) {
    /* Breadth-first expansion of the top of the tree, then the subtrees in
     * parallel.  Every node derives its stream from the root state, so the
     * split point does not change the result. */
    size_t cap = 1 << 16, n = 1;
    kron_node *front = (kron_node *)malloc(sizeof(kron_node) * cap);
    front[0] = (kron_node){0, nedges, k.total_nverts, 0, 0};
    while (n < (size_t)(64 * nthreads) && n > 0) {
      kron_node *next = (kron_node *)malloc(sizeof(kron_node) * (4 * n + 4));
      size_t m = 0;
      int internal = 0;
      for (size_t i = 0; i < n; ++i) {
        kron_node kids[4];
        int nk = kron_expand(&k, &front[i], kids);
        for (int j = 0; j < nk; ++j) next[m++] = kids[j];
        internal |= nk > 0;
      }
      free(front);
      front = next;
      n = m;
      if (!internal) break;
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (size_t i = 0; i < n; ++i) kron_subtree(&k, &front[i]);
    free(front);
  }
  /* Vertex permutation from the root stream, then (min, max) orientation. */
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)k.total_nverts);
  axbc_rand_permutation(&k.root, k.total_nverts, perm, nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t e = 0; e < nedges; ++e) {
    int64_t a = perm[k.out[2 * e]], b = perm[k.out[2 * e + 1]];
    k.out[2 * e] = a < b ? a : b;
    k.out[2 * e + 1] = a < b ? b : a;
  }
  free(perm);
  return (uint64_t *)k.out;
}
