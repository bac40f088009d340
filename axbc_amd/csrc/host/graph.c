/*
 * Graph ingest for the 1x1 sampled-Brandes path: mirror, sort/normalise,
 * one-pass 1-degree reduction, CSC build, text reader and lattice generator.
 *
 * The semantics follow AxBC's host code at -p 1x1 (one rank, so the 2D
 * ownership order of cmpedge degenerates to plain (u, v) order), including
 * its quirks, because the CSC and reach[] produced here are the hot path's
 * input and must equal the reference's:
 *   - norm_graph keeps the first sorted edge even when it is a self-loop
 *     (src/axbc2d.c:818-841: the self-loop test starts at index 1);
 *   - the degree reduction is a single pass whose removed-edge bookkeeping
 *     counts only the first `pnod` sorted removed entries into reach[]
 *     (src/axbc2d.c:695-807, 866-876).
 * Edges are packed 64-bit keys (u << 32 | v); see axbc_host.h.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axbc_host.h"

static void *xalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "axbc: cannot allocate %zu bytes\n", n);
    abort();
  }
  return p;
}

uint64_t *axbc_pack_pairs(const uint64_t *pairs, uint64_t ned) {
  uint64_t *k = (uint64_t *)xalloc(sizeof(uint64_t) * ned);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < ned; ++i) k[i] = AXBC_EDGE(pairs[2 * i], pairs[2 * i + 1]);
  return k;
}

uint64_t *axbc_mirror(uint64_t *ed, uint64_t *ned, int undirected) {
  /* src/axbc2d.c:333-351: reversed copies of the non-loop edges go after
   * the originals. */
  if (!undirected) return ed;
  uint64_t n0 = *ned, n = 0;
  ed = (uint64_t *)realloc(ed, sizeof(uint64_t) * 2 * (n0 ? n0 : 1));
  if (!ed) {
    fprintf(stderr, "axbc: cannot grow edge list\n");
    abort();
  }
  for (uint64_t i = 0; i < n0; ++i) {
    uint32_t u = AXBC_SRC(ed[i]), v = AXBC_DST(ed[i]);
    if (u != v) ed[n0 + n++] = AXBC_EDGE(v, u);
  }
  *ned = n0 + n;
  return ed;
}

/* Parallel LSD radix sort of 64-bit keys, 16-bit digits, only over the
 * digits that can be non-zero. */
void axbc_sort_edges(uint64_t *a, uint64_t n, int nthreads) {
  if (n < 2) return;
  uint64_t maxk = 0;
#pragma omp parallel for reduction(max : maxk) schedule(static)
  for (uint64_t i = 0; i < n; ++i) maxk = a[i] > maxk ? a[i] : maxk;
  int passes = 0;
  while (passes < 4 && (maxk >> (16 * passes))) ++passes;
  if (nthreads <= 0) nthreads = 1;
  uint64_t *tmp = (uint64_t *)xalloc(sizeof(uint64_t) * n);
  size_t *hist = (size_t *)xalloc(sizeof(size_t) * 65536 * (size_t)nthreads);
  uint64_t *src = a, *dst = tmp;
  for (int p = 0; p < passes; ++p) {
    const int sh = 16 * p;
    memset(hist, 0, sizeof(size_t) * 65536 * (size_t)nthreads);
#pragma omp parallel num_threads(nthreads)
    {
      int t = 0, nt = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      nt = omp_get_num_threads();
#endif
      uint64_t lo = n * (uint64_t)t / (uint64_t)nt, hi = n * (uint64_t)(t + 1) / (uint64_t)nt;
      size_t *h = hist + 65536 * (size_t)t;
      for (uint64_t i = lo; i < hi; ++i) h[(src[i] >> sh) & 0xFFFF]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int d = 0; d < 65536; ++d)
          for (int tt = 0; tt < nt; ++tt) {
            size_t c = hist[65536 * (size_t)tt + d];
            hist[65536 * (size_t)tt + d] = run;
            run += c;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) dst[h[(src[i] >> sh) & 0xFFFF]++] = src[i];
    }
    uint64_t *s = src;
    src = dst;
    dst = s;
  }
  if (src != a) memcpy(a, src, sizeof(uint64_t) * n);
  free(tmp);
  free(hist);
}

uint64_t axbc_norm_graph(uint64_t *ed, uint64_t ned, int nthreads) {
  if (ned == 0) return 0;
  axbc_sort_edges(ed, ned, nthreads);
  /* src/axbc2d.c:826-840: keep ed[0]; afterwards keep an edge when it
   * differs from its predecessor and is not a loop. */
  uint64_t l = 1;
  for (uint64_t n = 1; n < ned; ++n) {
    uint64_t e = ed[n];
    if (e != ed[n - 1] && AXBC_SRC(e) != AXBC_DST(e)) ed[l++] = e;
  }
  return l;
}

uint64_t axbc_degree_reduction(uint64_t *ed, uint64_t ned, uint64_t N, uint32_t *reach,
                               uint32_t *rem_ed, int nthreads) {
  (void)N;
  *rem_ed = 0;
  if (ned == 0) return 0;
  axbc_sort_edges(ed, ned, nthreads);
  uint64_t *kept = (uint64_t *)xalloc(sizeof(uint64_t) * ned);
  uint64_t *gone = (uint64_t *)xalloc(sizeof(uint64_t) * ned);
  uint64_t ne = 0, nod = 0;
  const uint64_t NONE = UINT64_MAX;
  uint64_t prev_u = NONE;
  /* First step (src/axbc2d.c:725-757): an edge (u,v) whose source has no
   * other edge (the next edge has another source and the previous edge
   * had another source) is removed and recorded as (v,u).  Loops and
   * duplicates are skipped but still update prev_u. */
  for (uint64_t i = 0; i + 1 < ned; ++i) {
    uint64_t u = AXBC_SRC(ed[i]), v = AXBC_DST(ed[i]);
    uint64_t nu = AXBC_SRC(ed[i + 1]);
    if (u == v || ed[i] == ed[i + 1]) {
      prev_u = u;
      continue;
    }
    if (u != nu && u != prev_u)
      gone[nod++] = AXBC_EDGE(v, u);
    else
      kept[ne++] = ed[i];
    prev_u = u;
  }
  {
    uint64_t u = AXBC_SRC(ed[ned - 1]), v = AXBC_DST(ed[ned - 1]);
    if (u == prev_u)
      kept[ne++] = ed[ned - 1];
    else if (u != v)
      gone[nod++] = AXBC_EDGE(v, u);
  }
  axbc_sort_edges(gone, nod, nthreads);
  /* Second step (src/axbc2d.c:769-795): merge-walk the kept edges against
   * the sorted removed list, dropping the reverse copies.  Entries of the
   * removed list past its end read as zero, as the zero-filled buffer of
   * the reference does. */
#define GONE_AT(k) ((k) < nod ? gone[k] : 0)
  uint64_t pnod = 0, out = 0;
  if (nod > 0) {
    for (uint64_t i = 0; i < ne; ++i) {
      while (AXBC_SRC(kept[i]) > AXBC_SRC(GONE_AT(pnod)) && pnod < nod) pnod++;
      if (kept[i] == GONE_AT(pnod)) {
        pnod++;
        continue;
      }
      ed[out++] = kept[i];
    }
  } else {
    memcpy(ed, kept, sizeof(uint64_t) * ne);
    out = ne;
  }
#undef GONE_AT
  /* init_bc_1degree (src/axbc2d.c:866-876): reach counts the first pnod
   * removed entries by their first endpoint. */
  for (uint64_t i = 0; i < pnod && i < nod; ++i) reach[AXBC_SRC(gone[i])]++;
  *rem_ed = (uint32_t)pnod;
  free(kept);
  free(gone);
  return out;
}

int axbc_build_csc(const uint64_t *ed, uint64_t ned, uint64_t N, uint32_t **colp, uint32_t **rowp) {
  if (ned > UINT32_MAX) return -1;
  uint32_t *col = (uint32_t *)calloc(N + 1, sizeof(uint32_t));
  uint32_t *row = (uint32_t *)xalloc(sizeof(uint32_t) * (ned ? ned : 1));
  if (!col) abort();
  for (uint64_t i = 0; i < ned; ++i) col[AXBC_DST(ed[i]) + 1]++;
  for (uint64_t j = 0; j < N; ++j) col[j + 1] += col[j];
  uint32_t *fill = (uint32_t *)xalloc(sizeof(uint32_t) * (N ? N : 1));
  memcpy(fill, col, sizeof(uint32_t) * N);
  for (uint64_t i = 0; i < ned; ++i) row[fill[AXBC_DST(ed[i])]++] = AXBC_SRC(ed[i]);
  free(fill);
  *colp = col;
  *rowp = row;
  return 0;
}

uint64_t *axbc_read_edge_file(const char *path, uint64_t N, int undirected, uint64_t *ned,
                              char *errbuf, size_t errlen) {
  /* src/axbc2d.c:356-445 at one rank: lines containing '#' are comments
   * (a "# Nodes: n Edges: m" header is echoed as "N=n E=m"), other lines
   * hold "i j"; ids must be < N.  Lines that hold no pair are skipped: the
   * reference appends the previous line's pair again for them, a duplicate
   * that its normalisation removes before the CSC is built. */
  FILE *f = fopen(path, "r");
  if (!f) {
    snprintf(errbuf, errlen, "Cannot open file %s...", path);
    return NULL;
  }
  uint64_t cap = 1 << 16, n = 0, line = 0;
  uint64_t *e = (uint64_t *)xalloc(sizeof(uint64_t) * cap);
  char buf[1024];
  while (fgets(buf, sizeof(buf), f)) {
    if (strchr(buf, '#')) {
      if (strstr(buf, "Nodes:")) {
        unsigned long long hn = 0, he = ~0ull;
        (void)sscanf(buf, "# Nodes: %llu Edges: %llu", &hn, &he);
        printf("N=%llu E=%llu\n", hn, he);
      }
      continue;
    }
    unsigned long long i, j;
    if (sscanf(buf, "%llu %llu", &i, &j) != 2) continue;
    ++line;
    if (i >= N || j >= N) {
      snprintf(errbuf, errlen, "In file %s line %llu found invalid edge for N=%llu: (%llu, %llu)", path,
               (unsigned long long)line, (unsigned long long)N, i, j);
      free(e);
      fclose(f);
      return NULL;
    }
    if (n == cap) {
      cap *= 2;
      e = (uint64_t *)realloc(e, sizeof(uint64_t) * cap);
      if (!e) abort();
    }
    e[n++] = AXBC_EDGE(i, j);
  }
  fclose(f);
  *ned = n;
  return axbc_mirror(e, ned, undirected);
}

static inline uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

uint64_t *axbc_lattice_edges(uint32_t side, double p_del, uint64_t seed, uint64_t *ned) {
  /* Edge k of the lattice, in the order "all horizontal edges row-major,
   * then all vertical edges row-major", survives when the k-th splitmix64
   * output, as a double in [0,1), is >= p_del. */
  uint64_t n = (uint64_t)side * side;
  uint64_t nh = (uint64_t)side * (side - 1), total = 2 * nh;
  uint64_t *e = (uint64_t *)xalloc(sizeof(uint64_t) * (total ? total : 1));
  uint64_t m = 0, s = seed;
  for (uint64_t k = 0; k < total; ++k) {
    double r = (double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
    if (r < p_del) continue;
    uint64_t u, v;
    if (k < nh) {
      uint64_t r0 = k / (side - 1), c0 = k % (side - 1);
      u = r0 * side + c0;
      v = u + 1;
    } else {
      uint64_t kk = k - nh;
      u = kk;
      v = kk + side;
    }
    e[m++] = AXBC_EDGE(u, v);
  }
  /* Largest connected component (by vertex count), ties to the smaller
   * root id. */
  uint32_t *par = (uint32_t *)xalloc(sizeof(uint32_t) * n);
  for (uint64_t i = 0; i < n; ++i) par[i] = (uint32_t)i;
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = uf_find(par, AXBC_SRC(e[i])), b = uf_find(par, AXBC_DST(e[i]));
    if (a != b) {
      if (a < b)
        par[b] = a;
      else
        par[a] = b;
    }
  }
  uint32_t *cnt = (uint32_t *)calloc(n, sizeof(uint32_t));
  uint32_t best = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r = uf_find(par, (uint32_t)i);
    cnt[r]++;
  }
  for (uint64_t i = 0; i < n; ++i)
    if (cnt[i] > cnt[best]) best = (uint32_t)i;
  uint64_t kept = 0;
  for (uint64_t i = 0; i < m; ++i)
    if (uf_find(par, AXBC_SRC(e[i])) == best) e[kept++] = e[i];
  free(par);
  free(cnt);
  *ned = kept;
  return e;
}

int axbc_graph_from_edges(uint64_t *ed, uint64_t ned, uint64_t N, int heuristic, int nthreads,
                          axbc_graph *g) {
  memset(g, 0, sizeof(*g));
  g->N = N;
  g->reach = (uint32_t *)calloc(N ? N : 1, sizeof(uint32_t));
  if (!g->reach) return -1;
  if (heuristic != 0) ned = axbc_norm_graph(ed, ned, nthreads);
  if (heuristic == 1) ned = axbc_degree_reduction(ed, ned, N, g->reach, &g->rem_ed, nthreads);
  ned = axbc_norm_graph(ed, ned, nthreads);
  g->ned = ned;
  int rc = axbc_build_csc(ed, ned, N, &g->col, &g->row);
  free(ed);
  return rc;
}

int axbc_graph_kronecker(int scale, int edgef, uint64_t seed1, uint64_t seed2, int undirected,
                         int heuristic, int nthreads, axbc_graph *g) {
  const double ini[4] = {.57, .19, .19, .05};
  int64_t m = ((int64_t)1 << scale) * edgef;
  uint64_t *pairs = axbc_kronecker_edges(scale, m, seed1, seed2, ini, nthreads);
  if (!pairs) return -1;
  uint64_t *ed = axbc_pack_pairs(pairs, (uint64_t)m);
  free(pairs);
  uint64_t ned = (uint64_t)m;
  ed = axbc_mirror(ed, &ned, undirected);
  int rc = axbc_graph_from_edges(ed, ned, (uint64_t)1 << scale, heuristic, nthreads, g);
  g->generated = (uint64_t)m;
  return rc;
}

void axbc_graph_free(axbc_graph *g) {
  free(g->col);
  free(g->row);
  free(g->reach);
  memset(g, 0, sizeof(*g));
}
