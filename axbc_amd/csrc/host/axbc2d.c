/*
 * axbc2d -- command-line driver of the MI355X sampled-Brandes engine.
 *
 * Same flags and pivot-loop semantics as AxBC's main() at -p 1x1
 * (src/axbc2d.c:2214-3223): graph from the Kronecker generator (-S/-E) or a
 * text edge list (-f/-n), optional 1-degree reduction (-H 1), pivots drawn
 * with srand48/lrand48 in the reference's order by the chosen heuristic
 * (-z), adaptive stop test (-c/-x), -N attempts, per-pivot log line and
 * result table, -o result file.  The per-pivot work runs on the GPU
 * through libaxbc_hip.so (include/axbc.h).
 *
 * Flags of this engine only:
 *   -B W   roots per batch: 0 one root per round (default), 64, 128 or 256
 *          (256: two bit-packed sweeps side by side, or one workgroup per
 *          root on deep graphs)
 *   -s S   generator seed pair (S, S+1) (default: the reference's 23, 24)
 *   -T T   host threads for graph construction
 *   -q     no per-pivot lines
 *   -G f   write the generated graph as a text edge list to f and exit
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axbc.h"
#include "axbc_host.h"

#define DEFAULT_SEED 4343543

static double now_s(void) {
  struct timeval t;
  gettimeofday(&t, NULL);
  return (double)t.tv_sec + 1e-6 * (double)t.tv_usec;
}

static void die(const char *fmt, const char *arg) {
  fprintf(stderr, fmt, arg);
  fputc('\n', stderr);
  exit(EXIT_FAILURE);
}

static void check(int rc, const char *what) {
  if (rc != AXBC_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, axbc_last_error());
    exit(EXIT_FAILURE);
  }
}

static void usage(const char *p) {
  fprintf(stderr,
          "Usage:\n"
          "\t %s -p 1x1 [-o outfile] [-D] [-d] [-N <# of searches>]\n"
          "\t -> to visit a graph read from file:\n"
          "\t\t -f <graph file> -n <# vertices> [-r <start vert>]\n"
          "\t -> to visit an RMAT graph:\n"
          "\t\t -S <scale> [-E <edge factor>]\n"
          "\t Where:\n"
          "\t\t -c <c> -x <x>  adaptive stop: bc/2 >= c*N on x check vertices\n"
          "\t\t -z <0..6>      pivot distribution\n"
          "\t\t -H <0|1>       1-degree reduction\n"
          "\t\t -U             DO NOT make graph Undirected\n"
          "\t\t -R <0|1>       fixed / time-based seed\n"
          "\t\t -B <0|64|128|256>  roots per GPU batch\n"
          "\t\t -g             graph generation and preprocessing on the GPU\n",
          p);
  exit(EXIT_FAILURE);
}

/* Preprocessing on the device (axbc_ingest): the same CSC, reach[] and
 * removed-edge count as axbc_graph_from_edges; frees `ed`. */
static int device_ingest(uint64_t *ed, uint64_t ned, uint64_t N, int undirected, int heuristic, axbc_graph *g) {
  memset(g, 0, sizeof(*g));
  g->N = N;
  g->col = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1));
  g->row = (uint32_t *)malloc(sizeof(uint32_t) * (2 * ned + 1));
  g->reach = (uint32_t *)malloc(sizeof(uint32_t) * (N ? N : 1));
  if (!g->col || !g->row || !g->reach) die("%s", "host allocation failed");
  uint64_t out = 0;
  const int rc = axbc_ingest(0, ed, ned, N, undirected, heuristic, g->col, g->row, 2 * ned + 1, g->reach, &out,
                             &g->rem_ed);
  free(ed);
  g->ned = out;
  g->generated = ned;
  return rc;
}

/* Generation and preprocessing on the device (axbc_kronecker_ingest): the
 * graph of axbc_graph_kronecker without the edges leaving the GPU. */
static int device_generate(int scale, int edgef, uint64_t s1, uint64_t s2, int undirected, int heuristic,
                           axbc_graph *g) {
  const double ini[4] = {.57, .19, .19, .05};
  const uint64_t N = (uint64_t)1 << scale, m = N * (uint64_t)edgef;
  memset(g, 0, sizeof(*g));
  g->N = N;
  g->col = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1));
  g->row = (uint32_t *)malloc(sizeof(uint32_t) * (2 * m + 1));
  g->reach = (uint32_t *)malloc(sizeof(uint32_t) * N);
  if (!g->col || !g->row || !g->reach) die("%s", "host allocation failed");
  uint64_t out = 0;
  const int rc = axbc_kronecker_ingest(0, scale, m, s1, s2, ini, undirected, heuristic, g->col, g->row, 2 * m + 1,
                                       g->reach, &out, &g->rem_ed);
  g->ned = out;
  g->generated = m;
  return rc;
}

/* srand48 / lrand48 of the reference's loop, on a state of its own
 * (nrand48: the same generator), so that the device can compute values
 * ahead of it; pos counts the values drawn. */
static unsigned short g_rng[3];
static uint64_t g_rng_pos;
static void drv_srand48(long seed) {
  g_rng[0] = 0x330E;
  g_rng[1] = (unsigned short)(seed & 0xFFFF);
  g_rng[2] = (unsigned short)((seed >> 16) & 0xFFFF);
  g_rng_pos = 0;
}
static long drv_lrand48(void) {
  ++g_rng_pos;
  return nrand48(g_rng);
}
static uint64_t drv_state48(void) {
  return (uint64_t)g_rng[0] | ((uint64_t)g_rng[1] << 16) | ((uint64_t)g_rng[2] << 32);
}

/* uniform_int(max, min) = lrand48() % (max - min + 1) + min in uint32
 * (src/axbc2d.c:2156-2164); an empty modulus uses the raw draw. */
static uint32_t uniform_int(uint32_t max, uint32_t min) {
  uint32_t m = max - min + 1u;
  long r = drv_lrand48();
  return m ? (uint32_t)((uint32_t)r % m) + min : (uint32_t)r + min;
}

/* r = lrand48()/RAND_MAX * total (uniform_double, src/axbc2d.c:2167-2172)
 * on the fixed-point scale of the weights, clamped to total */
static uint64_t fixed_draw(uint64_t total) {
  double rd = (double)drv_lrand48() / 2147483647.0 * (double)total;
  return rd >= (double)total ? total : (uint64_t)rd;
}

/* The vertex a first draw of table 1 (-z 1, over the deg^2 prefix) or 2
 * (-z 2, over the LCC prefix) picks: searched on the device a window of
 * generator positions at a time (axbc_draw_window); consumes the draw and
 * checks it against the device's value. */
#define DRAW_WINDOW 1024
typedef struct {
  uint64_t base;
  uint32_t count;
  uint32_t lr[DRAW_WINDOW], v[DRAW_WINDOW];
} draw_window;
static draw_window g_win[3];

static uint32_t window_pick(axbc_ctx *ctx, int table, uint64_t total) {
  draw_window *w = &g_win[table];
  if (!w->count || g_rng_pos < w->base || g_rng_pos >= w->base + w->count) {
    check(axbc_draw_window(ctx, table, drv_state48(), total, DRAW_WINDOW, w->lr, w->v), "axbc_draw_window");
    w->base = g_rng_pos;
    w->count = DRAW_WINDOW;
  }
  const uint64_t k = g_rng_pos - w->base;
  if ((uint32_t)drv_lrand48() != w->lr[k]) die("%s", "device draw window out of step with the generator");
  return w->v[k];
}

/* -d: the generated edge list as the reference dumps it (dump_rmat,
 * src/axbc2d.c:550-576, called at 2556-2559): the generator's edges
 * mirrored, normalised (duplicates and self-loops dropped) when -H is not 0,
 * before the 1-degree reduction, sorted by (u, v), one "u<TAB>v" line each,
 * into rmat_S<scale>_EF<edgef>.txt in the working directory.  The list is
 * generated again on the host so the dump does not depend on -g. */
static void dump_rmat(int scale, int edgef, uint64_t s1, uint64_t s2, int undirected, int heuristic, int nthreads) {
  const double ini[4] = {.57, .19, .19, .05};
  const uint64_t m = ((uint64_t)1 << scale) * (uint64_t)edgef;
  uint64_t *pairs = axbc_kronecker_edges(scale, (int64_t)m, s1, s2, ini, nthreads);
  if (!pairs) die("%s", "edge generation failed");
  uint64_t *ed = axbc_pack_pairs(pairs, m);
  free(pairs);
  uint64_t ned = m;
  ed = axbc_mirror(ed, &ned, undirected);
  printf("Dump file...\n");
  if (heuristic != 0)
    ned = axbc_norm_graph(ed, ned, nthreads);
  else
    axbc_sort_edges(ed, ned, nthreads);
  char fname[256];
  snprintf(fname, sizeof(fname), "rmat_S%d_EF%d.txt", scale, edgef);
  printf("DUMP THE GENERATED RMAT FILE IN %s\n", fname);
  FILE *f = fopen(fname, "w+");
  if (!f) die("error opening %s", fname);
  for (uint64_t i = 0; i < ned; ++i)
    fprintf(f, "%" PRIu64 "\t%" PRIu64 "\n", (uint64_t)AXBC_SRC(ed[i]), (uint64_t)AXBC_DST(ed[i]));
  fclose(f);
  free(ed);
}

/* -a: the reference's degree statistics (analyze_deg, src/axbc2d.c:1061-1115,
 * and dump_array, 1040-1057): the CSC column degrees after preprocessing
 * counted into 16 bins with lower limits 0..10, 100, 1000, 10^4, 10^5, 10^7
 * (bin b holds lim[b] <= d < lim[b+1]), appended as one line " c0, c1, ...,"
 * to degree_stats_<id>_0, where <id> is the run's time stamp. */
static void analyze_deg(const axbc_graph *g, unsigned id) {
  static const uint32_t lim[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 100, 1000, 10000, 100000, 10000000};
  uint32_t bins[16] = {0};
  for (uint64_t v = 0; v < g->N; ++v) {
    const uint32_t d = g->col[v + 1] - g->col[v];
    int b = 15;
    while (d < lim[b]) --b;
    ++bins[b];
  }
  char fname[256];
  snprintf(fname, sizeof(fname), "degree_stats_%u_0", id);
  FILE *f = fopen(fname, "a");
  if (!f) die("error opening %s", fname);
  for (int b = 0; b < 16; ++b) fprintf(f, " %d,", (int)bins[b]);
  fprintf(f, "\n");
  fclose(f);
}

/* Numeric option values rejected as the reference's sscanf checks reject
 * them (src/axbc2d.c:2345-2404): a value that does not start with a number
 * ends the run with the reference's message. */
static uint64_t num_u64(const char *arg, const char *msg) {
  char *end = NULL;
  while (*arg == ' ' || *arg == '\t') ++arg;
  if (*arg == '-' || *arg == '+') {
    if (arg[1] < '0' || arg[1] > '9') die(msg, arg);
  } else if (*arg < '0' || *arg > '9') {
    die(msg, arg);
  }
  const uint64_t v = strtoull(arg, &end, 10);
  return v;
}

static int num_int(const char *arg, const char *msg) { return (int)num_u64(arg, msg); }

typedef struct {
  uint32_t v0;
  double p;
  uint64_t ui;
  uint32_t mix;
} round_t;

int main(int argc, char **argv) {
  int R = 0, C = 0, scale = 21, edgef = 16, gread = -1, heuristic = 0, undirected = 1, random_mode = 0;
  int dump = 0, debug = 0, analyze = 0, quiet = 0, width = 0, nthreads = 0;
  uint64_t N = 0, nbfs = 1, startv = 0, seed_s = 0;
  uint32_t bader = 1, approx_n = 1;
  int distribution = 0, rootset = 0, have_seed = 0;
  char *gfile = NULL, *ofile = NULL, *gout = NULL;
  int device_build = 0;
  int c;
  if (argc == 1) usage(argv[0]);
  setvbuf(stdout, NULL, _IOLBF, 0);
  printf("MPI Tasks 1\n");
  for (int i = 0; i < argc; ++i) printf(" %s", argv[i]);
  printf("\n\n");
  while ((c = getopt(argc, argv, "o:p:amhDR:dUf:n:r:S:E:N:H:c:x:z:B:s:T:qG:g")) != -1) {
    switch (c) {
      case 'c': bader = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'x': approx_n = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'z': distribution = atoi(optarg); break;
      case 'H':
        heuristic = num_int(optarg, "Invalid Heuristic Option (-H): %s");
        if (heuristic >= 2) die("H2 and H3 are not implemented in Approx BC(-H): %s", optarg);
        break;
      case 'o': ofile = strdup(optarg); break;
      case 'p':
        if (sscanf(optarg, "%dx%d", &R, &C) != 2) die("Invalid proc mesh field: %s", optarg);
        break;
      case 'd': dump = 1; break;
      case 'R':
        random_mode = atoi(optarg);
        if (random_mode < 0 || random_mode > 3) die("Invalid random option (-R): %s", optarg);
        break;
      case 'D': debug = 1; break;
      case 'm': die("%s", "-m selects the reference's 2D multi-GPU path; this engine runs the 1x1 path"); break;
      case 'f':
        if (gread == 0) die("Unexpected option -%s!", "f");
        gread = 1;
        gfile = strdup(optarg);
        break;
      case 'n':
        if (gread == 0) die("Unexpected option -%s!", "n");
        gread = 1;
        N = num_u64(optarg, "Invalid number of vertices (-n): %s");
        break;
      case 'r':
        startv = num_u64(optarg, "Invalid root vertex (-r): %s");
        rootset = 1;
        break;
      case 'S':
        if (gread == 1) die("Unexpected option -%s!", "S");
        gread = 0;
        scale = num_int(optarg, "Invalid scale (-S): %s");
        if (scale < 0 || scale > 32) die("Invalid scale (-S): %s", optarg);
        N = 1ull << scale;
        break;
      case 'E':
        if (gread == 1) die("Unexpected option -%s!", "E");
        gread = 0;
        edgef = num_int(optarg, "Invalid edge factor (-S): %s");
        break;
      case 'N': nbfs = num_u64(optarg, "Invalid number of bfs (-N): %s"); break;
      case 'U': undirected = 0; break;
      case 'a': analyze = 1; break;
      case 'B': width = atoi(optarg); break;
      case 's':
        seed_s = strtoull(optarg, NULL, 10);
        have_seed = 1;
        break;
      case 'T': nthreads = atoi(optarg); break;
      case 'q': quiet = 1; break;
      case 'g': device_build = 1; break;
      case 'G': gout = strdup(optarg); break;
      default: usage(argv[0]);
    }
  }
  (void)debug;
  if (gread == -1) {  /* defaults: -S 21 -E 16 */
    gread = 0;
    N = 1ull << scale;
  }
  if (approx_n > N) die("%s", "Number of vertices required is lager than N");
  if (distribution > 6 || distribution < 0) die("%s", "Select a strategy between 0 and 6");
  if (gread && (!gfile || !N)) die("%s", "Graph file (-f) and number of vertices (-n) must be specified for file based bfs.");
  if (R != 1 || C != 1) die("%s", "R and C must be 1: this engine runs the 1x1 path on one GPU (-p 1x1)");
  if (startv >= N) die("Invalid start vertex: %s", "");
  if (N > 0xFFFFFFF0ull) die("%s", "Number of vertices too big for 32-bit ids");
  if (width != 0 && width != 64 && width != 128 && width != 256) die("%s", "-B must be 0, 64, 128 or 256");
  if (width && !(distribution == 0 || distribution == 1))
    die("%s", "-B batches need a static distribution (-z 0 or -z 1)");
  if (rootset) random_mode = 0;
  if (nthreads <= 0) {
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
  }
  uint64_t s1 = 23, s2 = 24;  /* gen_graph's seeds (src/axbc2d.c:524) */
  if (have_seed) {
    s1 = seed_s;
    s2 = seed_s + 1;
  }
  printf("Total number of vertices (N): %" PRIu64 "\n", N);
  printf("Processor mesh rows (R): 1\nProcessor mesh columns (C): 1\n");
  if (gread)
    printf("Reading graph from file: %s\n", gfile);
  else
    printf("RMAT graph scale: %d\nRMAT graph edge factor: %d\nNumber of bc rounds: %" PRIu64 "\n", scale, edgef, nbfs);
  printf("HEURISTICs: %s: %d\n\n", heuristic ? "1-degree reduction ON" : "OFF", heuristic);

  /* ---- graph ---------------------------------------------------------- */
  double t0 = now_s();
  axbc_graph g;
  uint64_t generated = 0;
  if (gread) {
    char err[512];
    uint64_t ned = 0;
    uint64_t *ed = axbc_read_edge_file(gfile, N, undirected, &ned, err, sizeof(err));
    if (!ed) die("%s", err);
    generated = ned;
    if (device_build) {
      if (device_ingest(ed, ned, N, 0, heuristic, &g)) die("device graph construction failed: %s", axbc_last_error());
    } else if (axbc_graph_from_edges(ed, ned, N, heuristic, nthreads, &g)) {
      die("%s", "graph construction failed");
    }
  } else if (device_build) {
    if (device_generate(scale, edgef, s1, s2, undirected, heuristic, &g))
      die("device graph construction failed: %s", axbc_last_error());
    generated = g.generated;
  } else {
    if (axbc_graph_kronecker(scale, edgef, s1, s2, undirected, heuristic, nthreads, &g))
      die("%s", "graph construction failed");
    generated = g.generated;
  }
  printf("Edges %s: %" PRIu64 "\n", gread ? "read from file (after mirror)" : "generated", generated);
  printf("Edges in the CSC: %" PRIu64 " (built in %.2f s)\n", g.ned, now_s() - t0);
  if (heuristic == 1) printf("task 0 Total edges removed %u\n", g.rem_ed);
  if (dump && !gread) dump_rmat(scale, edgef, s1, s2, undirected, heuristic, nthreads);
  if (analyze) analyze_deg(&g, (unsigned)time(NULL));
  if (gout) {
    FILE *f = fopen(gout, "w");
    if (!f) die("cannot open %s", gout);
    for (uint64_t v = 0; v < N; ++v)
      for (uint32_t k = g.col[v]; k < g.col[v + 1]; ++k) fprintf(f, "%u\t%" PRIu64 "\n", g.row[k], v);
    fclose(f);
    return 0;
  }

  /* ---- device --------------------------------------------------------- */
  axbc_ctx *ctx = NULL;
  size_t dev_bytes = 0;
  check(axbc_ctx_create(0, N, g.ned, g.col, g.row, g.reach, heuristic, &ctx, &dev_bytes), "axbc_ctx_create");
  if (width) check(axbc_set_batch_width(ctx, width), "axbc_set_batch_width");
  printf("Device memory allocated (MB): %zu\n", dev_bytes >> 20);

  /* ---- pivot loop (src/axbc2d.c:2700-3156) --------------------------------- */
  drv_srand48(random_mode == 1 ? (long)(getpid() + time(NULL)) : DEFAULT_SEED);
  uint32_t *approx = (uint32_t *)calloc(approx_n ? approx_n : 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < approx_n; ++i) approx[i] = uniform_int((uint32_t)N, 0);
  if (bader > 0) {
    printf("Selected Verts to Approximate (TODO REMOVE PRINT): ");
    for (uint32_t i = 0; i < approx_n; ++i) printf("%u ", approx[i]);
    printf("\n");
  } else {
    printf("Stop computation after %" PRIu64 " good samples\n", nbfs);
  }
  check(axbc_set_checks(ctx, approx, approx_n), "axbc_set_checks");
  uint32_t *sdn = NULL, sum_deg = 0;
  uint64_t *lw = NULL, ltotal = 0;
  uint32_t dist_mix = (uint32_t)distribution;
  if (distribution == 0) {
    (void)uniform_int((uint32_t)N, 0);
    printf("UNIFORM sampling strategy\n");
  } else if (distribution == 6) {
    dist_mix = uniform_int(4, 0);
    printf("Mixture sampling strategy: %u\n", dist_mix);
  } else {
    static const char *names[] = {"", "High-Degree Neighborhood ", "LCC ", "BC", "Delta", "BC inverse"};
    printf("%s sampling strategy\n", names[distribution]);
  }
  if (distribution == 2 || distribution == 6) {
    double tl = now_s();
    lw = (uint64_t *)malloc(sizeof(uint64_t) * N);
    check(axbc_lcc_build(ctx, NULL, lw, NULL, &ltotal), "axbc_lcc_build");
    printf("LCC done in %f secs\n", now_s() - tl);
  }
  if (distribution == 1 || distribution == 6) {
    sdn = (uint32_t *)malloc(sizeof(uint32_t) * N);
    check(axbc_hdn_build(ctx, NULL, sdn, &sum_deg), "axbc_hdn_build");
  }
  nbfs = nbfs < N ? nbfs : N;
  const uint32_t lim = (uint32_t)(N * N); /* int LIMxSAMPLE = N*N vs uint32 BSM */
  uint8_t *vs = (uint8_t *)calloc(N, 1);
  uint32_t nrounds = 0, skipped = 0;
  uint64_t sum_er = 0;
  double loop_t0 = now_s(), min_t = 1e30, max_t = 0;
  int stopped = 0;
  const int B = width ? width : 1;
  round_t *pending = (round_t *)calloc((size_t)B, sizeof(round_t));
  uint32_t *bv = (uint32_t *)calloc((size_t)B, sizeof(uint32_t));
  double *bp = (double *)calloc((size_t)B, sizeof(double));
  uint64_t *skip_at = (uint64_t *)calloc((size_t)B + 1, sizeof(uint64_t));
  axbc_pivot_stats *bst = (axbc_pivot_stats *)calloc((size_t)B, sizeof(axbc_pivot_stats));
  int npend = 0;
  uint32_t skipped_pend = 0;
  uint64_t ui;
  for (ui = 0; ui < nbfs; ++ui) {
    if (bader > 0 && npend == 0) {
      int stop = 0;
      check(axbc_stop_test(ctx, bader, &stop), "axbc_stop_test");
      if (stop) {
        printf("exit bader criterion\n");
        stopped = 1;
        break;
      }
    }
    /* draw until a root not used before (or the retry cap) */
    uint32_t v0 = 0, bsm = 0;
    double p = 0;
    int good = 0;
    uint64_t wtotal = 0;
    int wready = 0;
    while (!good) {
      if (bsm > lim) break;
      const uint32_t k = distribution == 6 ? dist_mix : (uint32_t)distribution;
      const int dyn = k >= 3 && k <= 5;
      if (k == 0 || (dyn && nrounds == 0)) {
        v0 = uniform_int((uint32_t)N, 0);
        p = 1.0 / (double)N;
      } else if (k == 1) {
        uint32_t u = window_pick(ctx, 1, sum_deg);
        uint32_t d = g.col[u + 1] - g.col[u];
        if (d > 0) {
          uint32_t nb = d == 1 ? 0 : uniform_int(d - 1, 0);
          u = g.row[g.col[u] + nb];
        }
        v0 = u;
        p = (double)sdn[v0] / (double)sum_deg;
      } else if (k == 2) {
        v0 = window_pick(ctx, 2, ltotal);
        p = (double)lw[v0] / (double)ltotal;
      } else {
        if (!wready || distribution == 6) {
          check(axbc_weights(ctx, (int)k, &wtotal), "axbc_weights");
          wready = 1;
        }
        uint64_t r = fixed_draw(wtotal);
        check(axbc_weights_pick(ctx, r, &v0, &p), "axbc_weights_pick");
      }
      if (distribution == 6) { /* mixture probability, next component (3030-3051) */
        const int in = v0 < N;
        p = 1.0 / (double)N;
        if (nrounds == 0) {
          p = p + 1.0 / (double)N;
          p = p + 1.0 / (double)N;
        } else {
          uint64_t t = 0, w = 0;
          check(axbc_weights(ctx, 3, &t), "axbc_weights");
          if (in) check(axbc_weights_at(ctx, v0, &w), "axbc_weights_at");
          p = p + (double)w / (double)t;
          w = 0;
          check(axbc_weights(ctx, 4, &t), "axbc_weights");
          if (in) check(axbc_weights_at(ctx, v0, &w), "axbc_weights_at");
          p = p + (double)w / (double)t;
        }
        p = p + (in ? (double)sdn[v0] : 0.0) / (double)sum_deg;
        p = p + (in ? (double)lw[v0] : 0.0) / (double)ltotal;
        p = p / 5.0;
        dist_mix = uniform_int(2, 0);
      }
      if (v0 < N && vs[v0])
        ++bsm;
      else
        good = 1;
    }
    if (bsm > lim) continue;
    if (v0 >= N || g.col[v0 + 1] == g.col[v0]) { /* degree-0 root */
      ++skipped_pend;
      if (v0 < N) vs[v0] = 1;
      if (!width) {
        skipped += skipped_pend;
        skipped_pend = 0;
      }
      continue;
    }
    vs[v0] = 1;
    if (!quiet)
      printf("Perform %u/%" PRIu64 " BC from %u (ui=%" PRIu64 " strategy (%d,%u) p =%lf)\n", nrounds + npend + 1,
             nbfs, v0, ui, distribution, dist_mix, p);
    if (!width) {
      axbc_pivot_stats st;
      double t1 = now_s();
      check(axbc_pivot(ctx, v0, p, &st), "axbc_pivot");
      double dt = now_s() - t1;
      min_t = dt < min_t ? dt : min_t;
      max_t = dt > max_t ? dt : max_t;
      sum_er += st.e_r;
      ++nrounds;
      continue;
    }
    skip_at[npend] = skipped_pend;
    pending[npend].v0 = v0;
    pending[npend].p = p;
    pending[npend].ui = ui;
    bv[npend] = v0;
    bp[npend] = p;
    ++npend;
    if (npend == B || ui + 1 == nbfs) {
      int applied = 0;
      double t1 = now_s();
      check(axbc_batch(ctx, bv, bp, npend, bader, &applied, bst), "axbc_batch");
      double dt = now_s() - t1;
      min_t = dt < min_t ? dt : min_t;
      max_t = dt > max_t ? dt : max_t;
      for (int k = 0; k < applied; ++k) sum_er += bst[k].e_r;
      nrounds += (uint32_t)applied;
      if (applied < npend) { /* stopped after round applied-1 */
        skipped += (uint32_t)skip_at[applied - 1];
        ui = pending[applied - 1].ui + 1;
        printf("exit bader criterion\n");
        stopped = 1;
        npend = 0;
        break;
      }
      skipped += skipped_pend;
      skipped_pend = 0;
      npend = 0;
    }
  }
  if (width && npend > 0) { /* loop ended with a partial batch */
    int applied = 0;
    check(axbc_batch(ctx, bv, bp, npend, bader, &applied, bst), "axbc_batch");
    for (int k = 0; k < applied; ++k) sum_er += bst[k].e_r;
    nrounds += (uint32_t)applied;
    skipped += applied < npend ? (uint32_t)skip_at[applied - 1] : skipped_pend;
  } else if (width) {
    skipped += skipped_pend;
  }
  check(axbc_sync(ctx), "axbc_sync");
  const double all_time = now_s() - loop_t0;

  /* ---- results ---------------------------------------------------------- */
  double *bc = (double *)malloc(sizeof(double) * N);
  check(axbc_get_bc(ctx, bc), "axbc_get_bc");
  if (ofile) {
    char resname[1024];
    snprintf(resname, sizeof(resname), "%s_1X1_0.log", ofile);
    printf("Result written to file: %s\n", resname);
    FILE *f = fopen(resname, "w");
    if (f) {
      for (uint64_t k = 0; k < N; ++k) fprintf(f, "%" PRIu64 "\t%.2f\n", k, nrounds ? bc[k] / 2.0 / nrounds : 0.0);
      fclose(f);
    }
  }
  printf("\n------------- RESULTS ---------------\n");
  printf("ClusterId\tSkip\tRounds\tExecTime\tRoundsTime\tMax\tMin\tMean\n");
  printf("0\t%u\t%u\t%lf\t%lf\t%lf\t%lf\t%lf\n", skipped, nrounds, all_time, all_time, max_t,
         min_t > 1e29 ? 0.0 : min_t, nrounds ? all_time / nrounds : 0.0);
  printf("task 0 BC skipped: %u \n", skipped);
  printf("task 0 BC rounds: %u  \n", nrounds);
  printf("task 0 BC rounds total time: %lf sec\n", all_time);
  printf("task 0 pivots/s: %.3f  GTEPS (fwd+bwd): %.4f  stopped: %d  batch: %d\n",
         all_time > 0 ? nrounds / all_time : 0.0, all_time > 0 ? 2.0 * (double)sum_er / all_time / 1e9 : 0.0, stopped,
         width);
  free(bc);
  free(vs);
  free(approx);
  free(sdn);
  free(lw);
  free(pending);
  free(bv);
  free(bp);
  free(skip_at);
  free(bst);
  axbc_ctx_destroy(ctx);
  axbc_graph_free(&g);
  return 0;
}
