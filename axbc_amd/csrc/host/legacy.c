/*
 * The reference's device entry points (src/bc2d.h:131-190) on one default
 * engine context, so that AxBC's own host code (src/axbc2d.c, 1x1 "mono"
 * path) links against the engine instead of src/cuda_kernels.cu.  Reads the
 * host program's globals N and heuristic as the reference's device file
 * does.  Errors print and exit like MY_CUDA_CHECK (src/cudamacro.h:1-16).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "axbc.h"
#include "axbc_legacy.h"

extern uint64_t N;     /* number of vertices, src/axbc2d.c:30 */
extern int heuristic;  /* -H, src/axbc2d.c:50 */

static axbc_ctx *g_ctx = NULL;
static uint32_t g_v0 = 0;

static void legacy_check(int rc, const char *what) {
  if (rc != AXBC_OK) {
    fprintf(stderr, "axbc error in %s: %s.\n", what, axbc_last_error());
    exit(EXIT_FAILURE);
  }
}

int assignDeviceToProcess(void) { return 0; }

size_t initcuda(uint64_t ned, LOCINT_T *col, LOCINT_T *row) {
  size_t bytes = 0;
  legacy_check(axbc_ctx_create(assignDeviceToProcess(), N, ned, col, row, NULL, heuristic, &g_ctx, &bytes),
               "initcuda");
  return bytes;
}

void init_bc_1degree_device(LOCINT_T *reach) { legacy_check(axbc_set_reach(g_ctx, reach), "init_bc_1degree_device"); }

void setcuda(uint64_t ned, LOCINT_T *col, LOCINT_T *row, LOCINT_T reach_v0) {
  (void)ned, (void)col, (void)row, (void)reach_v0; /* reach_v0 is re-read on the device side */
  legacy_check(axbc_begin_round(g_ctx), "setcuda");
}

void set_mlp_cuda(LOCINT_T row, int level, int sigma) {
  if (level != 0 || sigma != 1) {
    fprintf(stderr, "set_mlp_cuda: only the root (level 0, sigma 1) is supported\n");
    exit(EXIT_FAILURE);
  }
  g_v0 = row;
  legacy_check(axbc_set_root(g_ctx, row), "set_mlp_cuda");
}

LOCINT_T scan_col_csc_cuda_mono(int ncol, int level) {
  (void)ncol;
  uint32_t nfrt = 0;
  legacy_check(axbc_forward_level(g_ctx, level, &nfrt), "scan_col_csc_cuda_mono");
  return nfrt;
}

LOCINT_T scan_frt_csc_cuda_mono(int offset, int ncol, int depth) {
  (void)offset;
  legacy_check(axbc_backward_depth(g_ctx, depth), "scan_frt_csc_cuda_mono");
  return (LOCINT_T)ncol;
}

void pre_update_bc_cuda(LOCINT_T *reach, LOCINT_T v0, LOCINT_T *all) {
  (void)reach, (void)v0;
  legacy_check(axbc_reach_sum(g_ctx, all), "pre_update_bc_cuda");
}

void update_bc_cuda(LOCINT_T v0, double p, int ncol, const uint64_t nvisited) {
  (void)ncol;
  legacy_check(axbc_update_bc(g_ctx, v0, p, nvisited), "update_bc_cuda");
}

void get_deg(LOCINT_T *deg) { legacy_check(axbc_get_deg(g_ctx, deg), "get_deg"); }

void get_bc_f64(double *bc) { legacy_check(axbc_get_bc(g_ctx, bc), "get_bc"); }

void get_delta_f64(double *delta) { legacy_check(axbc_get_delta(g_ctx, delta), "get_delta"); }

static void to_float(void (*get)(double *), float *out) {
  double *tmp = (double *)malloc(sizeof(double) * N);
  if (!tmp) {
    fprintf(stderr, "Cannot allocate %zu bytes...\n", (size_t)(sizeof(double) * N));
    exit(EXIT_FAILURE);
  }
  get(tmp);
  for (uint64_t i = 0; i < N; ++i) out[i] = (float)tmp[i];
  free(tmp);
}

void get_bc(float *bc) { to_float(get_bc_f64, bc); }

void get_delta(float *delta) { to_float(get_delta_f64, delta); }

/* Per-vertex state of the last round (get_lvl cu:1431, get_sigma cu:1462
 * copy the same arrays off the device): levels and path counts of the
 * default context. */
void get_lvl(int *lvl) { legacy_check(axbc_get_lvl(g_ctx, lvl), "get_lvl"); }

void get_sigma(LOCINT_T *sigma) { legacy_check(axbc_get_sigma(g_ctx, sigma), "get_sigma"); }

/* The 2D process grid (-p RxC with R*C > 1, bc_func in src/axbc2d.c) is not
 * part of this engine: its frontier exchanges and state overwrites are
 * linked so that AxBC's host code builds unchanged, and stop the program
 * with a message if a 2D run reaches them (bc_func_mono never does). */
static void two_d_only(const char *what) {
  fprintf(stderr, "axbc error in %s: the 2D process grid is not supported by this engine; run with -p 1x1.\n", what);
  exit(EXIT_FAILURE);
}

LOCINT_T scan_col_csc_cuda(LOCINT_T *rbuf, LOCINT_T ld, int *rnum, int np, LOCINT_T *sbuf, int *snum, LOCINT_T *frt,
                           LOCINT_T *frt_sigma, int level) {
  (void)rbuf, (void)ld, (void)rnum, (void)np, (void)sbuf, (void)snum, (void)frt, (void)frt_sigma, (void)level;
  two_d_only("scan_col_csc_cuda");
  return 0;
}

LOCINT_T scan_frt_csc_cuda(const LOCINT_T *frt, int ncol, int depth, float *hSRbuf) {
  (void)frt, (void)ncol, (void)depth, (void)hSRbuf;
  two_d_only("scan_frt_csc_cuda");
  return 0;
}

LOCINT_T write_delta_cuda(LOCINT_T ncol, float *hRFbuf, float *hSFbuf) {
  (void)ncol, (void)hRFbuf, (void)hSFbuf;
  two_d_only("write_delta_cuda");
  return 0;
}

LOCINT_T append_rows_cuda(LOCINT_T *rbuf, LOCINT_T ld, int *rnum, int np, LOCINT_T *frt, LOCINT_T *frt_sigma,
                          LOCINT_T nfrt, int level) {
  (void)rbuf, (void)ld, (void)rnum, (void)np, (void)frt, (void)frt_sigma, (void)nfrt, (void)level;
  two_d_only("append_rows_cuda");
  return 0;
}

void set_get_overlap(LOCINT_T *sigma, int *lvl) {
  (void)sigma, (void)lvl;
  two_d_only("set_get_overlap");
}

void set_sigma(LOCINT_T *sigma) {
  (void)sigma;
  two_d_only("set_sigma");
}

void set_lvl(int *lvl) {
  (void)lvl;
  two_d_only("set_lvl");
}

int set_delta_cuda(float *hSRbuf, int nrow) {
  (void)hSRbuf, (void)nrow;
  two_d_only("set_delta_cuda");
  return 0;
}

void fincuda(void) {
  legacy_check(axbc_ctx_destroy(g_ctx), "fincuda");
  g_ctx = NULL;
}

void *CudaMallocHostSet(size_t size, int val) {
  void *p = axbc_host_alloc(size);
  if (!p) {
    fprintf(stderr, "Cannot allocate %zu pinned bytes\n", size);
    exit(EXIT_FAILURE);
  }
  memset(p, val, size);
  return p;
}

void CudaFreeHost(void *ptr) { axbc_host_free(ptr); }
