/*
 * Test host for libaxbc_legacy.so: drives the reference's device entry
 * points in the order AxBC's bc_func_mono (src/axbc2d.c:1514-1659) and
 * main() (2607-2686, 3112-3131) call them, on a CSC read from a file, and
 * writes the final FP64 bc.  It defines the globals the legacy layer reads
 * (N, heuristic) as the reference host does.
 *
 * usage: legacy_host GRAPH.bin PIVOTS.bin OUT.bin
 *   GRAPH.bin  u64 N, u64 ned, i32 heuristic, u32 col[N+1], u32 row[ned], u32 reach[N]
 *   PIVOTS.bin u64 count, then count x (u32 v0, f64 p)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "axbc_legacy.h"

uint64_t N = 0;
int heuristic = 0;

static void *rd(FILE *f, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (bytes && fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
  return p;
}

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  FILE *f = fopen(argv[1], "rb");
  uint64_t ned;
  int32_t h;
  if (!f || fread(&N, 8, 1, f) != 1 || fread(&ned, 8, 1, f) != 1 || fread(&h, 4, 1, f) != 1) return 2;
  heuristic = h;
  LOCINT_T *col = (LOCINT_T *)rd(f, 4 * (N + 1));
  LOCINT_T *row = (LOCINT_T *)rd(f, 4 * ned);
  LOCINT_T *reach = (LOCINT_T *)CudaMallocHostSet(4 * N, 0);
  if (fread(reach, 4, N, f) != N) return 2;
  fclose(f);
  f = fopen(argv[2], "rb");
  uint64_t np;
  if (!f || fread(&np, 8, 1, f) != 1) return 2;
  initcuda(ned, col, row);
  init_bc_1degree_device(reach);
  int *hf = (int *)calloc(N + 3, sizeof(int));
  int *hFnum = hf + 1; /* hFnum[-1] readable, as the reference's do-while reaches it */
  for (uint64_t k = 0; k < np; ++k) {
    uint32_t v0;
    double p;
    if (fread(&v0, 4, 1, f) != 1 || fread(&p, 8, 1, f) != 1) return 2;
    setcuda(ned, col, row, reach[v0]);
    int level = 0;
    uint64_t nfrt = 1, nvisited = 1;
    for (uint64_t i = 0; i < N + 2; ++i) hf[i] = 0;
    set_mlp_cuda(v0, 0, 1);
    while (1) {
      level++;
      hFnum[level] = hFnum[level - 1] + (int)nfrt;
      nfrt = scan_col_csc_cuda_mono((int)nfrt, level);
      if (!nfrt) break;
      nvisited += nfrt;
    }
    int depth = level - 2;
    do {
      int ncol = hFnum[depth + 1] - hFnum[depth];
      scan_frt_csc_cuda_mono(hFnum[depth], ncol, depth);
      depth--;
    } while (depth > 0);
    LOCINT_T all = 0;
    if (heuristic == 1 || heuristic == 3) pre_update_bc_cuda(reach, v0, &all);
    all += reach[v0];
    update_bc_cuda(v0, p, (int)N, nvisited + all);
  }
  fclose(f);
  double *bc = (double *)malloc(8 * N);
  get_bc_f64(bc);
  f = fopen(argv[3], "wb");
  fwrite(bc, 8, N, f);
  fclose(f);
  fincuda();
  CudaFreeHost(reach);
  return 0;
}
