/*
 * axbc_legacy.h -- the reference's device entry points (src/bc2d.h:131-190)
 * on top of the engine, for a drop-in build of AxBC's own host code at
 * -p 1x1 ("mono" path).  Implemented in libaxbc_legacy.so, which reads the
 * host program's globals exactly as src/cuda_kernels.cu does:
 *   uint64_t N; int heuristic;              (src/axbc2d.c:30, 50)
 * All calls act on one default engine context (the reference keeps its
 * device state in file-static globals, src/cuda_kernels.cu:41-80).
 * Errors print the engine's message and exit, as MY_CUDA_CHECK does
 * (src/cudamacro.h:1-16).  delta/bc are FP64 on the device; the float*
 * getters round them to float, the *_f64 getters return them unrounded.
 */
#ifndef AXBC_LEGACY_H
#define AXBC_LEGACY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint32_t LOCINT_T; /* LOCINT at the reference's default width */

/* initcuda: src/cuda_kernels.cu:1558 */
size_t initcuda(uint64_t ned, LOCINT_T *col, LOCINT_T *row);
/* init_bc_1degree_device: cu:1515 */
void init_bc_1degree_device(LOCINT_T *reach);
/* setcuda: cu:1532 */
void setcuda(uint64_t ned, LOCINT_T *col, LOCINT_T *row, LOCINT_T reach_v0);
/* set_mlp_cuda: cu:862 (root at level 0 with sigma 1) */
void set_mlp_cuda(LOCINT_T row, int level, int sigma);
/* scan_col_csc_cuda_mono: cu:1088, returns the size of the new frontier */
LOCINT_T scan_col_csc_cuda_mono(int ncol, int level);
/* scan_frt_csc_cuda_mono: cu:1013 */
LOCINT_T scan_frt_csc_cuda_mono(int offset, int ncol, int depth);
/* pre_update_bc_cuda: cu:358 */
void pre_update_bc_cuda(LOCINT_T *reach, LOCINT_T v0, LOCINT_T *all);
/* update_bc_cuda: cu:289 */
void update_bc_cuda(LOCINT_T v0, double p, int ncol, const uint64_t nvisited);
/* getters: cu:1458-1473 */
void get_deg(LOCINT_T *deg);
void get_bc(float *bc);
void get_delta(float *delta);
void get_bc_f64(double *bc);
void get_delta_f64(double *delta);
/* per-vertex levels / path counts of the last round: cu:1431, 1462 */
void get_lvl(int *lvl);
void get_sigma(LOCINT_T *sigma);
/* Entry points of the 2D process grid (bc_func, src/axbc2d.c:1239-1429;
 * declared at src/bc2d.h:142-176).  Exported so that AxBC's host code links
 * unchanged; a 2D run that reaches one prints a message and exits (the 1x1
 * path, bc_func_mono, never calls them). */
LOCINT_T scan_col_csc_cuda(LOCINT_T *rbuf, LOCINT_T ld, int *rnum, int np, LOCINT_T *sbuf, int *snum,
                           LOCINT_T *frt, LOCINT_T *frt_sigma, int level);
LOCINT_T scan_frt_csc_cuda(const LOCINT_T *frt, int ncol, int depth, float *hSRbuf);
LOCINT_T write_delta_cuda(LOCINT_T ncol, float *hRFbuf, float *hSFbuf);
LOCINT_T append_rows_cuda(LOCINT_T *rbuf, LOCINT_T ld, int *rnum, int np, LOCINT_T *frt, LOCINT_T *frt_sigma,
                          LOCINT_T nfrt, int level);
void set_get_overlap(LOCINT_T *sigma, int *lvl);
void set_sigma(LOCINT_T *sigma);
void set_lvl(int *lvl);
int set_delta_cuda(float *hSRbuf, int nrow);
/* fincuda: cu:1659 */
void fincuda(void);
/* pinned host buffers: cu:825-846 */
void *CudaMallocHostSet(size_t size, int val);
void CudaFreeHost(void *ptr);
/* device of this process: src/adtp.c:17 (one GPU: device 0) */
int assignDeviceToProcess(void);

#ifdef __cplusplus
}
#endif

#endif
