"""AxBC's own host code links against the legacy layer: a C translation unit
that declares every src/bc2d.h:135-176 function src/axbc2d.c calls, with the
reference's prototypes (LOCINT = 32-bit unsigned), and defines the globals
the device side reads (N, heuristic), links against libaxbc_legacy.so and
libaxbc_hip.so with no undefined symbol, and its dynamic symbols resolve at
load time.  No device call is made (runs without a GPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "axbc_amd", "lib")

# every bc2d.h function with a live call in src/axbc2d.c (calls inside
# comments left out), with the prototype bc2d.h gives it
PROTOS = """
#include <stddef.h>
#include <stdint.h>
typedef uint32_t LOCINT;
void setcuda(uint64_t ned, LOCINT *col, LOCINT *row, LOCINT reach_v0);
size_t initcuda(uint64_t ned, LOCINT *col, LOCINT *row);
void set_mlp_cuda(LOCINT row, int level, int sigma);
LOCINT scan_col_csc_cuda(LOCINT *rbuf, LOCINT ld, int *rnum, int np, LOCINT *sbuf, int *snum,
                         LOCINT *frt, LOCINT *frt_sigma, int level);
LOCINT scan_col_csc_cuda_mono(int ncol, int level);
LOCINT scan_frt_csc_cuda(const LOCINT *frt, int ncol, int depth, float *hSRbuf);
LOCINT scan_frt_csc_cuda_mono(int offset, int ncol, int depth);
LOCINT write_delta_cuda(LOCINT ncol, float *hRFbuf, float *hSFbuf);
LOCINT append_rows_cuda(LOCINT *rbuf, LOCINT ld, int *rnum, int np, LOCINT *frt, LOCINT *frt_sigma,
                        LOCINT nfrt, int level);
void update_bc_cuda(LOCINT v0, double p, int ncol, const uint64_t nvisited);
void pre_update_bc_cuda(LOCINT *reach, LOCINT v0, LOCINT *all);
void init_bc_1degree_device(LOCINT *reach);
void get_lvl(int *lvl);
void set_lvl(int *lvl);
void get_deg(LOCINT *deg);
void get_bc(float *bc);
void get_delta(float *delta);
void get_sigma(LOCINT *sigma);
void set_sigma(LOCINT *sigma);
void set_get_overlap(LOCINT *sigma, int *lvl);
int set_delta_cuda(float *hSRbuf, int nrow);
void fincuda();
void *CudaMallocHostSet(size_t size, int val);
void CudaFreeHost(void *ptr);
"""
NAMES = ["setcuda", "initcuda", "set_mlp_cuda", "scan_col_csc_cuda", "scan_col_csc_cuda_mono",
         "scan_frt_csc_cuda", "scan_frt_csc_cuda_mono", "write_delta_cuda", "append_rows_cuda",
         "update_bc_cuda", "pre_update_bc_cuda", "init_bc_1degree_device", "get_lvl", "set_lvl", "get_deg",
         "get_bc", "get_delta", "get_sigma", "set_sigma", "set_get_overlap", "set_delta_cuda", "fincuda",
         "CudaMallocHostSet", "CudaFreeHost"]


def test_reference_host_symbol_set_links(tmp_path):
    src = tmp_path / "host.c"
    body = "\n".join(f"  (void *){n}," for n in NAMES)
    src.write_text(PROTOS + f"""
uint64_t N = 0;
int heuristic = 0;
int main(int argc, char **argv) {{
  void *volatile fns[] = {{
{body}
  }};
  (void)argv;
  /* never called: only referenced, so that the linker must resolve them */
  return argc > 100 ? (int)(size_t)fns[0] : 0;
}}
""")
    exe = tmp_path / "host"
    r = subprocess.run(["gcc", "-std=gnu11", "-Wall", "-Werror", str(src), "-o", str(exe), "-L", LIB,
                        "-laxbc_legacy", "-laxbc_hip", "-Wl,--no-undefined", "-Wl,-rpath," + LIB],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, LD_BIND_NOW="1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


def test_legacy_exports_every_name():
    r = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libaxbc_legacy.so")],
                       capture_output=True, text=True, check=True)
    have = {ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()}
    missing = [n for n in NAMES if n not in have]
    assert not missing, missing
