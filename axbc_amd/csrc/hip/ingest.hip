// Graph construction on the device: mirror, sort + normalise, the one-pass
// 1-degree reduction and the CSC build, with the results of the host
// pipeline (graph.c) bit for bit.
//
// Edges are packed 64-bit keys (u << 32 | v).  The steps and their
// reference semantics (src/axbc2d.c at -p 1x1):
//   mirror       reversed copies of the non-loop edges (333-351)
//   norm_graph   sort; keep the first edge even if it is a loop, then every
//                edge that differs from its predecessor and is not a loop
//                (818-841)
//   reduction    an edge (u,v) is removed, and recorded as (v,u), when u has
//                no other edge (neither neighbour in the sorted list shares
//                its source); loops and duplicates are skipped (725-757).
//                The recorded list is sorted and merge-walked against the
//                kept edges (769-795): a pointer into the recorded list
//                skips entries whose source is below the kept edge's, and a
//                kept edge equal to the entry under the pointer is dropped
//                and advances it.  Within one source group the pointer thus
//                matches the longest prefix of the group's recorded entries
//                that are all kept edges, then stops; that prefix rule is
//                what the parallel form below computes.  reach counts the
//                recorded entries before the pointer's final position
//                (866-876).
//   build_csc    column = second endpoint, rows of a column in increasing
//                first endpoint (883-907): a sort by (v, u) and a count.
// Sorting, selection and scans use hipCUB device primitives.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>

#include "../../../include/axbc.h"
#include "engine_internal.h"

using namespace axbc;

namespace {

__device__ __forceinline__ uint32_t src_of(uint64_t e) { return (uint32_t)(e >> 32); }
__device__ __forceinline__ uint32_t dst_of(uint64_t e) { return (uint32_t)e; }
__device__ __forceinline__ uint64_t edge(uint32_t u, uint32_t v) { return ((uint64_t)u << 32) | v; }

int grid_of(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 8192));
}

// first index in a[0, n) with a[i] >= x
__device__ __forceinline__ uint64_t lower_bound_u64(const uint64_t *__restrict__ a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_reverse(const uint64_t *__restrict__ e, uint64_t n, uint64_t *rev, uint8_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = e[i];
    rev[i] = edge(dst_of(x), src_of(x));
    flag[i] = src_of(x) != dst_of(x);
  }
}

__global__ void k_norm_flags(const uint64_t *__restrict__ e, uint64_t n, uint8_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = e[i];
    flag[i] = i == 0 || (x != e[i - 1] && src_of(x) != dst_of(x));
  }
}

// 1-degree reduction, first step: kept / recorded / skipped per edge of
// the sorted list; the recorded key is the reversed edge.
__global__ void k_reduce_classify(const uint64_t *__restrict__ e, uint64_t n, uint8_t *kept, uint8_t *gone,
                                  uint64_t *rev) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = e[i];
    const uint32_t u = src_of(x), v = dst_of(x);
    const bool has_prev = i > 0;
    const uint32_t pu = has_prev ? src_of(e[i - 1]) : 0u;
    bool k = false, g = false;
    if (i + 1 < n) {
      const uint64_t y = e[i + 1];
      if (!(u == v || x == y)) {
        if (u != src_of(y) && !(has_prev && u == pu))
          g = true;
        else
          k = true;
      }
    } else {  // the last edge: kept when its source equals the previous one
      if (has_prev && u == pu)
        k = true;
      else if (u != v)
        g = true;
    }
    kept[i] = k;
    gone[i] = g;
    rev[i] = edge(v, u);
  }
}

// Recorded entry j: in_kept, and the first index of its source group.
__global__ void k_gone_props(const uint64_t *__restrict__ gone, uint64_t nod, const uint64_t *__restrict__ kept,
                             uint64_t nk, int64_t *badpos, uint64_t *gstart) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nod; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = gone[j];
    const uint64_t p = lower_bound_u64(kept, nk, g);
    const bool in = p < nk && kept[p] == g;
    badpos[j] = in ? -1 : (int64_t)j;
    gstart[j] = lower_bound_u64(gone, nod, (uint64_t)src_of(g) << 32);
  }
}

struct MaxI64 {
  __device__ __forceinline__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};

// matched[j]: entry j lies in the all-kept prefix of its source group
// (no entry missing from the kept list in [group start, j]).
__global__ void k_matched(const int64_t *__restrict__ lastbad, const uint64_t *__restrict__ gstart, uint64_t nod,
                          uint8_t *matched) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nod; j += (uint64_t)gridDim.x * blockDim.x)
    matched[j] = lastbad[j] < (int64_t)gstart[j];
}

// A kept edge survives unless it equals a matched recorded entry.
__global__ void k_kept_survives(const uint64_t *__restrict__ kept, uint64_t nk, const uint64_t *__restrict__ gone,
                                uint64_t nod, const uint8_t *__restrict__ matched, uint8_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nk; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = kept[i];
    const uint64_t p = lower_bound_u64(gone, nod, x);
    flag[i] = !(p < nod && gone[p] == x && matched[p]);
  }
}

// Final pointer position: the start of the last kept edge's source group
// in the recorded list plus its matched prefix.
__global__ void k_pointer_end(const uint64_t *__restrict__ kept, uint64_t nk, const uint64_t *__restrict__ gone,
                              uint64_t nod, const uint8_t *__restrict__ matched, uint64_t *pnod) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (nk == 0 || nod == 0) {
    *pnod = 0;
    return;
  }
  const uint32_t s = src_of(kept[nk - 1]);
  uint64_t p = lower_bound_u64(gone, nod, (uint64_t)s << 32);
  while (p < nod && src_of(gone[p]) == s && matched[p]) ++p;
  *pnod = p;
}

__global__ void k_reach(const uint64_t *__restrict__ gone, const uint64_t *__restrict__ pnod, uint32_t *reach) {
  const uint64_t m = *pnod;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&reach[src_of(gone[j])], 1u);
}

__global__ void k_swap_ends(const uint64_t *__restrict__ e, uint64_t n, uint64_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = edge(dst_of(e[i]), src_of(e[i]));
}

// Sorted (v, u) keys: rows, and column counts at col[v + 1].
__global__ void k_csc_rows(const uint64_t *__restrict__ vu, uint64_t n, uint32_t *row, uint32_t *cnt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = vu[i];
    row[i] = dst_of(x);
    atomicAdd(&cnt[src_of(x) + 1], 1u);
  }
}

struct Ingest {
  hipStream_t s = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int end_bit = 64;
  ~Ingest() {
    if (tmp) (void)hipFree(tmp);
  }
  // temporary storage of hipCUB calls, grown on demand
  int reserve(size_t bytes) {
    if (bytes <= tmp_bytes) return AXBC_OK;
    if (tmp) (void)hipFree(tmp);
    tmp = nullptr;
    tmp_bytes = 0;
    if (hipMalloc(&tmp, bytes) != hipSuccess) return fail(AXBC_ENOMEM, "ingest scratch of %s bytes (line %d)", "many", __LINE__);
    tmp_bytes = bytes;
    return AXBC_OK;
  }
  int sort(const uint64_t *in, uint64_t *out, uint64_t n) {
    size_t b = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, b, in, out, (int)n, 0, end_bit, s));
    int rc = reserve(b);
    if (rc) return rc;
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, b, in, out, (int)n, 0, end_bit, s));
    return AXBC_OK;
  }
  template <typename T>
  int select(const T *in, const uint8_t *flag, T *out, uint64_t n, uint64_t *count, uint64_t *dnum) {
    size_t b = 0;
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, b, in, flag, out, dnum, (int)n, s));
    int rc = reserve(b);
    if (rc) return rc;
    HIP_TRY(hipcub::DeviceSelect::Flagged(tmp, b, in, flag, out, dnum, (int)n, s));
    HIP_TRY(hipMemcpyAsync(count, dnum, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return AXBC_OK;
  }
};

template <typename T>
int dev_alloc(T **p, uint64_t count) {
  if (hipMalloc((void **)p, sizeof(T) * std::max<uint64_t>(count, 1)) != hipSuccess)
    return fail(AXBC_ENOMEM, "ingest allocation of %s elements failed (line %d)", "many", __LINE__);
  return AXBC_OK;
}

struct DevBufs {
  uint64_t *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *dnum = nullptr;
  uint8_t *f1 = nullptr, *f2 = nullptr;
  int64_t *i1 = nullptr, *i2 = nullptr;
  uint32_t *row = nullptr, *cnt = nullptr, *col = nullptr, *reach = nullptr;
  ~DevBufs() {
    void *p[] = {a, b, c, d, dnum, f1, f2, i1, i2, row, cnt, col, reach};
    for (void *x : p)
      if (x) (void)hipFree(x);
  }
};

// sort + norm_graph on buf (n keys); result in buf, count returned
int normalise(Ingest &in, DevBufs &m, uint64_t *&buf, uint64_t *&spare, uint64_t &n) {
  if (n == 0) return AXBC_OK;
  int rc;
  if ((rc = in.sort(buf, spare, n))) return rc;
  hipLaunchKernelGGL(k_norm_flags, dim3(grid_of(n)), dim3(256), 0, in.s, spare, n, m.f1);
  return in.select(spare, m.f1, buf, n, &n, m.dnum);
}

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};

// Stream, key width and buffers for ned input edges of N vertices.
int ingest_setup(int device, uint64_t ned, uint64_t N, int undirected, Ingest &in, StreamGuard &guard, DevBufs &m) {
  const uint64_t cap = undirected ? 2 * ned : ned;
  if (cap > 0x7FFFFFFFull) return fail(AXBC_EINVAL, "too many edges for one pass%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&in.s, hipStreamNonBlocking));
  guard.s = in.s;
  int bits = 0;
  while (bits < 32 && (N - 1) >> bits) ++bits;
  in.end_bit = 32 + std::max(bits, 1);  // keys are below N << 32
  int rc;
  if ((rc = dev_alloc(&m.a, cap)) || (rc = dev_alloc(&m.b, cap)) || (rc = dev_alloc(&m.c, cap)) ||
      (rc = dev_alloc(&m.d, cap)) || (rc = dev_alloc(&m.dnum, 1)) || (rc = dev_alloc(&m.f1, cap)) ||
      (rc = dev_alloc(&m.f2, cap)) || (rc = dev_alloc(&m.reach, N)))
    return rc;
  HIP_TRY(hipMemsetAsync(m.reach, 0, sizeof(uint32_t) * N, in.s));
  return AXBC_OK;
}

// The preprocessing of the ned edges in m.a (which holds room for their
// mirror); results into the host arrays.
int ingest_run(Ingest &in, DevBufs &m, uint64_t ned, uint64_t N, int undirected, int heuristic, uint32_t *col,
               uint32_t *row, uint64_t row_cap, uint32_t *reach, uint64_t *ned_out, uint32_t *rem_ed) {
  int rc;
  hipStream_t s = in.s;
  uint64_t n = ned;
  uint64_t *buf = m.a, *spare = m.b;
  // mirror: the reversed non-loop edges after the originals
  if (undirected && ned) {
    hipLaunchKernelGGL(k_reverse, dim3(grid_of(ned)), dim3(256), 0, s, buf, ned, spare, m.f1);
    uint64_t added = 0;
    if ((rc = in.select(spare, m.f1, buf + ned, ned, &added, m.dnum))) return rc;
    n = ned + added;
  }
  uint64_t pnod = 0;
  if (heuristic != 0 && (rc = normalise(in, m, buf, spare, n))) return rc;
  if (heuristic == 1 && n) {
    // buf holds the sorted list; m.c = kept, m.d = recorded (sorted)
    hipLaunchKernelGGL(k_reduce_classify, dim3(grid_of(n)), dim3(256), 0, s, buf, n, m.f1, m.f2, spare);
    uint64_t nk = 0, nod = 0;
    if ((rc = in.select(buf, m.f1, m.c, n, &nk, m.dnum))) return rc;
    if ((rc = in.select(spare, m.f2, buf, n, &nod, m.dnum))) return rc;
    if (nod) {
      if ((rc = in.sort(buf, m.d, nod))) return rc;
      if ((rc = dev_alloc(&m.i1, nod)) || (rc = dev_alloc(&m.i2, nod))) return rc;
      uint64_t *gstart = spare;  // free now
      hipLaunchKernelGGL(k_gone_props, dim3(grid_of(nod)), dim3(256), 0, s, m.d, nod, m.c, nk, m.i1, gstart);
      size_t b = 0;
      HIP_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, b, m.i1, m.i2, MaxI64(), (int)nod, s));
      if ((rc = in.reserve(b))) return rc;
      HIP_TRY(hipcub::DeviceScan::InclusiveScan(in.tmp, b, m.i1, m.i2, MaxI64(), (int)nod, s));
      uint8_t *matched = m.f2;
      hipLaunchKernelGGL(k_matched, dim3(grid_of(nod)), dim3(256), 0, s, m.i2, gstart, nod, matched);
      hipLaunchKernelGGL(k_kept_survives, dim3(grid_of(nk)), dim3(256), 0, s, m.c, nk, m.d, nod, matched, m.f1);
      hipLaunchKernelGGL(k_pointer_end, dim3(1), dim3(64), 0, s, m.c, nk, m.d, nod, matched, m.dnum);
      hipLaunchKernelGGL(k_reach, dim3(grid_of(nod)), dim3(256), 0, s, m.d, m.dnum, m.reach);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(&pnod, m.dnum, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if ((rc = in.select(m.c, m.f1, buf, nk, &n, m.dnum))) return rc;
    } else {
      HIP_TRY(hipMemcpyAsync(buf, m.c, sizeof(uint64_t) * nk, hipMemcpyDeviceToDevice, s));
      n = nk;
    }
  }
  if ((rc = normalise(in, m, buf, spare, n))) return rc;
  // CSC: sort by (v, u); rows in that order, columns by count
  if (row_cap < n) return fail(AXBC_EINVAL, "row buffer too small%s (line %d)", "", __LINE__);
  if ((rc = dev_alloc(&m.row, n)) || (rc = dev_alloc(&m.cnt, N + 1))) return rc;
  HIP_TRY(hipMemsetAsync(m.cnt, 0, sizeof(uint32_t) * (N + 1), s));
  if (n) {
    hipLaunchKernelGGL(k_swap_ends, dim3(grid_of(n)), dim3(256), 0, s, buf, n, spare);
    if ((rc = in.sort(spare, buf, n))) return rc;
    hipLaunchKernelGGL(k_csc_rows, dim3(grid_of(n)), dim3(256), 0, s, buf, n, m.row, m.cnt);
    HIP_TRY(hipGetLastError());
  }
  if ((rc = dev_alloc(&m.col, N + 1))) return rc;
  size_t b = 0;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, b, m.cnt, m.col, (int)(N + 1), s));
  if ((rc = in.reserve(b))) return rc;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(in.tmp, b, m.cnt, m.col, (int)(N + 1), s));
  HIP_TRY(hipMemcpyAsync(col, m.col, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, s));
  if (n) HIP_TRY(hipMemcpyAsync(row, m.row, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(reach, m.reach, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *ned_out = n;
  if (rem_ed) *rem_ed = (uint32_t)pnod;
  return AXBC_OK;
}

}  // namespace

extern "C" {

int axbc_ingest(int device, const uint64_t *edges, uint64_t ned, uint64_t N, int undirected, int heuristic,
                uint32_t *col, uint32_t *row, uint64_t row_cap, uint32_t *reach, uint64_t *ned_out, uint32_t *rem_ed) {
  if (!col || !reach || !ned_out || (ned && !edges) || N == 0 || N > 0xFFFFFFFFull)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  if (heuristic < 0 || heuristic > 1) return fail(AXBC_EINVAL, "heuristic must be 0 or 1%s (line %d)", "", __LINE__);
  for (uint64_t i = 0; i < ned; ++i)
    if ((edges[i] >> 32) >= N || (edges[i] & 0xFFFFFFFFull) >= N)
      return fail(AXBC_EINVAL, "edge endpoint out of range%s (line %d)", "", __LINE__);
  Ingest in;
  StreamGuard guard;
  DevBufs m;
  int rc;
  if ((rc = ingest_setup(device, ned, N, undirected, in, guard, m))) return rc;
  if (ned) HIP_TRY(hipMemcpyAsync(m.a, edges, sizeof(uint64_t) * ned, hipMemcpyHostToDevice, in.s));
  return ingest_run(in, m, ned, N, undirected, heuristic, col, row, row_cap, reach, ned_out, rem_ed);
}

int axbc_kronecker_ingest(int device, int scale, uint64_t nedges, uint64_t seed1, uint64_t seed2,
                          const double initiator[4], int undirected, int heuristic, uint32_t *col, uint32_t *row,
                          uint64_t row_cap, uint32_t *reach, uint64_t *ned_out, uint32_t *rem_ed) {
  if (!initiator || !col || !reach || !ned_out || scale < 1 || scale > 30)
    return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  if (heuristic < 0 || heuristic > 1) return fail(AXBC_EINVAL, "heuristic must be 0 or 1%s (line %d)", "", __LINE__);
  const uint64_t N = 1ull << scale;
  Ingest in;
  StreamGuard guard;
  DevBufs m;
  int rc;
  if ((rc = ingest_setup(device, nedges, N, undirected, in, guard, m))) return rc;
  // the generated edges go straight into the first buffer
  if ((rc = kronecker_device_edges(in.s, scale, nedges, seed1, seed2, initiator, m.a))) return rc;
  return ingest_run(in, m, nedges, N, undirected, heuristic, col, row, row_cap, reach, ned_out, rem_ed);
}

}  // extern "C"
