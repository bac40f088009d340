// axbc_batch_utility_noalloc: the batch utility (axbc_batch_utility) for
// callers that cannot allocate device memory.  Its staging and partial-row
// buffers live in one file-scope device array per GPU, `scratch`, held until
// the process ends; the entry point calls no hipMalloc / hipFree.
//
// Host code never passes the array's name to a kernel or a copy: on the host
// that name is only a stand-in for the device symbol.  Each device's address
// comes from hipGetSymbolAddress, taken once per device with that device
// current, and is used everywhere the host needs the array.
//
// Every call on one GPU shares that GPU's array: calls that overlap on the
// same GPU (two host threads, two contexts) race on it.  Each call writes
// every byte it later reads (targets, partial rows and the output row are
// stored before they are read), so nothing carries over between calls.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../../include/axbc.h"
#include "engine_internal.h"

using namespace axbc;

namespace {

constexpr size_t kScrAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kScrAlign - 1) / kScrAlign * kScrAlign; }
// layout: partial rows (kUtilWaves x 128 doubles), output row (128 doubles),
// target ids, target weights
constexpr size_t kScrPart = 0;
constexpr size_t kScrOut = align_up(kScrPart + sizeof(double) * kUtilWaves * 128);
constexpr size_t kScrVert = align_up(kScrOut + sizeof(double) * 128);
constexpr size_t kScrWeight = align_up(kScrVert + sizeof(uint32_t) * AXBC_NOALLOC_UTILITY_MAX);
constexpr size_t kScrBytes = align_up(kScrWeight + sizeof(double) * AXBC_NOALLOC_UTILITY_MAX);

static __device__ __align__(256) unsigned char scratch[kScrBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The array's address on `device`, which the caller has made current;
// resolved on the first call there.
int scratch_of_device(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(AXBC_EINVAL, "device id out of range%s (line %d)", "", __LINE__);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(AXBC_EHIP, "hipGetSymbolAddress of the scratch array failed: %s (line %d)", hipGetErrorString(e),
                  __LINE__);
    g_scratch_ptr[device] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[device];
  return AXBC_OK;
}

}  // namespace

extern "C" int axbc_batch_utility_noalloc(axbc_ctx *c, const uint32_t *vertex, const double *weight, uint32_t count,
                                          double *out) {
  if (!c || !c->batch || !vertex || !weight || !out) return fail(AXBC_EINVAL, "bad argument%s (line %d)", "", __LINE__);
  if (count > AXBC_NOALLOC_UTILITY_MAX)
    return fail(AXBC_EINVAL, "count above the %s targets of the preallocated scratch (line %d)", "65536", __LINE__);
  for (uint32_t i = 0; i < count; ++i)
    if (vertex[i] >= c->n) return fail(AXBC_EINVAL, "vertex out of range%s (line %d)", "", __LINE__);
  HIP_TRY(hipSetDevice(c->device));
  unsigned char *base = nullptr;
  int rc;
  if ((rc = scratch_of_device(c->device, &base))) return rc;
  return batch_utility_any(c, (uint32_t *)(base + kScrVert), (double *)(base + kScrWeight),
                           (double *)(base + kScrPart), (double *)(base + kScrOut), vertex, weight, count, out);
}
