// Wave64 primitives for gfx950 (CDNA4).  All helpers must be called with
// every lane of the wave active (they use cross-lane shuffles).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axbc {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Inclusive prefix sum over the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// Owner search for load-balanced edge rounds: given the inclusive degree
// prefix `incl` held one value per lane (non-decreasing), return the number
// of lanes j with incl[j] <= e, i.e. the lane whose range contains edge e.
__device__ __forceinline__ int wave_owner(uint32_t incl, uint32_t e) {
  int pos = 0;
#pragma unroll
  for (int step = kWave / 2; step > 0; step >>= 1) {
    uint32_t v = __shfl(incl, pos + step - 1);
    if (v <= e) pos += step;
  }
  return pos;
}

// Append `v` from every lane with `pred` to queue[*tail ...] with one atomic
// per wave.  Order within the wave follows lane order.
__device__ __forceinline__ void wave_append(bool pred, uint32_t v, uint32_t *queue,
                                            uint32_t *tail) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(tail, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (pred) queue[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

}  // namespace axbc
