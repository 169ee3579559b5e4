// Device-side pieces of the stable LSD radix sort (mpm_sort.hip) that a key
// kernel elsewhere can share: the tile shape and the tile histogram of the
// first pass, so the kernel that computes the keys also counts their first
// digit and the sort starts at that pass's scan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
// 11-bit digits: the 22-bit cell keys of a 128^3 grid sort in two passes
// (three with 8-bit digits); the scatter's per-wave digit counters take
// 4 x 2048 words of LDS.
constexpr int RS_BITS = 11;
constexpr int RS_BINS = 1 << RS_BITS;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Lanes (among `valid`) whose digit equals this lane's.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t m = valid;
#pragma unroll
  for (int b = 0; b < RS_BITS; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

// Histogram of digit (key >> shift) & (RS_BINS - 1) over this workgroup's
// tile, whose thread t holds the keys of tile slots r * RS_THREADS + t
// (r < RS_ITEMS; slots at or past n are not counted), written tile-major
// (one contiguous row per tile) to hist[blockIdx.x * RS_BINS + d].  h:
// RS_BINS words of LDS.  Every thread of the workgroup calls it.
__device__ __forceinline__ void rs_tile_histogram(const uint32_t (&k)[RS_ITEMS], int64_t n, int shift,
                                                  uint32_t* h, uint32_t* __restrict__ hist) {
  const int t = threadIdx.x;
  for (int i = t; i < RS_BINS; i += RS_THREADS) h[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const unsigned lane = __lane_id();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + (int64_t)r * RS_THREADS + t;
    const bool valid = idx < n;
    const uint32_t d = valid ? (k[r] >> shift) & (uint32_t)(RS_BINS - 1) : 0u;
    const uint64_t vm = __ballot(valid);
    const uint64_t m = match_digit(d, vm);
    if (valid && lane == 63u - (unsigned)__clzll(m)) atomicAdd(&h[d], (uint32_t)__popcll(m));
  }
  __syncthreads();
  for (int i = t; i < RS_BINS; i += RS_THREADS) hist[(int64_t)blockIdx.x * RS_BINS + i] = h[i];
}

}  // namespace mpm
