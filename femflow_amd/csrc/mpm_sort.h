// Host-side declarations of the binning primitives (mpm_sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mpm355.h"

namespace mpm {

size_t radix_workspace_bytes(int64_t n);
int radix_passes(uint32_t max_key);
// Chunk numbering of the binned particles: idx[i] = number of chunk starts
// before sorted key i (a start: offset from its block's first key is a
// multiple of chunk); *nchunks = their total.
void chunk_index(const uint32_t* keys, const uint32_t* bstart, int64_t n, int chunk, uint32_t* idx,
                 uint32_t* part, uint32_t* nchunks, hipStream_t st);
void radix_sort(uint32_t* keys, uint32_t* vals, int64_t n, uint32_t max_key, void* ws,
                bool iota_values, uint32_t** kout, uint32_t** vout, hipStream_t st);
void bin_keys(const float* x, int64_t n, int dim, int res, float inv_dx, int layout, uint32_t* keys,
              hipStream_t st);

}  // namespace mpm
