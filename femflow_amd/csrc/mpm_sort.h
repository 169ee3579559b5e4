// Host-side declarations of the binning primitives (mpm_sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mpm355.h"

namespace mpm {

size_t radix_workspace_bytes(int64_t n);
int radix_passes(uint32_t max_key);
void exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* part, uint32_t* total,
                    hipStream_t st);
void radix_sort(uint32_t* keys, uint32_t* vals, int64_t n, uint32_t max_key, void* ws,
                bool iota_values, uint32_t** kout, uint32_t** vout, hipStream_t st);
void bin_keys(const float* x, int64_t n, int dim, int res, float inv_dx, int layout, uint32_t* keys,
              hipStream_t st);

}  // namespace mpm
