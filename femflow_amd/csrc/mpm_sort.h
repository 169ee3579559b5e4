// Host-side declarations of the binning primitives (mpm_sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mpm355.h"

namespace mpm {

size_t radix_workspace_bytes(int64_t n);
int radix_passes(uint32_t max_key);
// Chunk list of a binning: occupied block b (bocc[b] != 0, particles
// bstart[b] .. bend[b]) gets ceil(count / chunk) chunks of nearly equal size
// (start, count, b, block_scene[b]), numbered after those of the blocks
// before it; bfirst[b]
// and bcount[b] are its first chunk and chunk count (0 for empty blocks);
// *nchunks = the total.  part: chunk_list_part_words(nblocks) words.
size_t chunk_list_part_words(int64_t nblocks);
void chunk_list(const uint32_t* bocc, const uint32_t* bstart, const uint32_t* bend, const int* block_scene,
                int64_t nblocks, int chunk, uint32_t* part, int4* chunks, uint32_t* bfirst, uint32_t* bcount,
                uint32_t* nchunks, hipStream_t st);
// Tiles of the sort (one workgroup each) and the first pass's tile
// histogram inside the workspace, for a key kernel that also counts the
// first digit (rs_tile_histogram, mpm_sort_dev.h) and passes
// first_hist_done.
int64_t radix_tiles(int64_t n);
uint32_t* radix_first_hist(void* ws, int64_t n);
void radix_sort(uint32_t* keys, uint32_t* vals, int64_t n, uint32_t max_key, void* ws,
                bool iota_values, uint32_t** kout, uint32_t** vout, hipStream_t st,
                bool first_hist_done = false);
void bin_keys(const float* x, int64_t n, int dim, int res, float inv_dx, int layout, uint32_t* keys,
              hipStream_t st);

}  // namespace mpm
