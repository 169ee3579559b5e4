// Cell binning: key generation, a stable LSD radix sort on 32-bit keys
// (11-bit digits; three launches per pass: tile histograms, a per-digit scan
// over tiles, the stable scatter), and the chunk list (a two-launch scan of
// per-block chunk counts) that gives the engine its work items for the
// transfer kernels.
//
// Stability: inside a tile each wave owns a contiguous run of keys and walks
// it in index order; ranks among equal digits come from a 64-lane match mask
// built with one ballot per digit bit plus per-wave running counters, and
// waves are ordered by a prefix over their counters.  Tiles are ordered by
// the per-digit exclusive scan over tiles of the per-tile histograms.  Each
// pass is therefore a stable counting sort, and the whole sort equals a
// stable sort by key.
#include "mpm_sort.h"
#include "mpm_sort_dev.h"

namespace mpm {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Per-tile digit histogram, written tile-major: hist[tile * RS_BINS + d].
__global__ __launch_bounds__(RS_THREADS) void rs_upsweep(const uint32_t* __restrict__ keys,
                                                         int64_t n, int shift,
                                                         uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[RS_BINS];
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  // every key of the thread in flight at once: one memory round trip per
  // tile rather than one per few rounds
  uint32_t k[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + (int64_t)r * RS_THREADS + threadIdx.x;
    k[r] = idx < n ? keys[idx] : 0u;
  }
  rs_tile_histogram(k, n, shift, h, hist);
}

// Exclusive prefix sum over the 256 threads of a workgroup (one value each):
// wave-level inclusive scans by lane shuffles, then the wave totals.
__device__ __forceinline__ uint32_t excl_scan_256(uint32_t v, uint32_t* wave_tot) {
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, (unsigned)off);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wave_tot[w] = incl;
  __syncthreads();
  uint32_t base = 0;
#pragma unroll
  for (int i = 0; i < RS_WAVES; ++i) base += i < w ? wave_tot[i] : 0u;
  return base + incl - v;
}

// Per digit d: the exclusive prefix over tiles of the tile-major histogram
// hist[tile * RS_BINS + d] (written to offs in the same layout) and the
// digit's total (dtot[d]); the scatter adds the exclusive prefix of the
// totals over digits.  A workgroup takes 64 digits (lane = digit, so every
// read and write is one contiguous 256-byte row piece) and splits the tiles
// into DS_WAVES contiguous ranges, one per wave: each wave sums its range,
// the waves' sums are prefixed in LDS, and each wave then writes its range's
// running prefix.  Sixteen waves keep each wave's serial walk short (16
// tiles for 2^20 keys).
constexpr int DS_WAVES = 16;
__global__ __launch_bounds__(DS_WAVES * 64) void rs_digit_scan(const uint32_t* __restrict__ hist, int ntiles,
                                                              uint32_t* __restrict__ offs,
                                                              uint32_t* __restrict__ dtot) {
  __shared__ uint32_t part[DS_WAVES][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int d = blockIdx.x * 64 + lane;
  const int per = (ntiles + DS_WAVES - 1) / DS_WAVES;
  const int t0 = w * per < ntiles ? w * per : ntiles;
  const int t1 = t0 + per < ntiles ? t0 + per : ntiles;
  uint32_t sum = 0;
#pragma unroll 16
  for (int tl = t0; tl < t1; ++tl) sum += hist[(int64_t)tl * RS_BINS + d];
  part[w][lane] = sum;
  __syncthreads();
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < DS_WAVES; ++i) run += i < w ? part[i][lane] : 0u;
  if (w == DS_WAVES - 1) dtot[d] = run + sum;
#pragma unroll 16
  for (int tl = t0; tl < t1; ++tl) {
    const uint32_t v = hist[(int64_t)tl * RS_BINS + d];
    offs[(int64_t)tl * RS_BINS + d] = run;
    run += v;
  }
}

__global__ __launch_bounds__(RS_THREADS) void rs_downsweep(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ offs,
    const uint32_t* __restrict__ dtot, int ntiles, int iota_values) {
  __shared__ uint32_t cnt[RS_WAVES][RS_BINS];
  __shared__ uint32_t goff[RS_BINS];
  __shared__ uint32_t wt[RS_WAVES];
  const int t = threadIdx.x;
  const int w = t >> 6;
  const unsigned lane = __lane_id();
  constexpr int DPT = RS_BINS / RS_THREADS;  // digits per thread
#pragma unroll
  for (int i = 0; i < RS_WAVES; ++i)
    for (int d = t; d < RS_BINS; d += RS_THREADS) cnt[i][d] = 0;
  // global start of digit d in this tile: keys of smaller digits (all tiles)
  // plus keys of digit d in earlier tiles; thread t takes digits
  // t * DPT .. t * DPT + DPT - 1
  {
    uint32_t tot[DPT], sum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      tot[j] = dtot[t * DPT + j];
      sum += tot[j];
    }
    uint32_t run = excl_scan_256(sum, wt);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = t * DPT + j;
      goff[d] = run + offs[(int64_t)blockIdx.x * RS_BINS + d];
      run += tot[j];
    }
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * 64 * RS_ITEMS;
  uint32_t k[RS_ITEMS], val[RS_ITEMS];
  uint32_t rank[RS_ITEMS];
  const uint64_t lt = lanemask_lt();
  // keys and values of the wave's run loaded together, ahead of the ranking
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + r * 64 + lane;
    const bool valid = idx < n;
    k[r] = valid ? kin[idx] : 0u;
    val[r] = iota_values ? (uint32_t)idx : (valid ? vin[idx] : 0u);
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + r * 64 + lane;
    const bool valid = idx < n;
    const uint32_t d = (k[r] >> shift) & (uint32_t)(RS_BINS - 1);
    const uint64_t vm = __ballot(valid);
    const uint64_t m = match_digit(d, vm);
    uint32_t prev = 0;
    if (valid) prev = cnt[w][d];
    rank[r] = prev + (uint32_t)__popcll(m & lt);
    if (valid && lane == 63u - (unsigned)__clzll(m)) cnt[w][d] = prev + (uint32_t)__popcll(m);
  }
  __syncthreads();
  for (int d = t; d < RS_BINS; d += RS_THREADS) {
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < RS_WAVES; ++i) {
      const uint32_t c = cnt[i][d];
      cnt[i][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + r * 64 + lane;
    if (idx < n) {
      const uint32_t d = (k[r] >> shift) & (uint32_t)(RS_BINS - 1);
      const uint32_t dst = goff[d] + cnt[w][d] + rank[r];
      kout[dst] = k[r];
      vout[dst] = val[r];
    }
  }
}

// ---------------------------------------------------------------- scan ----
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const uint32_t a = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const uint32_t incl = sh[t];
  *total = sh[SCAN_THREADS - 1];
  __syncthreads();
  return incl - v;
}

// ------------------------------------------------------- chunk list ----
// Chunk list of the new binning, numbered over blocks: block b holds
// ceil(count_b / chunk) chunks of nearly equal size, numbered after those
// of the blocks before it.  Two launches over tiles of SCAN_THREADS blocks, one block per thread
// (coalesced reads, a few hundred tiles for a 128^3 grid): per-tile chunk
// counts, then each tile adds the counts of the tiles before it, scans its
// own blocks and writes their chunks, first chunk and chunk count.
__device__ __forceinline__ uint32_t block_chunks(const uint32_t* __restrict__ bocc,
                                                 const uint32_t* __restrict__ bstart,
                                                 const uint32_t* __restrict__ bend, int64_t b, int chunk) {
  return bocc[b] ? (bend[b] - bstart[b] + (uint32_t)chunk - 1u) / (uint32_t)chunk : 0u;
}

__global__ __launch_bounds__(SCAN_THREADS) void block_tile_kernel(const uint32_t* __restrict__ bocc,
                                                                  const uint32_t* __restrict__ bstart,
                                                                  const uint32_t* __restrict__ bend,
                                                                  int64_t nblocks, int chunk,
                                                                  uint32_t* __restrict__ part) {
  __shared__ uint32_t sh[SCAN_THREADS];
  const int64_t b = (int64_t)blockIdx.x * SCAN_THREADS + threadIdx.x;
  const uint32_t c = b < nblocks ? block_chunks(bocc, bstart, bend, b, chunk) : 0u;
  uint32_t tot;
  block_exclusive_scan(c, sh, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void block_chunks_kernel(
    const uint32_t* __restrict__ bocc, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
    const int* __restrict__ block_scene, int64_t nblocks, int chunk, const uint32_t* __restrict__ part,
    int4* __restrict__ chunks, uint32_t* __restrict__ bfirst, uint32_t* __restrict__ bcount,
    uint32_t* __restrict__ nchunks) {
  __shared__ uint32_t sh[SCAN_THREADS];
  uint32_t before = 0;
  for (uint32_t q = threadIdx.x; q < blockIdx.x; q += SCAN_THREADS) before += part[q];
  uint32_t tot_before;
  block_exclusive_scan(before, sh, &tot_before);
  const int64_t b = (int64_t)blockIdx.x * SCAN_THREADS + threadIdx.x;
  const uint32_t c = b < nblocks ? block_chunks(bocc, bstart, bend, b, chunk) : 0u;
  uint32_t tot;
  const uint32_t run = block_exclusive_scan(c, sh, &tot) + tot_before;
  if (b < nblocks) {
    bfirst[b] = run;
    bcount[b] = c;
    if (c) {
      const uint32_t p0 = bstart[b], cnt = bend[b] - p0;
      const int sc = block_scene[b];
      // c chunks of (nearly) equal size rather than full ones and a
      // remainder: less spread in per-chunk work, a shorter tail
      for (uint32_t j = 0; j < c; ++j) {
        const uint32_t s0 = (uint32_t)(((uint64_t)cnt * j) / c), s1 = (uint32_t)(((uint64_t)cnt * (j + 1)) / c);
        chunks[run + j] = make_int4((int)(p0 + s0), (int)(s1 - s0), (int)b, sc);
      }
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *nchunks = run + c;
}

// ---------------------------------------------------------------- keys ----
// base = trunc(fl32(fl32(x*inv_dx) - 0.5)), clamped to [0, res] (the same
// rounding the reference's (x*inv_dx - 0.5).astype(int64) sees on FP32 x).
__device__ __forceinline__ int cell_coord(float x, float inv_dx, int res) {
  const float s = __fmul_rn(x, inv_dx);
  const float t = __fsub_rn(s, 0.5f);
  int b;
  if (!(t > -1.0f)) b = 0;            // also NaN
  else if (!(t < (float)res)) b = res;
  else b = (int)t;                    // truncation toward zero
  return b < 0 ? 0 : (b > res ? res : b);
}

__global__ void bin_keys_kernel(const float* __restrict__ x, int64_t n, int dim, int res,
                                float inv_dx, int layout, uint32_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int b[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) b[d] = cell_coord(x[i * dim + d], inv_dx, res);
  const uint32_t r1 = (uint32_t)res + 1u;
  uint32_t key;
  if (layout == MPM_KEY_ROWMAJOR) {
    key = (uint32_t)b[0];
    for (int d = 1; d < dim; ++d) key = key * r1 + (uint32_t)b[d];
  } else {
    const uint32_t nb = (r1 + 3u) / 4u;
    uint32_t bid = (uint32_t)(b[0] >> 2), cid = (uint32_t)(b[0] & 3);
    for (int d = 1; d < dim; ++d) {
      bid = bid * nb + (uint32_t)(b[d] >> 2);
      cid = cid * 4u + (uint32_t)(b[d] & 3);
    }
    key = bid * (dim == 3 ? 64u : 16u) + cid;
  }
  keys[i] = key;
}

// ------------------------------------------------------------ host side ----
static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

size_t radix_workspace_bytes(int64_t n) {
  const int64_t ntiles = cdiv(n > 0 ? n : 1, RS_TILE);
  const int64_t hist = (int64_t)RS_BINS * ntiles;
  // keys/vals ping-pong (4 arrays of n) + hist + offs + digit totals
  return sizeof(uint32_t) * (4 * (size_t)n + 2 * (size_t)(hist > n ? hist : n) + RS_BINS + 64);
}

size_t chunk_list_part_words(int64_t nblocks) { return (size_t)cdiv(nblocks, SCAN_THREADS) + 1; }

void chunk_list(const uint32_t* bocc, const uint32_t* bstart, const uint32_t* bend, const int* block_scene,
                int64_t nblocks, int chunk, uint32_t* part, int4* chunks, uint32_t* bfirst, uint32_t* bcount,
                uint32_t* nchunks, hipStream_t st) {
  if (nblocks <= 0) return;
  const unsigned tiles = (unsigned)cdiv(nblocks, SCAN_THREADS);
  hipLaunchKernelGGL(block_tile_kernel, dim3(tiles), dim3(SCAN_THREADS), 0, st, bocc, bstart, bend, nblocks,
                     chunk, part);
  hipLaunchKernelGGL(block_chunks_kernel, dim3(tiles), dim3(SCAN_THREADS), 0, st, bocc, bstart, bend,
                     block_scene, nblocks, chunk, part, chunks, bfirst, bcount, nchunks);
}

int radix_passes(uint32_t max_key) {
  int bits = 0;
  while (bits < 32 && (max_key >> bits) != 0u) ++bits;
  return bits == 0 ? 1 : (bits + RS_BITS - 1) / RS_BITS;
}

int64_t radix_tiles(int64_t n) { return cdiv(n > 0 ? n : 1, RS_TILE); }

uint32_t* radix_first_hist(void* ws, int64_t n) { return (uint32_t*)ws + 2 * n; }

// Stable sort of keys[0..n) carrying values; result in (*kout, *vout), which
// point into either the input arrays or the workspace.  iota_values: the
// input values are 0..n-1 (not read).  first_hist_done: the first pass's
// tile histogram is already in the workspace (radix_first_hist).
void radix_sort(uint32_t* keys, uint32_t* vals, int64_t n, uint32_t max_key, void* ws,
                bool iota_values, uint32_t** kout, uint32_t** vout, hipStream_t st, bool first_hist_done) {
  const int64_t ntiles = radix_tiles(n);
  const int64_t hist_n = (int64_t)RS_BINS * ntiles;
  const int64_t big = hist_n > n ? hist_n : n;
  uint32_t* w = (uint32_t*)ws;
  uint32_t* k2 = w;
  uint32_t* v2 = w + n;
  uint32_t* hist = radix_first_hist(ws, n);
  uint32_t* offs = hist + big;
  uint32_t* dtot = offs + big;
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  const int passes = radix_passes(max_key);
  for (int p = 0; p < passes; ++p) {
    const int shift = RS_BITS * p;
    if (p > 0 || !first_hist_done)
      hipLaunchKernelGGL(rs_upsweep, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, ka, n, shift, hist);
    hipLaunchKernelGGL(rs_digit_scan, dim3(RS_BINS / 64), dim3(DS_WAVES * 64), 0, st, hist, (int)ntiles, offs,
                       dtot);
    hipLaunchKernelGGL(rs_downsweep, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, ka, va, kb, vb,
                       n, shift, offs, dtot, (int)ntiles, (iota_values && p == 0) ? 1 : 0);
    uint32_t* t;
    t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  *kout = ka;
  *vout = va;
}

void bin_keys(const float* x, int64_t n, int dim, int res, float inv_dx, int layout, uint32_t* keys,
              hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(bin_keys_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, x, n, dim, res,
                     inv_dx, layout, keys);
}

}  // namespace mpm
