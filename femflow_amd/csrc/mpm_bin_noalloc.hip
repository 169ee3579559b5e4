// mpm_bin without allocation: the same keys and stable radix sort as mpm_bin
// (mpm_capi.hip), with the sort's workspace in one file-scope device array
// instead of per-call hipMalloc/hipFree, for callers that cannot allocate.
//
// The array holds, for n <= NOALLOC_MAX_N particles: the sort's working copy
// of the keys (n words), the key/value ping-pong buffers (2n), and the
// per-tile digit histograms with their offsets (2 * max(RS_BINS * tiles, n))
// and the RS_BINS digit totals.  Each GPU holds its own copy for the life of the
// process; its address is looked up once per device (hipGetSymbolAddress)
// and only that address is handed to kernels.  Nothing in it needs to start
// at zero: every word a call reads was written earlier in the same call.
// Calls that overlap on one GPU share the array and race.
#include <mutex>
#include <string>

#include "mpm355.h"
#include "mpm_sort.h"
#include "mpm_sort_dev.h"

namespace mpm {
void set_error(const std::string& msg);

constexpr int64_t NOALLOC_MAX_N = int64_t(1) << 20;
constexpr int64_t NOALLOC_TILES = (NOALLOC_MAX_N + RS_TILE - 1) / RS_TILE;
constexpr int64_t NOALLOC_BIG =
    RS_BINS * NOALLOC_TILES > NOALLOC_MAX_N ? RS_BINS * NOALLOC_TILES : NOALLOC_MAX_N;
constexpr int64_t NOALLOC_WORDS = NOALLOC_MAX_N + 2 * NOALLOC_MAX_N + 2 * NOALLOC_BIG + RS_BINS + 64;
static_assert(NOALLOC_WORDS * 4 <= (int64_t(24) << 20), "scratch stays under 24 MiB");

static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this device's scratch array (looked up once per
// device, with that device current).  Null on failure, error set.
uint32_t* scratch_for_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    set_error("mpm_bin_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("mpm_bin_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace
}  // namespace mpm

extern "C" int32_t mpm_bin_noalloc(mpm_ctx* ctx, int32_t dim, int32_t res, double inv_dx, int32_t layout,
                                   const float* x, int64_t n, uint32_t* keys, uint32_t* perm) {
  (void)ctx;
  if ((dim != 2 && dim != 3) || res < 1 || res > 4096 || n < 0 || (n > 0 && (!x || !keys || !perm)) ||
      (layout != MPM_KEY_BLOCKED && layout != MPM_KEY_ROWMAJOR)) {
    mpm::set_error("mpm_bin_noalloc: bad argument");
    return MPM_ERR_ARG;
  }
  if (n > mpm::NOALLOC_MAX_N) {
    mpm::set_error("mpm_bin_noalloc: n = " + std::to_string(n) + " exceeds the fixed workspace's " +
                   std::to_string(mpm::NOALLOC_MAX_N) + " particles (use mpm_bin)");
    return MPM_ERR_ARG;
  }
  if (n == 0) return MPM_OK;
  const uint64_t r1 = (uint64_t)res + 1;
  uint64_t max_key;
  if (layout == MPM_KEY_ROWMAJOR) {
    max_key = dim == 3 ? r1 * r1 * r1 - 1 : r1 * r1 - 1;
  } else {
    const uint64_t nb = (r1 + 3) / 4;
    max_key = dim == 3 ? nb * nb * nb * 64 - 1 : nb * nb * 16 - 1;
  }
  if (max_key >= (1ull << 32)) {
    mpm::set_error("mpm_bin_noalloc: key space exceeds 32 bits");
    return MPM_ERR_ARG;
  }
  uint32_t* s = mpm::scratch_for_current_device();
  if (!s) return MPM_ERR_HIP;
  hipStream_t st = nullptr;
  uint32_t* tmp = s;            // n words: the sort's copy of the keys
  uint32_t* ws = s + mpm::NOALLOC_MAX_N;  // radix_sort's workspace
  mpm::bin_keys(x, n, dim, res, (float)inv_dx, layout, keys, st);
  if (hipMemcpyAsync(tmp, keys, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    mpm::set_error("mpm_bin_noalloc: copying the keys failed");
    return MPM_ERR_HIP;
  }
  uint32_t *ks, *vs;
  mpm::radix_sort(tmp, perm, n, (uint32_t)max_key, ws, true, &ks, &vs, st);
  if (vs != perm) hipMemcpyAsync(perm, vs, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
  const hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    mpm::set_error(std::string("mpm_bin_noalloc: ") + hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}
