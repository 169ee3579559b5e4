// Host-side engine object behind the mpm_ctx handle of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mpm355.h"
#include "mpm_engine_kernels.h"

namespace mpm {

void set_error(const std::string& msg);

// A captured segment fixes the particle count (launch grids, sort sizes),
// the segment length, whether it rebins first, which state buffer is current,
// the sums destination, and whether it continues the run's previous segment
// (cont) and ends the run (fin); any of them changing needs another graph.
struct GraphKey {
  int64_t n;
  int len, parity, bin;
  uintptr_t sums;
  int nsteps, cont, fin;
  bool operator<(const GraphKey& o) const {
    return std::tie(n, len, parity, bin, sums, nsteps, cont, fin) <
           std::tie(o.n, o.len, o.parity, o.bin, o.sums, o.nsteps, o.cont, o.fin);
  }
};
// Captured graphs kept per context; beyond this the cache is emptied first.
constexpr size_t MAX_GRAPHS = 16;

struct CorrExtract {
  uint16_t* X;      // rows x 352 BF16 features
  uint32_t* nodes;  // global node index (block * 64 + node in block)
  float* vel;       // rows x 3 updated grid velocity
  int64_t max_rows;
};

// Per-substep state output (caller order): slot k of each array receives the
// state after substep k.  grids (optional, single-scene engines): slot k of
// 7 (res+1)^3 floats receives substep k's dense grids, the P2G momentum
// ((res+1)^3 x 3), the P2G mass ((res+1)^3) and the updated velocity
// ((res+1)^3 x 3), zero at nodes no particle reached.
struct StateSink {
  float *x, *v, *F, *C, *Jp;
  float* grids;
};

struct TimedLaunch {
  int kind;  // 0 P2G, 1 G2P2G, 2 G2P, 3 grid update
  hipEvent_t e0, e1;
};

struct Engine {
  // configuration
  int device = 0, num_cus = 256;
  int64_t n_max = 0, n = 0, total_blocks = 0;
  int nmat = 1;
  int chunk = 512, sort_every = 48, grid_mult = 8;  // the measured headline optimum
  int spread = 9;  // scatter start offsets per cell (3 or 9, see transfer_kernel)
  // fused passes take chunks from per-XCD ticket queues (set for engines of
  // 8 or more scenes, whose many uneven chunks the queue balances)
  bool dyn_sched = false;
  // set while a segment with per-substep state output is enqueued: the fused
  // pass then writes v and C every substep
  bool store_vc = false;
  bool use_graph = true, timing = false;
  // substeps run since the particles were last binned; -1: not binned (after
  // an upload or a rebin request), so the next segment rebins first
  int since_bin = -1;
  int last_list = 0;  // active-block list the last substep's P2G appended to
  std::vector<SceneDev> scenes_host;
  // streams and graphs
  hipStream_t user_stream = nullptr, own_stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::vector<hipEvent_t> tevents;
  size_t tnext = 0;
  std::vector<TimedLaunch> tpairs;
  // device memory
  std::vector<void*> allocs;
  EngineState st[2] = {};
  int cur = 0;
  SceneDev* d_scenes = nullptr;
  MatParams* d_mats = nullptr;
  int* d_block_scene = nullptr;
  float4* d_acc = nullptr;
  float4* d_vel = nullptr;
  uint32_t *d_flags = nullptr, *d_list = nullptr, *d_bstart = nullptr, *d_bend = nullptr;
  uint32_t* d_counters = nullptr;
  unsigned long long *d_oob = nullptr, *d_fallback = nullptr;
  int4* d_chunks = nullptr;
  uint32_t *d_nchunks = nullptr, *d_keys = nullptr, *d_vals = nullptr,
           *d_bocc = nullptr, *d_part = nullptr;
  uint32_t* d_sorted_keys = nullptr;
  // sort permutation of the last rebin, consumed by the next P2G (which
  // moves the particles into binned order); null once consumed
  const uint32_t* pending_perm = nullptr;
  uint32_t* d_hard = nullptr;
  uint32_t *d_cand_list = nullptr, *d_cand_flag = nullptr;
  uint2* d_cand_src = nullptr;
  float4* d_stage = nullptr;
  // learned corrector
  bool corr_on = false;
  float corr_mref = 1.f, corr_vref = 1.f;
  float4* d_pm = nullptr;
  float4* d_cellf = nullptr;
  float* d_corr_params = nullptr;
  uint16_t* d_corr_packed = nullptr;
  uint32_t* d_corr_count = nullptr;
  // sticky half-space colliders (all scenes, at most MAX_PLANES)
  std::vector<PlaneDev> planes_host;
  PlaneDev* d_planes = nullptr;
  uint32_t stage_cap = 0;
  uint32_t *d_bfirst = nullptr, *d_bcount = nullptr;
  void* d_ws = nullptr;
  size_t ws_bytes = 0;

  ~Engine();
  void release();
  int alloc(void** p, size_t bytes);
  int init(int64_t n_max, int nscenes, const mpm_scene* sc, int nmat, const mpm_material* mats);
  EngineArgs args(double* sums, int max_steps, int sel) const;
  int upload(const mpm_particles* p, const int32_t* scene_ids);
  int download(mpm_particles* p);
  // rebin; also starts the segment of len substeps (counters, sums index)
  void enqueue_bin(hipStream_t s, int len, int with_sums);
  void enqueue_segment(int len, bool bin, hipStream_t s, double* sums, int max_steps, float* frames,
                       int64_t step0, const StateSink* sink, bool cont, bool fin);
  // destroy every captured graph (after anything they baked in changes)
  void drop_graphs();
  int active_blocks(int64_t* out);
  void enqueue_frame(hipStream_t s, float* frames, int64_t step);
  void enqueue_state(hipStream_t s, const StateSink* sink, int64_t step);
  int positions(float* x);
  // separate_deferred: finish the FP64-deferred particles in their own
  // launch right after the transfer (needed where something reads the
  // particles before the grid update, and after the last G2P); otherwise
  // the following launch_grid(..., deferred = 1) finishes them
  void launch_transfer(int mode, hipStream_t s, const EngineArgs& a, bool separate_deferred = true);
  void launch_grid(hipStream_t s, const EngineArgs& a, int with_sums, int len, int k, int deferred = 0);
  void fill_u32(hipStream_t s, uint32_t* p, int64_t count, uint32_t v);
  void launch_cellf(hipStream_t s, const EngineArgs& a);
  // corrector after a grid update; ex != null: write features instead (sampling)
  void launch_corrector(hipStream_t s, const EngineArgs& a, const struct CorrExtract* ex);
  int set_corrector(const float* params, float m_ref, float v_ref, int enable);
  int corrector_sample(uint16_t* X, uint32_t* nodes, float* vel, int64_t max_rows, int64_t* rows);
  int grid_velocity(const uint32_t* nodes, int64_t count, float* out);
  int set_planes(int scene, int count, const double* points, const double* normals, int quirks);
  hipEvent_t tevent();
  int run(int nsteps, double* sums, float* frames = nullptr, const StateSink* sink = nullptr);
  int take_counter(int which, int64_t* out);
  int particle_sums(double* out);
  int kernel_times(double* ms, int64_t* counts);
};

}  // namespace mpm
