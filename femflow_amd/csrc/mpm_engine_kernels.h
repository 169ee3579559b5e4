// Device-side data structures and kernels of the particle-resident engine.
//
// Grid storage is block-sparse and block-major: every scene's (R+1)^3 node
// lattice is cut into 4x4x4-node blocks (nb = ceil((R+1)/4) per axis), the
// blocks of all scenes are concatenated, and a node lives at
//   (block_base + (bi*nb + bj)*nb + bk) * 64 + (ii*16 + jj*4 + kk)
// as one float4 (momentum or velocity xyz, mass).  Particles are binned by
// (global block, cell in block) and processed in chunks of at most CHUNK
// particles from one block.  A chunk's grid footprint is staged in an LDS
// arena of 8x8x8 nodes whose origin is the block origin minus one node, so
// particles that have drifted up to one cell out of their block (the grid
// update clamps |v| to 0.9 dx/dt per axis) still land in it; anything further
// takes the global-atomic path, so results never depend on binning.
#pragma once
#include "mpm_math.h"

namespace mpm {

constexpr int ENG_THREADS = 256;
constexpr int ARENA = 8;  // nodes per axis
constexpr int ARENA_NODES = ARENA * ARENA * ARENA;
// Staged part of a chunk's arena: nodes 0..6 per axis.  A particle inside
// its block reaches arena nodes 0..6 only; the eighth layer is reached by
// particles that drifted out of the block since binning and goes through
// global atomics.
constexpr int STAGE = 7;

// Staging record of one chunk.  Per axis the seven staged nodes split by
// the block they belong to: node 0 lies in the block before the chunk's own
// (part 0), nodes 1..4 in its own (part 1), nodes 5..6 in the block after
// (part 2).  The record is cut into 27 part boxes like an 8x8x8 array cut at
// 2 and 6 per axis (box extents 2, 4, 2: never below 8 slots, always a
// multiple of 8), in part-major order; a box holds its part's nodes
// (extents 1, 4, 2) row-major in its first slots.  So the grid update of
// one block reads each source chunk's contribution to it as one contiguous
// run of slots in its lanes' order, and every run starts on a 128-byte line
// (8 slots): it fetches no line of another box.  512 slots per record for
// 343 nodes; the unused slots are never written or read.
constexpr int STAGE_REC = 512;

// Slot of staged node (x0, x1, x2) (each 0..6) in its chunk's record.  The
// per-part constants are read from nibbles (part p at bits 4p): first node
// s (0, 1, 5), extent z (1, 4, 2), first box coordinate t (0, 2, 6), box
// extent e (2, 4, 2).
__device__ __forceinline__ int stage_slot(int x0, int x1, int x2) {
  const int p0 = x0 == 0 ? 0 : (x0 < 5 ? 1 : 2), p1 = x1 == 0 ? 0 : (x1 < 5 ? 1 : 2),
            p2 = x2 == 0 ? 0 : (x2 < 5 ? 1 : 2);
  const int s0 = (0x510 >> (4 * p0)) & 15, s1 = (0x510 >> (4 * p1)) & 15, s2 = (0x510 >> (4 * p2)) & 15;
  const int z1 = (0x241 >> (4 * p1)) & 15, z2 = (0x241 >> (4 * p2)) & 15;
  const int t0 = (0x620 >> (4 * p0)) & 15, t1 = (0x620 >> (4 * p1)) & 15, t2 = (0x620 >> (4 * p2)) & 15;
  const int e0 = (0x242 >> (4 * p0)) & 15, e1 = (0x242 >> (4 * p1)) & 15;
  return t0 * 64 + e0 * t1 * 8 + e0 * e1 * t2 + ((x0 - s0) * z1 + (x1 - s1)) * z2 + (x2 - s2);
}
constexpr int MAX_SCENES = 4096;

struct SceneDev {
  int res, nb, block_base, boundary, bound, quirks;
  float inv_dx, dx, dt, gravity, vol, v_allowed;
};

// Particle state in tiles of 64 slots: inside a tile, component d of a
// K-component field (x, v: K = 3; F, C: K = 9 row-major) occupies 64
// consecutive floats, so one wave's access to one component of 64 slots is
// contiguous, while one particle's components sit at a fixed 256-byte stride
// from a single address.  Scalar fields are plain arrays.  Allocations are
// padded to a whole number of tiles.
constexpr int PTILE = 64;
__host__ __device__ __forceinline__ int64_t tix(int64_t i, int K, int d) {
  return ((i >> 6) * K + d) * PTILE + (i & 63);
}

struct EngineState {
  float* x;
  float* v;
  float* F;
  float* C;
  float* Jp;
  float* mass;
  float* mu0;
  float* lam0;
  uint32_t* attr;   // material id | scene << 8
  uint32_t* order;  // caller index of this slot
};

struct EngineArgs {
  EngineState s;
  // The first P2G after a rebin also moves the particles into binned order:
  // slot p of s is read from slot perm[p] of src (perm null: no move).
  EngineState src;
  const uint32_t* perm;
  const SceneDev* scenes;
  const MatParams* mats;
  int nmat;
  float4* acc;
  float4* vel;
  uint32_t* flags;
  // The grid update's work list is the candidate blocks -- the 27
  // neighbours of every block that holds particles, built at each rebin:
  // every node a chunk's LDS arena reaches lies in them -- followed by the
  // blocks marked during the substep by the paths that reach past the arena
  // (global-atomic scatter of drifted particles, the deferred FP64 pass).
  const uint32_t* cand_list;
  const uint32_t* cand_count;
  const uint2* cand_src;  // per candidate entry: (first chunk, chunk count) of its 27 neighbours, 32 slots
  const uint32_t* cand_flag;  // == *cand_gen: the block is a candidate of the current binning
  const uint32_t* cand_gen;   // binning generation (counters[5])
  uint32_t* list;        // blocks marked during this substep (not candidates)
  uint32_t* list_count;  // their number
  uint32_t* other_count; // the other list's count (cleared by the grid update)
  uint32_t* counters;  // [0], [1] marked-block list counts, [2] sums step base,
                       // [3], [6] deferred particles (by substep parity; bit 31:
                       // finished), [4] candidate blocks, [5] binning
                       // generation, [7] grid-update start ticket,
                       // [8..15] per-XCD chunk queues of the fused pass
  int store_vc;          // fused pass: write v and C every substep (per-substep state output)
  uint32_t* hard_list;   // deferred particles: index << 2 | plasticity pending | no P2G
  uint32_t* hard_count;  // this substep's deferred count (counters[3] or [6] by parity)
  uint32_t* other_hard;  // the other parity's count (cleared by the grid update)
  uint32_t* defer_ticket;  // counters[7]: start order of the grid update's workgroups
  const int4* chunks;  // start, count, global block, scene (one per occupied block)
  const uint32_t* keys;  // binned (sorted) cell keys of the current order
  const uint32_t* nchunks;
  const int* block_scene;
  int nscenes;        // scene table length (1: every block is scene 0's)
  uint32_t list_cap;  // entries each active-block list has room for
  unsigned long long* oob;
  unsigned long long* fallback;
  double* sums;
  int max_steps;
  // Per-chunk P2G results: chunk c < stage_cap writes its whole 8x8x8 arena
  // (FP32 momentum + mass, arena order) to stage[c * 512 ...] with plain
  // stores; the grid update sums, for each node, the arenas of the chunks of
  // the blocks around it (first chunk and chunk count per block).  Chunks at
  // or beyond stage_cap add into `acc` with global atomics instead.
  float4* stage;
  uint32_t stage_cap;
  const uint32_t* bfirst;
  const uint32_t* bcount;
  // Learned-corrector features (null when the corrector is off): the grid
  // update writes each active node's P2G (momentum, mass) to pm; cellf holds,
  // per node, the sums over the particles whose stencil base is that node of
  // (m, m F[0..8]) as three float4 (m, mF00, mF01, mF02), (mF10, mF11,
  // mF12, mF20), (mF21, mF22, 0, 0).
  float4* pm;
  float4* cellf;
  // Sticky half-space colliders (null / 0 when none): see PlaneDev.
  const PlaneDev* planes;
  int nplanes;
  // Dense grid record of this substep (null: none): scene 0's P2G momentum
  // (nn x 3), P2G mass (nn) and updated velocity (nn x 3), nn = (res+1)^3.
  float* grid_rec;
};

__device__ __forceinline__ size_t node_addr(const SceneDev& sc, int i, int j, int k) {
  const int blk = sc.block_base + ((i >> 2) * sc.nb + (j >> 2)) * sc.nb + (k >> 2);
  return (size_t)blk * 64 + (size_t)(((i & 3) << 4) | ((j & 3) << 2) | (k & 3));
}

// Stencil validity and (base, fx) for one particle (three_d/p2g.py:50-53 with
// the bounds test of utils.py:138-150).
__device__ __forceinline__ bool locate3(const float* xp, const SceneDev& sc, int base[3],
                                        float fx[3]) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float s = xp[d] * sc.inv_dx;
    const float t = s - 0.5f;
    ok = ok && (t > -1.0f) && (t < (float)(sc.res - 1));
    base[d] = ok ? (int)t : 0;
    fx[d] = s - (float)base[d];
  }
  return ok;
}

// First staged chunk and chunk count of block (S0, S1, S2) of scene sc
// (0, 0 outside the scene's blocks), clamped to the staging capacity.
__device__ __forceinline__ uint2 neighbour_chunks(const EngineArgs& a, const SceneDev& sc, int S0, int S1,
                                                  int S2) {
  const int nb = sc.nb;
  uint32_t c0 = 0, cn = 0;
  if (S0 >= 0 && S0 < nb && S1 >= 0 && S1 < nb && S2 >= 0 && S2 < nb) {
    const uint32_t S = (uint32_t)(sc.block_base + (S0 * nb + S1) * nb + S2);
    // both loads in flight together (bfirst is stale where bcount is 0)
    cn = a.bcount[S];
    c0 = a.bfirst[S];
    c0 = cn ? c0 : 0u;
    cn = c0 + cn <= a.stage_cap ? cn : (c0 < a.stage_cap ? a.stage_cap - c0 : 0u);
  }
  return make_uint2(c0, cn);
}

// Block g must be in this substep's grid update: candidates always are.
__device__ __forceinline__ void mark_active(const EngineArgs& a, uint32_t g) {
  if (a.cand_flag[g] == *a.cand_gen) return;
  if (__hip_atomic_load(a.flags + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
    if (atomicOr(a.flags + g, 1u) == 0u) {
      const uint32_t idx = atomicAdd(a.list_count, 1u);
      a.list[idx] = g;
    }
  }
}

// Entry b of the grid update's work list (candidates, then marked blocks).
__device__ __forceinline__ uint32_t active_block(const EngineArgs& a, uint32_t b, uint32_t ncand) {
  return b < ncand ? a.cand_list[b] : a.list[b - ncand];
}

}  // namespace mpm
