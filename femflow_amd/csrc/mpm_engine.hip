// The particle-resident MLS-MPM engine: binning, LDS-arena P2G, the fused
// G2P(n)+P2G(n+1) transfer, block-sparse grid update, and the rollout driver
// (eager or replayed as a captured hipGraph).  Reference step semantics are
// femflow/solvers/mpm/mls_mpm.py:40-79 with three_d/{p2g,grid_op,g2p}.py.
#include <hip/hip_ext.h>
#include <string.h>

#include <string>
#include <vector>

#include "mpm_engine.h"
#include "mpm_engine_kernels.h"
#include "mpm_sort.h"
#include "mpm_sort_dev.h"

namespace mpm {

// ------------------------------------------------------------------ G2P ----
// Weight of stencil index i (uniform loop counter) along one axis, the same
// arithmetic as bspline(): 0.5 (fx-1.5)^2, 0.75 - (fx-1)^2, 0.5 (fx-0.5)^2.
__device__ __forceinline__ float weight_at(float fx, int i) {
  const float t = fx - (i == 0 ? 1.5f : (i == 1 ? 1.0f : 0.5f));
  return i == 1 ? 0.75f - t * t : 0.5f * t * t;
}

// Velocity arena (G2P): node (a0, a1, a2) of the 8x8x8 arena at float4
// a0*81 + a1*9 + a2.  The plain 64/8/1 layout puts cells that differ in a0
// or in a1 by two at multiples of 256 bytes, on the same LDS banks; the odd
// strides spread a block's 64 cells over all of them (one-to-one on the
// arena, and it stays inside the accumulation arena it shares LDS with).
constexpr int VEL_S0 = 81, VEL_S1 = 9, VEL_NODES = 7 * VEL_S0 + 7 * VEL_S1 + 7 + 1;

// v = sum w g; C = 4 inv_dx sum w g (ijk - fx)^T  (three_d/g2p.py:29-46),
// accumulated as C[:, a] = 4 inv_dx (sum w g ijk_a - fx_a v).  The sums are
// taken axis by axis (w = wi wj wk): per row (i, j) the k-sums
// sum wk g and sum k wk g, per slab i their j-weighted sums, so each node
// costs 5 FMAs instead of 13.
template <bool FROM_LDS>
__device__ __forceinline__ void gather(const float4* __restrict__ vel_s, int abase,
                                       const float4* __restrict__ vel, const SceneDev& sc,
                                       const int base[3], const float w[3][3], const float fx[3],
                                       float v[3], float C[9]) {
  float S[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    v[a] = 0.f;
#pragma unroll
    for (int d = 0; d < 3; ++d) S[a][d] = 0.f;
  }
  size_t X[3], Y[3], Z[3];
  if (!FROM_LDS) {
    const size_t nb = (size_t)sc.nb;
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      const int ni = base[0] + o, nj = base[1] + o, nk = base[2] + o;
      X[o] = (size_t)(ni >> 2) * nb * nb * 64 + (size_t)((ni & 3) << 4);
      Y[o] = (size_t)(nj >> 2) * nb * 64 + (size_t)((nj & 3) << 2);
      Z[o] = (size_t)(nk >> 2) * 64 + (size_t)(nk & 3);
    }
  }
  const size_t gbase = (size_t)sc.block_base * 64;
  const float w2k2 = 2.f * w[2][2];
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const float wi = weight_at(fx[0], i);
    size_t Xi = 0;
    if (!FROM_LDS) Xi = i == 0 ? X[0] : (i == 1 ? X[1] : X[2]);
    const int ai = abase + i * VEL_S0;
    // accumulators start from their first term (an FMA onto +0 cannot be
    // folded into a product under IEEE signed zeros)
    float Vi[3], Tj[3], Tk[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float s[3], tk[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float4 g;
        if (FROM_LDS) g = vel_s[ai + j * VEL_S1 + k];
        else g = vel[gbase + Xi + Y[j] + Z[k]];
        const float gg[3] = {g.x, g.y, g.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (k == 0) s[d] = w[2][0] * gg[d];
          else s[d] += w[2][k] * gg[d];
          if (k == 1) tk[d] = w[2][1] * gg[d];
          if (k == 2) tk[d] += w2k2 * gg[d];
        }
      }
      const float wj = w[1][j];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (j == 0) {
          Vi[d] = wj * s[d];
          Tk[d] = wj * tk[d];
        } else {
          Vi[d] += wj * s[d];
          Tk[d] += wj * tk[d];
        }
        if (j == 1) Tj[d] = wj * s[d];
        if (j == 2) Tj[d] += (2.f * wj) * s[d];
      }
    }
    const float iwi = (float)i * wi;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] += wi * Vi[d];
      S[0][d] += iwi * Vi[d];
      S[1][d] += wi * Tj[d];
      S[2][d] += wi * Tk[d];
    }
  }
  const float c4 = 4.f * sc.inv_dx;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) C[3 * d + a] = c4 * (S[a][d] - fx[a] * v[d]);
}

// True when every singular value of F lies strictly inside [lo, hi]: the
// eigenvalues of A = F^T F are then inside [lo^2, hi^2], i.e. A - lo^2 I and
// hi^2 I - A are positive definite (leading principal minors > 0).
__device__ __forceinline__ bool singular_values_inside(const float F[9], float lo, float hi) {
  float A[6];  // symmetric: a00 a01 a02 a11 a12 a22
  A[0] = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  A[1] = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  A[2] = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  A[3] = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  A[4] = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  A[5] = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  bool ok = true;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const float sg = side == 0 ? 1.f : -1.f;
    const float sh = side == 0 ? lo * lo : hi * hi;
    const float b00 = sg * (A[0] - sh), b11 = sg * (A[3] - sh), b22 = sg * (A[5] - sh);
    const float b01 = sg * A[1], b02 = sg * A[2], b12 = sg * A[4];
    const float m2 = b00 * b11 - b01 * b01;
    const float m3 = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                     b02 * (b01 * b12 - b11 * b02);
    ok = ok && b00 > 0.f && m2 > 0.f && m3 > 0.f;
  }
  return ok;
}

// x += dt v, F <- (I + dt C) F, then plasticity (three_d/g2p.py:44-58).
// Snow runs the FP32 SVD only when the cheap test above says a singular value
// must be clamped.  Returns false when the clamp needs the FP64 SVD (F is
// then left at the trial value and the particle goes to the deferred list).
__device__ __forceinline__ bool advance(const SceneDev& sc, const MatParams& mp, float x[3],
                                        const float v[3], const float C[9], float F[9], float& Jp) {
#pragma unroll
  for (int d = 0; d < 3; ++d) x[d] += sc.dt * v[d];
  float M[9], Fn[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) M[i] = sc.dt * C[i] + ((i % 4 == 0) ? 1.f : 0.f);
  matmul3(M, F, Fn);
  bool ok = true;
  if (mp.kind == MPM_SNOW) {
    const float lo = 1.f - mp.theta_c, hi = 1.f + mp.theta_s;
    const float oldJ = det3(Fn);
    if (!singular_values_inside(Fn, lo, hi)) {
      float U[9], s[3], V[9];
      svd3<float>(Fn, U, s, V, 5);
      if (s[2] >= 1e-2f * s[0] && tfinite(s[0] + s[1] + s[2])) {
        if (s[0] < lo || s[0] > hi || s[1] < lo || s[1] > hi || s[2] < lo || s[2] > hi) {
#pragma unroll
          for (int c = 0; c < 3; ++c) s[c] = tmin(tmax(s[c], lo), hi);
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              Fn[3 * i + j] = U[3 * i] * s[0] * V[3 * j] + U[3 * i + 1] * s[1] * V[3 * j + 1] +
                              U[3 * i + 2] * s[2] * V[3 * j + 2];
        }
      } else {
        ok = false;
      }
    }
    if (ok) Jp = tmin(tmax(Jp * oldJ / (det3(Fn) + 1e-10f), 0.6f), 20.f);
  } else if (mp.kind == MPM_WATER) {
    const float c = tcbrt(det3(Fn));
#pragma unroll
    for (int i = 0; i < 9; ++i) Fn[i] = (i % 4 == 0) ? c : 0.f;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = Fn[i];
  return ok;
}

// Append particle p to the deferred (FP64) list; bit 0 = plasticity pending,
// bit 1 = no P2G wanted (the G2P-only pass).
__device__ __forceinline__ void defer(const EngineArgs& a, int64_t p, uint32_t flags) {
  const uint32_t i = atomicAdd(a.hard_count, 1u);
  a.hard_list[i] = ((uint32_t)p << 2) | flags;
}

// ------------------------------------------------------------------ P2G ----
// Per node (i,j,k): w_ijk (m v + A (ijk - fx) dx) and w_ijk m (three_d/p2g.py:67-80),
// written as w (q0 + Adx ijk) with q0 = m v - A fx dx, Adx = A dx.
struct P2GTerms {
  float q0[3];
  float Adx[9];
  float m;
};

// FP32 terms; false when the FP32 polar factor is not trustworthy (F near
// singular or the iteration did not settle): the caller defers the particle.
// P F^T = 2 mu (F - R) F^T + lam (J - 1) J (utils.py:120-135): the first
// term from the series of stretch_series while F is within 3% of a
// rotation (every snow particle after its clamp, and elastic ones under
// moderate strain), else from the FP32 polar iteration.
__device__ __forceinline__ bool p2g_terms(const SceneDev& sc, const MatParams& mp, float mass,
                                          float mu0, float lam0, float Jp, const float F[9],
                                          const float C[9], const float v[3], const float fx[3],
                                          P2GTerms& t) {
  float mu, lam;
  lame<float>(mp.kind, mp.hardening, mu0, lam0, Jp, mu, lam);
  float dev[9];  // 2 mu (F - R) F^T
#pragma unroll
  for (int i = 0; i < 9; ++i) dev[i] = 0.f;
  if (mu != 0.f) {
    const float m2 = 2.f * mu;
    float S[6];
    if (stretch_series<float>(F, S)) {
      dev[0] = m2 * S[0];
      dev[1] = dev[3] = m2 * S[1];
      dev[2] = dev[6] = m2 * S[2];
      dev[4] = m2 * S[3];
      dev[5] = dev[7] = m2 * S[4];
      dev[8] = m2 * S[5];
    } else {
      float R[9];
      if (!polar_newton<float>(F, R, 12, 1e-5f)) return false;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float d0 = m2 * (F[3 * i] - R[3 * i]), d1 = m2 * (F[3 * i + 1] - R[3 * i + 1]),
                    d2 = m2 * (F[3 * i + 2] - R[3 * i + 2]);
#pragma unroll
        for (int j = 0; j < 3; ++j) dev[3 * i + j] = d0 * F[3 * j] + d1 * F[3 * j + 1] + d2 * F[3 * j + 2];
      }
    }
  }
  const float Jm1 = det3_minus_one(F);
  const float scale = -(sc.dt * sc.vol) * (4.f * sc.inv_dx * sc.inv_dx);
  const float vt = lam * Jm1 * (1.f + Jm1);
  const bool diag_only = sc.quirks == MPM_QUIRKS_TEXTBOOK || mp.kind == MPM_WATER;
  float A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float pf = dev[3 * i + j] + ((diag_only && i != j) ? 0.f : vt);
      A[3 * i + j] = scale * pf + mass * C[3 * i + j];
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t.q0[d] = mass * v[d] - sc.dx * (A[3 * d] * fx[0] + A[3 * d + 1] * fx[1] + A[3 * d + 2] * fx[2]);
#pragma unroll
    for (int e = 0; e < 3; ++e) t.Adx[3 * d + e] = A[3 * d + e] * sc.dx;
  }
  t.m = mass;
  return true;
}

__device__ __forceinline__ void scatter_global(const EngineArgs& a, const SceneDev& sc,
                                               const int base[3], const float w[3][3],
                                               const P2GTerms& t) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float wt = w[0][i] * w[1][j] * w[2][k];
        const size_t n = node_addr(sc, base[0] + i, base[1] + j, base[2] + k);
        float* dst = reinterpret_cast<float*>(a.acc + n);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          atomicAdd(dst + d, wt * (t.q0[d] + t.Adx[3 * d] * (float)i + t.Adx[3 * d + 1] * (float)j +
                                   t.Adx[3 * d + 2] * (float)k));
        atomicAdd(dst + 3, wt * t.m);
      }
  // mark the (up to 8) blocks the stencil touched
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = base[0] + ((c >> 2) & 1) * 2, j = base[1] + ((c >> 1) & 1) * 2, k = base[2] + (c & 1) * 2;
    const uint32_t g = (uint32_t)(sc.block_base + ((i >> 2) * sc.nb + (j >> 2)) * sc.nb + (k >> 2));
    mark_active(a, g);
  }
}

// ----------------------------------------------------------- particle I/O ----
// One address per particle and field; components at immediate offsets.
template <int K>
__device__ __forceinline__ void loadk(const float* f, int64_t p, float o[K]) {
  const float* b = f + tix(p, K, 0);
#pragma unroll
  for (int d = 0; d < K; ++d) o[d] = b[d * PTILE];
}
template <int K>
__device__ __forceinline__ void storek(float* f, int64_t p, const float o[K]) {
  float* b = f + tix(p, K, 0);
#pragma unroll
  for (int d = 0; d < K; ++d) b[d * PTILE] = o[d];
}

// --------------------------------------------------------- transfer kernel ----
// One work item = one chunk of at most CHUNK binned particles of one 4x4x4-cell
// block.  MODE 0: P2G only (right after binning).  MODE 1: fused G2P of this
// substep + P2G of the next.  MODE 2: G2P only.
//
// P2G accumulates in an 8x8x8-node LDS arena in 32-bit fixed point: each
// thread first computes its particles' P2G terms; the chunk sums, per
// channel, every particle's bound on its largest single-node contribution
// (which bounds any node's total from this chunk), scales so that sum maps to
// 2^30, rounds each contribution to int32 and adds two channels at a time
// with 64-bit integer LDS atomics (float LDS atomics are an order of
// magnitude slower on this hardware; one 64-bit add costs less than two
// 32-bit ones).  Integer sums are associative, so the in-chunk accumulation
// is order independent.  The flush converts each node back to FP32 and
// stores the arena's 7x7x7 core to the chunk's staging slot, which the grid
// update sums; the eighth node layer, chunks beyond the staging capacity and
// particles that left the arena (drift since binning) take global atomics.
// Persistent: each XCD (workgroup id mod 8) walks a contiguous eighth of the
// chunk list, so neighbouring chunks share that XCD's L2.
struct ChunkScale {
  float s[4];     // float -> fixed point
  double inv[4];  // fixed point -> float
};

__device__ __forceinline__ float warp_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Largest |w_o (q0 + Adx o)| and |w_o m| over the 27 nodes, bounded by the
// largest weight product times the largest |q0 + Adx o| (o in {0,1,2}^3).
__device__ __forceinline__ void contribution_bound(const float w[3][3], const P2GTerms& t, float b[4]) {
  float wm = 1.f;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    wm *= fmaxf(fabsf(w[d][0]), fmaxf(fabsf(w[d][1]), fabsf(w[d][2])));
#pragma unroll
  for (int d = 0; d < 3; ++d)
    b[d] = wm * (fabsf(t.q0[d]) + 2.f * (fabsf(t.Adx[3 * d]) + fabsf(t.Adx[3 * d + 1]) + fabsf(t.Adx[3 * d + 2])));
  b[3] = wm * fabsf(t.m);
}

// Accumulation arena: two 64-bit words per node (x|y and z|m channel pairs,
// see pack_fixed2); node (a0, a1, a2) of the 8x8x8 arena lives at
// a0*101 + a1*9 + a2, which spreads the 64 cells of a block over the LDS
// banks (a plain 64/8/1 layout puts them on a quarter of the banks).  The
// map must be one-to-one over all 512 nodes, not only over the 6x6x6 nodes
// that freshly binned particles reach: particles that drifted across a block
// face since binning reach layers 0 and 7, so the row stride must be at least
// 7*9+8 = 71 (with 69, node (e, 7, 6) shared a word with (e+1, 0, 0)).  101
// is the smallest such stride that is 69 mod 32, i.e. keeps 69's bank pattern.
constexpr int ACC_S0 = 101, ACC_S1 = 9, ACC_NODES = 7 * ACC_S0 + 7 * ACC_S1 + 7 + 1;

constexpr bool acc_layout_one_to_one() {
  bool used[ACC_NODES] = {};
  for (int e0 = 0; e0 < 8; ++e0)
    for (int e1 = 0; e1 < 8; ++e1)
      for (int e2 = 0; e2 < 8; ++e2) {
        const int w = e0 * ACC_S0 + e1 * ACC_S1 + e2;
        if (w < 0 || w >= ACC_NODES || used[w]) return false;
        used[w] = true;
      }
  return true;
}
static_assert(acc_layout_one_to_one(), "accumulation arena layout must be one-to-one on 8x8x8 nodes");
static_assert(VEL_S1 >= 8 && VEL_S0 >= 7 * VEL_S1 + 8, "velocity arena layout must be one-to-one");
static_assert(4 * VEL_NODES <= 4 * ACC_NODES, "velocity arena must fit the LDS it shares");

// Two int32 fixed-point values in one 64-bit word whose 64-bit sum carries
// both sums exactly: word = hi * 2^32 + lo (lo sign-extended).  Valid while
// each sum stays inside int32, which the chunk scale guarantees.
__device__ __forceinline__ unsigned long long pack_fixed2(int hi, int lo) {
  const uint32_t h = (uint32_t)(hi + (lo >> 31));
  return ((unsigned long long)h << 32) | (unsigned long long)(uint32_t)lo;
}
__device__ __forceinline__ void unpack_fixed2(unsigned long long w, int& hi, int& lo) {
  lo = (int)(uint32_t)w;
  hi = (int)(uint32_t)(w >> 32) - (lo >> 31);
}
// The same with a low channel that is never negative (the mass: the arena
// path only takes particles whose weights and mass are >= 0, and its sum
// stays below 2^31): no borrow, so packing is free.
__device__ __forceinline__ unsigned long long pack_fixed2_nn(int hi, int lo_nonneg) {
  return ((unsigned long long)(uint32_t)hi << 32) | (unsigned long long)(uint32_t)lo_nonneg;
}
__device__ __forceinline__ void unpack_fixed2_nn(unsigned long long w, int& hi, int& lo) {
  lo = (int)(uint32_t)w;
  hi = (int)(uint32_t)(w >> 32);
}

// round(w * q) to int32 for finite, in-range products.  A truncating
// conversion is biased by up to one fixed-point unit per contribution toward
// zero, which adds up over the ~100 contributions a node receives: where a
// particle's bound is dominated by its affine term (strong velocity gradients
// at walls) the node momenta were off by ~1e-5 relative within three
// substeps.  Rounding to nearest leaves a zero-mean error of at most half a
// unit.  floor(x + 0.5) in one instruction: the compiler only forms
// v_cvt_flr_i32_f32 when NaNs are ruled out for the whole translation unit,
// which would also fold away the kernels' finiteness checks.
__device__ __forceinline__ int round_fixed(float w, float q) {
  const float y = __builtin_fmaf(w, q, 0.5f);
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(y));
  return r;
}

// SPREAD (3 or 9): how many different starting nodes the lanes that share a
// cell use in the scatter.  3 rotates the node slab (q mod 3); 9 also
// rotates the row inside the slab ((q mod 9) / 3), which spreads dense
// cells' atomics over three times as many LDS words at the cost of a
// runtime row index (a few VALU per slab).  Measured at 2^19 particles:
// equal at 16 particles per cell, 25% / 34% faster at 128 / 1,157, and 1.5%
// faster on the three-cube 2^20 scene (~21 per cell), so 9 is the default.
// MODE: 0 = P2G (the first substep of a run; after a rebin it also moves
// the particles into the new binned order), 1 = fused G2P + P2G, 2 = G2P
// (the last substep of a run), 3 = the fused pass that also moves the
// particles (the first substep after a rebin inside a run).
// DYN: workgroups take their chunks from a per-XCD ticket queue instead of
// a fixed interleave.  With tens of thousands of uneven chunks (many scenes
// of different densities) the queue balances the tail; with a few thousand
// the 8 contended counters cost more than they save, so only the fused pass
// of multi-scene engines uses it.
template <int MODE, int P, int SPREAD, bool DYN = false>
__global__ __launch_bounds__(ENG_THREADS) void transfer_kernel(EngineArgs a) {
  static_assert(SPREAD == 3 || SPREAD == 9, "scatter spread is 3 or 9");
  // The velocity arena (read by the G2P half) and the accumulation arena
  // (written by the P2G half after every gather of the chunk is done) share
  // one LDS region.
  constexpr int VEL_WORDS = MODE == 0 ? 4 : 4 * VEL_NODES;
  constexpr int ACC_WORDS = MODE == 2 ? 4 : 4 * ACC_NODES;
  __shared__ __align__(16) uint32_t arena_s[VEL_WORDS > ACC_WORDS ? VEL_WORDS : ACC_WORDS];
  float4* const vel_s = reinterpret_cast<float4*>(arena_s);
  unsigned long long(*const acc2)[MODE == 2 ? 1 : ACC_NODES] =
      reinterpret_cast<unsigned long long(*)[MODE == 2 ? 1 : ACC_NODES]>(arena_s);
  __shared__ float red_s[ENG_THREADS / 64][4];
  __shared__ ChunkScale cs;
  __shared__ uint32_t wg_oob, wg_fb;
  // P2G term table, four float4 per chunk entry, (fx0, fx1, fx2, m),
  // (q0[0..2], Adx0), (Adx1..4), (Adx5..8), read and written 16 bytes at a
  // time, and the arena cell of each entry (-1: not in the table).  Entry q
  // sits at slot q + q/16, which spreads the scatter phase's strided reads
  // over the banks.
  constexpr int TAB = MODE == 2 ? 1 : P * ENG_THREADS;
  constexpr int TABS = TAB + TAB / 16;
  __shared__ float4 pt4[4][TABS];
  __shared__ int pb_s[TABS];
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  if (t == 0) {
    wg_oob = 0;
    wg_fb = 0;
  }
  // before any wave counts into them (the first P2G has no barrier ahead of
  // its particle phase)
  __syncthreads();
  const uint32_t nch = *a.nchunks;
  const int xcd = blockIdx.x & 7;
  const int l = blockIdx.x >> 3;
  const int L = ((int)gridDim.x - xcd + 7) >> 3;
  const uint32_t lo = (uint32_t)(((uint64_t)nch * xcd) >> 3);
  const uint32_t hi = (uint32_t)(((uint64_t)nch * (xcd + 1)) >> 3);
  // a single-scene engine reads its scene once, not once per chunk (one
  // dependent load fewer ahead of each chunk's velocity arena)
  const bool one_scene = a.nscenes == 1;
  const SceneDev sc0 = a.scenes[0];
  __shared__ uint32_t next_s;
  uint32_t* const queue = a.counters + 8 + xcd;  // this XCD's ticket counter
  if constexpr (DYN) {
    if (t == 0) next_s = lo + atomicAdd(queue, 1u);
    __syncthreads();
  }
  for (uint32_t c = DYN ? next_s : lo + (uint32_t)l; c < hi;) {
    uint32_t nxt = 0u;
    if constexpr (DYN) {
      if (t == 0) nxt = lo + atomicAdd(queue, 1u);  // in flight while this chunk runs
    }
    const int4 ch = a.chunks[c];
    const SceneDev sc = one_scene ? sc0 : a.scenes[ch.w];
    const int local = ch.z - sc.block_base;
    const int nb = sc.nb;
    const int bi = local / (nb * nb), bj = (local / nb) % nb, bk = local % nb;
    const int O0 = bi * 4 - 1, O1 = bj * 4 - 1, O2 = bk * 4 - 1;
    if constexpr (MODE != 0) {
      for (int e = t; e < ARENA_NODES; e += ENG_THREADS) {
        const int gi = O0 + (e >> 6), gj = O1 + ((e >> 3) & 7), gk = O2 + (e & 7);
        const bool in = gi >= 0 && gj >= 0 && gk >= 0 && gi <= sc.res && gj <= sc.res && gk <= sc.res;
        vel_s[(e >> 6) * VEL_S0 + ((e >> 3) & 7) * VEL_S1 + (e & 7)] =
            in ? a.vel[node_addr(sc, gi, gj, gk)] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    __syncthreads();
    // Phase 1: particle q = r * ENG_THREADS + t of the chunk (coalesced).
    // G2P modes gather, advance and store the particle; P2G modes then
    // compute its terms into the LDS term table (or send it down the global
    // path when it left the arena or its terms are not finite).
    float bsum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int r = 0; r < P; ++r) {
      const int q = r * ENG_THREADS + t;
      if (q >= ch.y) continue;
      const int qs = q + (q >> 4);
      if constexpr (MODE != 2) pb_s[qs] = -1;
      const int64_t p = (int64_t)ch.x + q;
      float x[3], F[9], v[3], C[9], Jp, w[3][3], fx[3];
      float pmass, pmu0, plam0;
      int base[3];
      uint32_t attr;
      // MODE 3: the fused pass of a rebin inside a run; binned slot p
      // takes the particle of slot perm[p] of the previous order, whose G2P
      // inputs are read from there and whose constants are copied here
      const int64_t ps = MODE == 3 ? (int64_t)a.perm[p] : p;
      if constexpr (MODE == 3) {
        loadk<3>(a.src.x, ps, x);
        loadk<9>(a.src.F, ps, F);
        Jp = a.src.Jp[ps];
        attr = a.src.attr[ps];
        pmass = a.src.mass[ps];
        pmu0 = a.src.mu0[ps];
        plam0 = a.src.lam0[ps];
        a.s.attr[p] = attr;
        a.s.mass[p] = pmass;
        a.s.mu0[p] = pmu0;
        a.s.lam0[p] = plam0;
        a.s.order[p] = a.src.order[ps];
      } else if (MODE == 0 && a.perm) {
        // binned slot p takes the particle of slot perm[p] of the previous
        // order: every field the fused pass reads is copied here, before any
        // early exit.  v and C are read by the P2G below from registers and
        // recomputed by the next G2P, so the new slot gets them only where
        // they are read before that (FP64-deferred particles, per-substep
        // state output) or kept (particles that left the grid).
        const int64_t pm = a.perm[p];
        loadk<3>(a.src.x, pm, x);
        loadk<9>(a.src.F, pm, F);
        loadk<3>(a.src.v, pm, v);
        loadk<9>(a.src.C, pm, C);
        Jp = a.src.Jp[pm];
        attr = a.src.attr[pm];
        pmass = a.src.mass[pm];
        pmu0 = a.src.mu0[pm];
        plam0 = a.src.lam0[pm];
        const uint32_t ord = a.src.order[pm];
        storek<3>(a.s.x, p, x);
        storek<9>(a.s.F, p, F);
        if (a.store_vc) {
          storek<3>(a.s.v, p, v);
          storek<9>(a.s.C, p, C);
        }
        a.s.Jp[p] = Jp;
        a.s.attr[p] = attr;
        a.s.mass[p] = pmass;
        a.s.mu0[p] = pmu0;
        a.s.lam0[p] = plam0;
        a.s.order[p] = ord;
      } else {
        loadk<3>(a.s.x, p, x);
        loadk<9>(a.s.F, p, F);
        Jp = a.s.Jp[p];
        attr = a.s.attr[p];
        pmass = a.s.mass[p];
        pmu0 = a.s.mu0[p];
        plam0 = a.s.lam0[p];
        if constexpr (MODE == 0) {
          loadk<3>(a.s.v, p, v);
          loadk<9>(a.s.C, p, C);
        }
      }
      int mid = (int)(attr & 255u);
      mid = mid < a.nmat ? mid : a.nmat - 1;
      const MatParams mp = a.mats[mid];
      if constexpr (MODE != 0) {
        if (!locate3(x, sc, base, fx)) {
          if constexpr (MODE == 3) {  // the particle keeps its state in the new slot
            storek<3>(a.s.x, p, x);
            storek<9>(a.s.F, p, F);
            a.s.Jp[p] = Jp;
            loadk<3>(a.src.v, ps, v);
            loadk<9>(a.src.C, ps, C);
            storek<3>(a.s.v, p, v);
            storek<9>(a.s.C, p, C);
          }
          atomicAdd(&wg_oob, 1u);
          continue;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) bspline(fx[d], w[d]);
        const int a0 = base[0] - O0, a1 = base[1] - O1, a2 = base[2] - O2;
        if (a0 >= 0 && a0 <= 5 && a1 >= 0 && a1 <= 5 && a2 >= 0 && a2 <= 5)
          gather<true>(vel_s, a0 * VEL_S0 + a1 * VEL_S1 + a2, a.vel, sc, base, w, fx, v, C);
        else
          gather<false>(vel_s, 0, a.vel, sc, base, w, fx, v, C);
        const bool ok = advance(sc, mp, x, v, C, F, Jp);
        storek<3>(a.s.x, p, x);
        storek<9>(a.s.F, p, F);
        a.s.Jp[p] = Jp;
        // v and C are this G2P's output.  In the fused pass the P2G below
        // consumes them from registers and the next G2P recomputes both, so
        // they go to memory only where they are read: after the last G2P of
        // a segment, for per-substep state output, and for particles the
        // FP64 pass finishes.
        if (MODE == 2 || a.store_vc || !ok) {
          storek<3>(a.s.v, p, v);
          storek<9>(a.s.C, p, C);
        }
        if (!ok) {
          defer(a, p, MODE == 2 ? 3u : 1u);
          continue;
        }
      }
      // MODE 0 after a rebin: v and C not yet in the new slot
      const bool vc_unstored = MODE == 0 && a.perm && !a.store_vc;
      if constexpr (MODE != 2) {
        if (!locate3(x, sc, base, fx)) {
          if (vc_unstored) {
            storek<3>(a.s.v, p, v);
            storek<9>(a.s.C, p, C);
          }
          atomicAdd(&wg_oob, 1u);
          continue;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) bspline(fx[d], w[d]);
        P2GTerms tm;
        if (!p2g_terms(sc, mp, pmass, pmu0, plam0, Jp, F, C, v, fx, tm)) {
          if (((MODE == 1 || MODE == 3) && !a.store_vc) || vc_unstored) {  // the FP64 pass reads them
            storek<3>(a.s.v, p, v);
            storek<9>(a.s.C, p, C);
          }
          defer(a, p, 0u);
          continue;
        }
        const int a0 = base[0] - O0, a1 = base[1] - O1, a2 = base[2] - O2;
        // negative weights (the truncated base at a low wall, fx < 0.134) or
        // mass would make a mass contribution negative: float path
        bool inside = a0 >= 0 && a0 <= 5 && a1 >= 0 && a1 <= 5 && a2 >= 0 && a2 <= 5 &&
                      fminf(w[0][1], fminf(w[1][1], w[2][1])) >= 0.f && tm.m >= 0.f;
        if (inside) {
          float b[4];
          contribution_bound(w, tm, b);
          bool finite = true;
#pragma unroll
          for (int d = 0; d < 4; ++d) finite = finite && tfinite(b[d]);
          if (finite) {
#pragma unroll
            for (int d = 0; d < 4; ++d) bsum[d] += b[d];
            pt4[0][qs] = make_float4(fx[0], fx[1], fx[2], tm.m);
            pt4[1][qs] = make_float4(tm.q0[0], tm.q0[1], tm.q0[2], tm.Adx[0]);
            pt4[2][qs] = make_float4(tm.Adx[1], tm.Adx[2], tm.Adx[3], tm.Adx[4]);
            pt4[3][qs] = make_float4(tm.Adx[5], tm.Adx[6], tm.Adx[7], tm.Adx[8]);
            pb_s[qs] = a0 * ACC_S0 + a1 * ACC_S1 + a2;
          } else {  // NaN/Inf terms keep their value through the float path
            inside = false;
          }
        }
        if (!inside) scatter_global(a, sc, base, w, tm);
      }
    }
    if constexpr (MODE != 2) {
      // chunk-wide bound per channel -> fixed-point scales
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const float m = warp_sum(bsum[d]);
        if (lane == 0) red_s[wv][d] = m;
      }
      __syncthreads();
      for (int e = t; e < 2 * ACC_NODES; e += ENG_THREADS) acc2[0][e] = 0ull;
      if (t < 4) {
        float m = 0.f;
#pragma unroll
        for (int w2 = 0; w2 < ENG_THREADS / 64; ++w2) m += red_s[w2][t];
        m *= 1.0001f;  // round-up margin for the FP32 sums above
        const float s = m > 0.f ? 1073741824.f / m : 1.f;  // 2^30 / bound
        cs.s[t] = s;
        cs.inv[t] = 1.0 / (double)s;
      }
      __syncthreads();
      // Phase 2: scatter.  Particle q of the table for (lane, r), the lanes of
      // a wave S = ceil(count/64) entries apart, so that they rarely share a
      // cell; lanes that do share one start the 3x9-node stencil at different
      // node slabs (q mod 3), so their atomics hit different arena words.
      const float scl[4] = {cs.s[0], cs.s[1], cs.s[2], cs.s[3]};
      const int S = (ch.y + 63) >> 6;
      const int PP = (S + 3) >> 2;  // <= P
#pragma unroll 1
      for (int r = 0; r < PP; ++r) {
        const int jj = wv * PP + r;
        const int q = lane * S + jj;
        if (jj >= S || q >= ch.y) continue;
        const int qs = q + (q >> 4);
        const int pb = pb_s[qs];
        if (pb < 0) continue;
        const float4 t0 = pt4[0][qs], t1 = pt4[1][qs], t2 = pt4[2][qs], t3 = pt4[3][qs];
        const float fx0 = t0.x;
        float w1[3], w2[3];
        bspline(t0.y, w1);
        bspline(t0.z, w2);
        // Adx row-major: A0s[d] = Adx[3d], A1s[d] = Adx[3d+1], A2s[d] = Adx[3d+2]
        const float q0r[3] = {t1.x, t1.y, t1.z};
        const float A0r[3] = {t1.w, t2.z, t3.y};
        const float A1r[3] = {t2.x, t2.w, t3.z};
        const float A2r[3] = {t2.y, t3.x, t3.w};
        float q0s[3], A0s[3], A1s[3], A2s[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          q0s[d] = q0r[d] * scl[d];
          A0s[d] = A0r[d] * scl[d];
          A1s[d] = A1r[d] * scl[d];
          A2s[d] = A2r[d] * scl[d];
        }
        const float ms = t0.w * scl[3];
        const int i0 = SPREAD == 9 ? (q % 9) % 3 : q % 3;
        // row order of the slab (identity for SPREAD 3, where it folds to
        // the compile-time rows)
        float w1r[3], fjr[3];
        int jor[3];
        {
          const int j0 = SPREAD == 9 ? (q % 9) / 3 : 0;
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const int j = SPREAD == 9 ? (j0 + jj >= 3 ? j0 + jj - 3 : j0 + jj) : jj;
            w1r[jj] = SPREAD == 9 ? (j == 0 ? w1[0] : (j == 1 ? w1[1] : w1[2])) : w1[jj];
            fjr[jj] = (float)j;
            jor[jj] = j * ACC_S1;
          }
        }
        // the three slabs unrolled: the scheduler interleaves one slab's
        // conversions with the previous slab's LDS atomics (same VGPRs)
#pragma unroll
        for (int sl = 0; sl < 3; ++sl) {
          const int i = i0 + sl >= 3 ? i0 + sl - 3 : i0 + sl;
          const float wi = weight_at(fx0, i);
          const float fi = (float)i;
          const int ai = pb + i * ACC_S0;
          float qi[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) qi[d] = q0s[d] + A0s[d] * fi;
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const float wij = wi * w1r[jj];
            float qij[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) qij[d] = qi[d] + A1s[d] * fjr[jj];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const float wt = wij * w2[k];
              const int off = ai + jor[jj] + k;
              // rounded to nearest (fma with 0.5, floor conversion: the
              // same two instructions as a product and a truncation)
              // q + A2 k for k = 0, 1, 2 (no multiply by a literal 0 or 1,
              // which IEEE semantics keep from being folded)
              float qk[3];
#pragma unroll
              for (int d = 0; d < 3; ++d)
                qk[d] = k == 0 ? qij[d] : (k == 1 ? qij[d] + A2s[d] : qij[d] + 2.f * A2s[d]);
              const int c0 = round_fixed(wt, qk[0]);
              const int c1 = round_fixed(wt, qk[1]);
              const int c2 = round_fixed(wt, qk[2]);
              const int c3 = round_fixed(wt, ms);
              atomicAdd(&acc2[0][off], pack_fixed2(c1, c0));
              atomicAdd(&acc2[1][off], pack_fixed2_nn(c2, c3));
            }
          }
        }
      }
      __syncthreads();
      const double inv[4] = {cs.inv[0], cs.inv[1], cs.inv[2], cs.inv[3]};
      const bool staged = c < a.stage_cap;
      float4* stage = a.stage + (size_t)c * STAGE_REC;
#pragma unroll 1
      for (int e = t; e < ARENA_NODES; e += ENG_THREADS) {
        const int e0 = e >> 6, e1 = (e >> 3) & 7, e2 = e & 7;
        const int ph = e0 * ACC_S0 + e1 * ACC_S1 + e2;
        int q0, q1, q2, q3;
        unpack_fixed2(acc2[0][ph], q1, q0);
        unpack_fixed2_nn(acc2[1][ph], q2, q3);
        const float4 val = make_float4((float)((double)q0 * inv[0]), (float)((double)q1 * inv[1]),
                                       (float)((double)q2 * inv[2]), (float)((double)q3 * inv[3]));
        const bool to_stage = staged && e0 < STAGE && e1 < STAGE && e2 < STAGE;
        if (to_stage) stage[stage_slot(e0, e1, e2)] = val;
        // every arena node lies in one of the 27 blocks around this chunk's
        // block, all candidates of the grid update: no marking needed
        if (!to_stage && (q0 != 0 || q1 != 0 || q2 != 0 || q3 != 0)) {
          const int gi = O0 + e0, gj = O1 + e1, gk = O2 + e2;
          if (gi >= 0 && gj >= 0 && gk >= 0 && gi <= sc.res && gj <= sc.res && gk <= sc.res) {
            float* dst = reinterpret_cast<float*>(a.acc + node_addr(sc, gi, gj, gk));
            atomicAdd(dst + 0, val.x);
            atomicAdd(dst + 1, val.y);
            atomicAdd(dst + 2, val.z);
            atomicAdd(dst + 3, val.w);
          }
        }
      }
    }
    if constexpr (DYN) {
      if (t == 0) next_s = nxt;
    }
    __syncthreads();
    if constexpr (DYN) {
      c = next_s;
    } else {
      c += (uint32_t)L;
    }
  }
  __syncthreads();
  if (t == 0) {
    if (wg_oob) atomicAdd(a.oob, (unsigned long long)wg_oob);
    if (wg_fb) atomicAdd(a.fallback, (unsigned long long)wg_fb);
  }
}

// ------------------------------------------------------- deferred (FP64) ----
// Particles the FP32 transfer declined: snow plasticity whose SVD is too
// ill-conditioned for FP32 (s_min < 1e-2 s_max), and stresses whose FP32
// polar iteration did not settle (F near singular).  They are finished in
// FP64 (Jacobi SVD, Newton polar with SVD fallback) and scattered with
// global atomics, either by the grid update that follows the transfer (its
// first-started workgroup, before any workgroup reads the grid) or, where
// something reads the particles in between, by deferred_kernel.
__device__ __forceinline__ void finish_deferred(const EngineArgs& a, uint32_t e) {
  const int64_t p = e >> 2;
  const uint32_t attr = a.s.attr[p];
  int mid = (int)(attr & 255u);
  mid = mid < a.nmat ? mid : a.nmat - 1;
  const MatParams mp = a.mats[mid];
  const SceneDev sc = a.scenes[attr >> 8];
  double F[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = a.s.F[tix(p, 9, k)];
  double Jp = a.s.Jp[p];
  if ((e & 1u) && mp.kind == MPM_SNOW) {
    double U[9], sv[3], V[9];
    svd3<double>(F, U, sv, V, 10);
    const double oldJ = det3(F);
    const double lo = 1.0 - (double)mp.theta_c, hi = 1.0 + (double)mp.theta_s;
#pragma unroll
    for (int c = 0; c < 3; ++c) sv[c] = tmin(tmax(sv[c], lo), hi);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        F[3 * r + c] = U[3 * r] * sv[0] * V[3 * c] + U[3 * r + 1] * sv[1] * V[3 * c + 1] +
                       U[3 * r + 2] * sv[2] * V[3 * c + 2];
    Jp = tmin(tmax(Jp * oldJ / (det3(F) + 1e-10), 0.6), 20.0);
#pragma unroll
    for (int k = 0; k < 9; ++k) a.s.F[tix(p, 9, k)] = (float)F[k];
    a.s.Jp[p] = (float)Jp;
  }
  if (e & 2u) return;
  float xf[3];
  loadk<3>(a.s.x, p, xf);
  int base[3];
  float fxf[3];
  if (!locate3(xf, sc, base, fxf)) {
    atomicAdd(a.oob, 1ull);
    return;
  }
  double C[9], v[3], fx[3], w[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) C[k] = a.s.C[tix(p, 9, k)];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v[d] = a.s.v[tix(p, 3, d)];
    fx[d] = fxf[d];
    bspline(fx[d], w[d]);
  }
  const double m = a.s.mass[p];
  double mu, lam, A[9];
  lame<double>(mp.kind, (double)mp.hardening, (double)a.s.mu0[p], (double)a.s.lam0[p], Jp, mu, lam);
  affine3<double>(F, C, m, mu, lam, mp.kind, (double)sc.inv_dx, (double)sc.dt, (double)sc.vol,
                  sc.quirks == MPM_QUIRKS_TEXTBOOK, A);
  const double dx = sc.dx;
#pragma unroll
  for (int i0 = 0; i0 < 3; ++i0)
#pragma unroll
    for (int j0 = 0; j0 < 3; ++j0)
#pragma unroll
      for (int k0 = 0; k0 < 3; ++k0) {
        const double wt = w[0][i0] * w[1][j0] * w[2][k0];
        const double dp[3] = {(i0 - fx[0]) * dx, (j0 - fx[1]) * dx, (k0 - fx[2]) * dx};
        float* dst = reinterpret_cast<float*>(a.acc + node_addr(sc, base[0] + i0, base[1] + j0, base[2] + k0));
#pragma unroll
        for (int d = 0; d < 3; ++d)
          atomicAdd(dst + d, (float)(wt * (m * v[d] + A[3 * d] * dp[0] + A[3 * d + 1] * dp[1] + A[3 * d + 2] * dp[2])));
        atomicAdd(dst + 3, (float)(wt * m));
      }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = base[0] + ((c >> 2) & 1) * 2, j = base[1] + ((c >> 1) & 1) * 2, k = base[2] + (c & 1) * 2;
    mark_active(a, (uint32_t)(sc.block_base + ((i >> 2) * sc.nb + (j >> 2)) * sc.nb + (k >> 2)));
  }
}

__global__ __launch_bounds__(ENG_THREADS) void deferred_kernel(EngineArgs a) {
  const uint32_t count = *a.hard_count;
  uint32_t done_here = 0;
  for (uint32_t i = blockIdx.x * ENG_THREADS + threadIdx.x; i < count; i += gridDim.x * ENG_THREADS) {
    finish_deferred(a, a.hard_list[i]);
    ++done_here;
  }
  if (done_here) atomicAdd(a.fallback, (unsigned long long)done_here);
}

// Start of a segment: clear both marked-block list counters, both deferred
// counts, the grid update's start ticket and the chunk queues, advance the
// per-substep sums index by the segment length, and, when the segment
// rebins, start a new candidate list (count 0, next generation).
__device__ __forceinline__ void segment_begin(uint32_t* counters, int len, int with_sums, int bin) {
  counters[0] = 0u;
  counters[1] = 0u;
  counters[3] = 0u;
  counters[6] = 0u;
  counters[7] = 0u;
  for (int x = 0; x < 8; ++x) counters[8 + x] = 0u;
  if (with_sums) counters[2] += (uint32_t)len;
  if (bin) {
    counters[4] = 0u;
    counters[5] += 1u;
  }
}

__global__ void segment_begin_kernel(uint32_t* counters, int len, int with_sums, int bin) {
  if (threadIdx.x == 0 && blockIdx.x == 0) segment_begin(counters, len, with_sums, bin);
}

// ------------------------------------------------------------ grid update ----
// three_d/grid_op.py:25-47 on the active blocks only; also zeroes the
// accumulation grid and the block flags for the next P2G, and (optionally)
// adds FP64 conservation sums for this substep.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Deferred particles left by the transfer before this grid update: the
// workgroup that starts first (a ticket, so it is resident) finishes them
// and then sets bit 31 of the count; the others wait for that bit before
// they read the grid.  A count of 0 (the usual case) costs one load.  The
// wait is bounded; running out sets the error counter instead of hanging.
constexpr uint32_t DEFER_DONE = 0x80000000u;
constexpr uint32_t DEFER_SPIN_CAP = 1u << 24;

__device__ __forceinline__ void grid_deferred(const EngineArgs& a) {
  __shared__ uint32_t ticket_s;
  // written by the transfer before this launch; only its bit 31 changes
  // during it, so every thread sees 0 or every thread sees nonzero
  const uint32_t count = *a.hard_count;
  if (count == 0u) return;
  if (threadIdx.x == 0) ticket_s = atomicAdd(a.defer_ticket, 1u);
  __syncthreads();
  const uint32_t ticket = ticket_s;
  if (ticket == 0u && !(count & DEFER_DONE)) {
    const uint32_t c = count & ~DEFER_DONE;
    for (uint32_t i = threadIdx.x; i < c; i += ENG_THREADS) finish_deferred(a, a.hard_list[i]);
    __syncthreads();
    if (threadIdx.x == 0) {
      atomicAdd(a.fallback, (unsigned long long)c);
      __threadfence();
      __hip_atomic_store(a.hard_count, c | DEFER_DONE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    uint32_t spins = 0;
    while (!(__hip_atomic_load(a.hard_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & DEFER_DONE)) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins >= DEFER_SPIN_CAP) {
        atomicAdd(a.oob, 1ull << 40);  // reported as an error by the host
        break;
      }
    }
    __threadfence();
  }
  // the last workgroup to take a ticket resets it for the next launch
  if (threadIdx.x == 0 && ticket == gridDim.x - 1u) __hip_atomic_store(a.defer_ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
}

__global__ __launch_bounds__(ENG_THREADS) __attribute__((amdgpu_waves_per_eu(3))) void grid_update_kernel(
    EngineArgs a, int with_sums, int len, int k, int deferred) {
  __shared__ double red[ENG_THREADS / 64][8];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *a.other_count = 0u;  // the list the next substep appends to
    *a.other_hard = 0u;   // the deferred count the next substep appends to
    for (int x = 0; x < 8; ++x) a.counters[8 + x] = 0u;  // the next transfer's chunk queues
  }
  const int t = threadIdx.x;
  const int wv = t >> 6, inner = t & 63;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // One listed block per wave, one node per lane.  The kernel is a chain of
  // dependent loads (list entry, neighbour chunk slots, staged arenas), so
  // the first candidate entry is read together with the counts (every slot
  // below list_cap is addressable; entries past the count are ignored), and
  // a single-scene engine skips the block -> scene lookup.  The candidate
  // list and its table are fixed since the rebin, so they are read ahead of
  // the deferred pass; the marked list, the accumulator and the staged
  // arenas only after it.
  const bool one_scene = a.nscenes == 1;
  const uint32_t stride = gridDim.x * 4u;
  uint32_t b = blockIdx.x * 4u + (uint32_t)wv;
  uint32_t g_next = b < a.list_cap ? a.cand_list[b] : 0u;
  uint2 src_next = b < a.list_cap && inner < 27 ? a.cand_src[(size_t)b * 32 + inner] : make_uint2(0u, 0u);
  const uint32_t ncand = *a.cand_count;
  if (deferred) grid_deferred(a);
  const uint32_t count = ncand + *a.list_count;
  for (; b < count; b += stride) {
    const bool cand = b < ncand;
    const uint32_t g = cand ? g_next : a.list[b - ncand];
    const uint2 src = src_next;
    g_next = b + stride < a.list_cap ? a.cand_list[b + stride] : 0u;
    src_next = b + stride < a.list_cap && inner < 27 ? a.cand_src[(size_t)(b + stride) * 32 + inner]
                                                     : make_uint2(0u, 0u);
    const int sid = one_scene ? 0 : a.block_scene[g];
    const SceneDev sc = a.scenes[sid];
    const int local = (int)g - sc.block_base;
    const int nb = sc.nb;
    const int bi = local / (nb * nb), bj = (local / nb) % nb, bk = local % nb;
    const int o0 = inner >> 4, o1 = (inner >> 2) & 3, o2 = inner & 3;
    const int i = bi * 4 + o0, j = bj * 4 + o1, k = bk * 4 + o2;
    const size_t addr = (size_t)g * 64 + inner;
    // lanes 0..26: first chunk and chunk count of the 27 blocks around this
    // one (candidates: from the rebin's table; marked blocks: looked up)
    uint32_t my_c0 = src.x, my_cn = src.y;
    if (!cand && inner < 27) {
      const uint2 nc = neighbour_chunks(a, sc, bi - 1 + inner / 9, bj - 1 + (inner / 3) % 3, bk - 1 + inner % 3);
      my_c0 = nc.x;
      my_cn = nc.y;
    }
    float4 q = a.acc[addr];
    // the accumulator is nonzero only where a global-atomic path wrote
    const bool acc_used = q.x != 0.f || q.y != 0.f || q.z != 0.f || q.w != 0.f;
    // Staged chunk arenas that cover this node: per axis, node offset o < 3
    // lies in the arenas of blocks b-1 and b, o = 3 in those of b and b+1
    // (an arena spans nodes 4b-1 .. 4b+6).  The first chunk of each of the 8
    // sources is loaded together; further chunks (dense blocks) follow.
    uint32_t c0[8], cn[8];
    int ai[8];
#pragma unroll
    for (int sn = 0; sn < 8; ++sn) {
      const int s0 = (sn >> 2) - (o0 == 3 ? 0 : 1);
      const int s1 = ((sn >> 1) & 1) - (o1 == 3 ? 0 : 1);
      const int s2 = (sn & 1) - (o2 == 3 ? 0 : 1);
      const int ti = (s0 + 1) * 9 + (s1 + 1) * 3 + (s2 + 1);
      c0[sn] = (uint32_t)__shfl((int)my_c0, ti);
      cn[sn] = (uint32_t)__shfl((int)my_cn, ti);
      // arena index of this node in the source's arena; index 7 (the
      // unstaged eighth layer) came through the global accumulator
      const int x0 = o0 + 1 - 4 * s0, x1 = o1 + 1 - 4 * s1, x2 = o2 + 1 - 4 * s2;
      const bool unstaged = x0 == STAGE || x1 == STAGE || x2 == STAGE;
      ai[sn] = unstaged ? 0 : stage_slot(x0, x1, x2);
      if (unstaged) cn[sn] = 0u;
    }
    // the first two chunks of every source are loaded in one round trip
    // and summed source-major, chunk-minor; further chunks (a block of ~20
    // particles per cell has three) one per round, all sources together, so
    // a block's dependent rounds follow its densest source rather than the
    // sum over sources.  Two per source rather than three keeps the kernel
    // at 168 registers, three waves per SIMD: every listed block of a
    // 2^20-particle scene then has a resident wave from the start.
    float4 sv[2][8];
#pragma unroll
    for (int sn = 0; sn < 8; ++sn) {
      const float4* src = a.stage + (size_t)c0[sn] * STAGE_REC + ai[sn];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        sv[c][sn] = cn[sn] > (uint32_t)c ? src[(size_t)c * STAGE_REC] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int sn = 0; sn < 8; ++sn) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        q.x += sv[c][sn].x;
        q.y += sv[c][sn].y;
        q.z += sv[c][sn].z;
        q.w += sv[c][sn].w;
      }
    }
    uint32_t cmax = 0;
#pragma unroll
    for (int sn = 0; sn < 8; ++sn) cmax = cn[sn] > cmax ? cn[sn] : cmax;
    for (uint32_t c = 2; c < cmax; ++c) {
      float4 v2[8];
#pragma unroll
      for (int sn = 0; sn < 8; ++sn)
        v2[sn] = c < cn[sn] ? a.stage[(size_t)(c0[sn] + c) * STAGE_REC + ai[sn]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int sn = 0; sn < 8; ++sn) {
        q.x += v2[sn].x;
        q.y += v2[sn].y;
        q.z += v2[sn].z;
        q.w += v2[sn].w;
      }
    }
    if (a.pm) a.pm[addr] = q;
    const float m = q.w;
    float v[3] = {q.x, q.y, q.z};
    if (with_sums) {
      s[0] += m;
      s[1] += q.x;
      s[2] += q.y;
      s[3] += q.z;
      s[7] += m > 0.f ? 1.0 : 0.0;  // nodes with mass
    }
    if (m > 0.f) {
      const float im = 1.f / m;
      v[0] *= im;
      v[1] *= im;
      v[2] *= im;
      v[1] += sc.dt * sc.gravity;
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = tmin(tmax(v[d], -sc.v_allowed), sc.v_allowed);
    }
    const int I[3] = {i, j, k};
    if (sc.boundary == MPM_BC_REFERENCE) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (I[d] < 1 || I[d] >= sc.res - 1) v[d] = 0.f;
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (I[d] < sc.bound || I[d] > sc.res - sc.bound) v[0] = v[1] = v[2] = 0.f;
    }
    if (i > sc.res || j > sc.res || k > sc.res) v[0] = v[1] = v[2] = 0.f;
    for (int pl = 0; pl < a.nplanes; ++pl)
      if (a.planes[pl].scene == sid && plane_blocks(a.planes[pl], i, j, k)) v[0] = v[1] = v[2] = 0.f;
    a.vel[addr] = make_float4(v[0], v[1], v[2], m);
    if (acc_used) a.acc[addr] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a.grid_rec && i <= sc.res && j <= sc.res && k <= sc.res) {
      const size_t r1 = (size_t)sc.res + 1, nn = r1 * r1 * r1;
      const size_t di = ((size_t)i * r1 + (size_t)j) * r1 + (size_t)k;
      float* gp = a.grid_rec;
      gp[3 * di] = q.x;
      gp[3 * di + 1] = q.y;
      gp[3 * di + 2] = q.z;
      gp[3 * nn + di] = m;
      float* gv = gp + 4 * nn;
      gv[3 * di] = v[0];
      gv[3 * di + 1] = v[1];
      gv[3 * di + 2] = v[2];
    }
    if (with_sums) {
      s[4] += (double)m * v[0];
      s[5] += (double)m * v[1];
      s[6] += (double)m * v[2];
    }
    if (inner == 0) a.flags[g] = 0u;
  }
  if (with_sums) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double r = wave_sum(s[c]);
      if ((t & 63) == 0) red[wv][c] = r;
    }
  }
  __syncthreads();
  if (t == 0 && with_sums) {
    const int step = (int)a.counters[2] - len + k;
    if (step >= 0 && step < a.max_steps) {
      for (int c = 0; c < 8; ++c) {
        double r = 0;
        for (int w = 0; w < ENG_THREADS / 64; ++w) r += red[w][c];
        atomicAdd(a.sums + (size_t)step * 8 + c, r);
      }
    }
  }
}

// ------------------------------------------------------- binning helpers ----
__device__ __forceinline__ int cell_of(float x, float inv_dx, int res) {
  const float t = __fsub_rn(__fmul_rn(x, inv_dx), 0.5f);
  int b;
  if (!(t > -1.0f)) b = 0;
  else if (!(t < (float)res)) b = res;
  else b = (int)t;
  return b < 0 ? 0 : (b > res ? res : b);
}

// Cell keys of the new binning, one sort tile (RS_TILE particles) per
// workgroup, with the tile histogram of the sort's first digit, so the sort
// starts at that pass's scan.  Also starts the segment that rebins (the
// counters' segment_begin) and clears the block occupancy that
// block_bounds_kernel sets: two launches fewer ahead of the sort.
__global__ __launch_bounds__(RS_THREADS) void engine_keys_kernel(EngineArgs a, int64_t n,
                                                                 uint32_t* __restrict__ keys,
                                                                 uint32_t* __restrict__ bocc, int64_t nblocks,
                                                                 int len, int with_sums,
                                                                 uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[RS_BINS];
  const int64_t gid = (int64_t)blockIdx.x * RS_THREADS + threadIdx.x;
  if (gid == 0) segment_begin(a.counters, len, with_sums, 1);
  for (int64_t j = gid; j < nblocks; j += (int64_t)gridDim.x * RS_THREADS) bocc[j] = 0u;
  const bool one_scene = a.nscenes == 1;
  const SceneDev sc0 = a.scenes[0];
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  // every position of the thread is loaded before the first key is stored
  // (the key stores could alias the positions as far as the compiler knows,
  // which would otherwise put one memory round trip per key in series)
  float px[RS_ITEMS][3];
  uint32_t sid[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + (int64_t)r * RS_THREADS + threadIdx.x;
    const int64_t j = i < n ? i : 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) px[r][d] = a.s.x[tix(j, 3, d)];
    sid[r] = one_scene ? 0u : a.s.attr[j] >> 8;
  }
  uint32_t k[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + (int64_t)r * RS_THREADS + threadIdx.x;
    const SceneDev sc = one_scene ? sc0 : a.scenes[sid[r]];
    const int b0 = cell_of(px[r][0], sc.inv_dx, sc.res);
    const int b1 = cell_of(px[r][1], sc.inv_dx, sc.res);
    const int b2 = cell_of(px[r][2], sc.inv_dx, sc.res);
    const uint32_t blk = (uint32_t)(sc.block_base + ((b0 >> 2) * sc.nb + (b1 >> 2)) * sc.nb + (b2 >> 2));
    k[r] = blk * 64u + (uint32_t)(((b0 & 3) << 4) | ((b1 & 3) << 2) | (b2 & 3));
    if (i < n) keys[i] = k[r];
  }
  rs_tile_histogram(k, n, 0, h, hist);
}

__global__ void block_bounds_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                    uint32_t* __restrict__ bstart, uint32_t* __restrict__ bend,
                                    uint32_t* __restrict__ bocc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t blk = keys[i] >> 6;
  if (i == 0 || (keys[i - 1] >> 6) != blk) {
    bstart[blk] = (uint32_t)i;
    bocc[blk] = 1u;
  }
  if (i == n - 1 || (keys[i + 1] >> 6) != blk) bend[blk] = (uint32_t)(i + 1);
}

// Candidate blocks of the grid update for the new binning (the segment's
// start has advanced the binning generation and zeroed the count): every
// chunk tags the 27 blocks around its own with the generation, and the first
// tag of a block appends it, so each block is listed once.  A block is a
// candidate while its tag equals the current generation, so the previous
// list needs no clearing.
__global__ void cand_build_kernel(EngineArgs a, uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                                  uint32_t* __restrict__ flag) {
  const uint32_t gen = *a.cand_gen;
  const uint32_t n = *a.nchunks * 27u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 ch = a.chunks[i / 27u];
    const int nbh = (int)(i % 27u);
    const SceneDev sc = a.scenes[ch.w];
    const int local = ch.z - sc.block_base;
    const int nb = sc.nb;
    const int bi = local / (nb * nb) - 1 + nbh / 9, bj = (local / nb) % nb - 1 + (nbh / 3) % 3,
              bk = local % nb - 1 + nbh % 3;
    if (bi < 0 || bj < 0 || bk < 0 || bi >= nb || bj >= nb || bk >= nb) continue;
    const uint32_t g = (uint32_t)(sc.block_base + (bi * nb + bj) * nb + bk);
    if (flag[g] != gen && atomicExch(flag + g, gen) != gen) list[atomicAdd(count, 1u)] = g;
  }
}

// Per candidate entry (one wave each, lanes 0..26): the staged chunks of
// its 27 neighbouring blocks.  They are fixed until the next rebin, so the
// grid update reads them together with the entry instead of looking them
// up through the block's neighbours.
__global__ void cand_src_kernel(EngineArgs a, uint2* __restrict__ out) {
  const uint32_t ncand = *a.cand_count;
  const int lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); b < ncand; b += waves) {
    if (lane >= 27) continue;
    const uint32_t g = a.cand_list[b];
    const int sid = a.nscenes == 1 ? 0 : a.block_scene[g];
    const SceneDev sc = a.scenes[sid];
    const int local = (int)g - sc.block_base;
    const int nb = sc.nb;
    const int bi = local / (nb * nb), bj = (local / nb) % nb, bk = local % nb;
    out[(size_t)b * 32 + lane] =
        neighbour_chunks(a, sc, bi - 1 + lane / 9, bj - 1 + (lane / 3) % 3, bk - 1 + lane % 3);
  }
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_f64_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ----------------------------------------------------------- upload/download ----
__global__ void upload_kernel(EngineState s, int64_t n, const float* __restrict__ x,
                              const float* __restrict__ v, const float* __restrict__ F,
                              const float* __restrict__ C, const float* __restrict__ Jp,
                              const float* __restrict__ mass, const float* __restrict__ mu0,
                              const float* __restrict__ lam0, int mat_scalar,
                              const int32_t* __restrict__ material, const int32_t* __restrict__ scene,
                              int nscenes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s.x[tix(i, 3, d)] = x[3 * i + d];
    s.v[tix(i, 3, d)] = v[3 * i + d];
  }
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    s.F[tix(i, 9, d)] = F[9 * i + d];
    s.C[tix(i, 9, d)] = C[9 * i + d];
  }
  s.Jp[i] = Jp[i];
  const int64_t mi = mat_scalar ? 0 : i;
  s.mass[i] = mass[mi];
  s.mu0[i] = mu0[mi];
  s.lam0[i] = lam0[mi];
  const uint32_t mat = material ? (uint32_t)(material[i] & 255) : 0u;
  int sc_i = scene ? scene[i] : 0;
  sc_i = sc_i < 0 ? 0 : (sc_i >= nscenes ? nscenes - 1 : sc_i);
  const uint32_t sc = (uint32_t)sc_i;
  s.attr[i] = mat | (sc << 8);
  s.order[i] = (uint32_t)i;
}

__global__ void download_kernel(EngineState s, int64_t n, float* __restrict__ x, float* __restrict__ v,
                                float* __restrict__ F, float* __restrict__ C, float* __restrict__ Jp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = s.order[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x[3 * o + d] = s.x[tix(i, 3, d)];
    v[3 * o + d] = s.v[tix(i, 3, d)];
  }
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    F[9 * o + d] = s.F[tix(i, 9, d)];
    C[9 * o + d] = s.C[tix(i, 9, d)];
  }
  Jp[o] = s.Jp[i];
}

// Positions in the caller's order (one frame of the rollout output).
__global__ void positions_kernel(EngineState s, int64_t n, float* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = s.order[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) x[3 * o + d] = s.x[tix(i, 3, d)];
}

__global__ void particle_sums_kernel(EngineState s, int64_t n, double* out) {
  __shared__ double red[ENG_THREADS / 64][8];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double m = s.mass[i];
    acc[0] += m;
    for (int d = 0; d < 3; ++d) {
      acc[1 + d] += m * (double)s.v[tix(i, 3, d)];
      acc[4 + d] += m * (double)s.x[tix(i, 3, d)];
    }
  }
  const int t = threadIdx.x;
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    const double r = wave_sum(acc[c]);
    if ((t & 63) == 0) red[t >> 6][c] = r;
  }
  __syncthreads();
  if (t < 7) {
    double r = 0;
    for (int w = 0; w < ENG_THREADS / 64; ++w) r += red[w][t];
    atomicAdd(out + t, r);
  }
}

// ================================================================ host ====
static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                  \
      return MPM_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

Engine::~Engine() { release(); }

void Engine::release() {
  for (auto& kv : graphs) {
    if (kv.second) hipGraphExecDestroy(kv.second);
  }
  graphs.clear();
  for (void* p : allocs) hipFree(p);
  allocs.clear();
  if (ev_in) hipEventDestroy(ev_in);
  if (ev_out) hipEventDestroy(ev_out);
  for (auto& e : tevents) hipEventDestroy(e);
  tevents.clear();
  if (own_stream) hipStreamDestroy(own_stream);
  ev_in = ev_out = nullptr;
  own_stream = nullptr;
}

int Engine::alloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  allocs.push_back(*p);
  return MPM_OK;
}

int Engine::init(int64_t n_max_, int nscenes, const mpm_scene* sc, int nmat_, const mpm_material* mats) {
  if (n_max_ < 0 || n_max_ >= (int64_t(1) << 30) || nscenes <= 0 || nscenes > MAX_SCENES || !sc || nmat_ <= 0 ||
      nmat_ > 256 || !mats) {
    set_error("mpm_ctx_create: bad argument (0 <= n_max < 2^30, 1 <= nscenes <= 4096, 1 <= nmat <= 256)");
    return MPM_ERR_ARG;
  }
  n_max = n_max_;
  nmat = nmat_;
  HIPCHK(hipGetDevice(&device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  num_cus = prop.multiProcessorCount;
  HIPCHK(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
  // scene table
  std::vector<SceneDev> hs(nscenes);
  std::vector<int> bscene;
  int64_t blocks = 0;
  for (int s = 0; s < nscenes; ++s) {
    const mpm_scene& q = sc[s];
    if (q.res < 4 || q.res > 1024 || !(q.dt > 0)) {
      set_error("mpm_ctx_create: scene res must be in [4, 1024] and dt > 0");
      return MPM_ERR_ARG;
    }
    SceneDev d{};
    d.res = q.res;
    d.nb = (q.res + 1 + 3) / 4;
    d.block_base = (int)blocks;
    d.boundary = q.boundary;
    d.bound = q.bound;
    d.quirks = q.quirks;
    d.dx = (float)(1.0 / q.res);
    d.inv_dx = (float)q.res;
    d.dt = (float)q.dt;
    d.gravity = (float)q.gravity;
    d.vol = (float)q.volume;
    d.v_allowed = (float)((1.0 / q.res) * 0.9 / q.dt);
    hs[s] = d;
    const int64_t nbs = (int64_t)d.nb * d.nb * d.nb;
    bscene.insert(bscene.end(), (size_t)nbs, s);
    blocks += nbs;
  }
  if (blocks * 64 >= (int64_t)1 << 32) {
    set_error("mpm_ctx_create: total grid blocks exceed the 32-bit key space");
    return MPM_ERR_ARG;
  }
  total_blocks = blocks;
  scenes_host = hs;
  dyn_sched = nscenes >= 8;
  int rc;
  const int64_t n = ((n_max > 0 ? n_max : 1) + PTILE - 1) / PTILE * PTILE;
  if ((rc = alloc((void**)&d_scenes, sizeof(SceneDev) * nscenes))) return rc;
  if ((rc = alloc((void**)&d_mats, sizeof(MatParams) * nmat))) return rc;
  if ((rc = alloc((void**)&d_block_scene, sizeof(int) * blocks))) return rc;
  HIPCHK(hipMemcpy(d_scenes, hs.data(), sizeof(SceneDev) * nscenes, hipMemcpyHostToDevice));
  std::vector<MatParams> hm(nmat);
  for (int i = 0; i < nmat; ++i) {
    hm[i].kind = mats[i].kind;
    hm[i].hardening = mats[i].hardening;
    hm[i].theta_c = mats[i].theta_c;
    hm[i].theta_s = mats[i].theta_s;
  }
  HIPCHK(hipMemcpy(d_mats, hm.data(), sizeof(MatParams) * nmat, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_block_scene, bscene.data(), sizeof(int) * blocks, hipMemcpyHostToDevice));
  // particle state (two copies for reordering)
  for (int b = 0; b < 2; ++b) {
    EngineState& s = st[b];
    if ((rc = alloc((void**)&s.x, 12 * n)) || (rc = alloc((void**)&s.v, 12 * n)) ||
        (rc = alloc((void**)&s.F, 36 * n)) || (rc = alloc((void**)&s.C, 36 * n)) ||
        (rc = alloc((void**)&s.Jp, 4 * n)) || (rc = alloc((void**)&s.mass, 4 * n)) ||
        (rc = alloc((void**)&s.mu0, 4 * n)) || (rc = alloc((void**)&s.lam0, 4 * n)) ||
        (rc = alloc((void**)&s.attr, 4 * n)) || (rc = alloc((void**)&s.order, 4 * n)))
      return rc;
  }
  // grids
  if ((rc = alloc((void**)&d_acc, sizeof(float4) * 64 * blocks))) return rc;
  if ((rc = alloc((void**)&d_vel, sizeof(float4) * 64 * blocks))) return rc;
  if ((rc = alloc((void**)&d_flags, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_list, sizeof(uint32_t) * 2 * blocks))) return rc;
  if ((rc = alloc((void**)&d_bstart, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_bend, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_counters, sizeof(uint32_t) * 16))) return rc;
  if ((rc = alloc((void**)&d_oob, sizeof(unsigned long long) * 2))) return rc;
  d_fallback = d_oob + 1;
  // binning workspace
  const int64_t max_chunks = n / 256 + blocks + 1;
  if ((rc = alloc((void**)&d_chunks, sizeof(int4) * (max_chunks < n + 1 ? max_chunks : n + 1)))) return rc;
  if ((rc = alloc((void**)&d_nchunks, sizeof(uint32_t) * 4))) return rc;
  // staged P2G arenas: room for every chunk at the smallest chunk size unless
  // nearly every particle sits in its own block
  const int64_t cap_chunks = (max_chunks < n + 1 ? max_chunks : n + 1);
  const int64_t want = 2 * ((n + 255) / 256) + 8192;
  stage_cap = (uint32_t)(want < cap_chunks ? want : cap_chunks);
  if ((rc = alloc((void**)&d_stage, sizeof(float4) * STAGE_REC * (size_t)stage_cap))) return rc;
  if ((rc = alloc((void**)&d_bfirst, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_bcount, sizeof(uint32_t) * blocks))) return rc;
  HIPCHK(hipMemset(d_bcount, 0, sizeof(uint32_t) * blocks));
  if ((rc = alloc((void**)&d_keys, sizeof(uint32_t) * n))) return rc;
  if ((rc = alloc((void**)&d_vals, sizeof(uint32_t) * n))) return rc;
  if ((rc = alloc((void**)&d_bocc, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_part, sizeof(uint32_t) * chunk_list_part_words(blocks)))) return rc;
  if ((rc = alloc((void**)&d_hard, sizeof(uint32_t) * n))) return rc;
  if ((rc = alloc((void**)&d_cand_list, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_cand_flag, sizeof(uint32_t) * blocks))) return rc;
  if ((rc = alloc((void**)&d_cand_src, sizeof(uint2) * 32 * blocks))) return rc;
  HIPCHK(hipMemset(d_cand_flag, 0, sizeof(uint32_t) * blocks));
  ws_bytes = radix_workspace_bytes(n);
  if ((rc = alloc(&d_ws, ws_bytes))) return rc;
  HIPCHK(hipMemset(d_acc, 0, sizeof(float4) * 64 * blocks));
  HIPCHK(hipMemset(d_vel, 0, sizeof(float4) * 64 * blocks));
  HIPCHK(hipMemset(d_flags, 0, sizeof(uint32_t) * blocks));
  HIPCHK(hipMemset(d_counters, 0, sizeof(uint32_t) * 16));
  HIPCHK(hipMemset(d_oob, 0, sizeof(unsigned long long) * 2));
  HIPCHK(hipMemset(d_nchunks, 0, sizeof(uint32_t) * 4));
  // the memsets above run on the null stream; the engine's stream does not
  // synchronise with it, so wait for them here
  HIPCHK(hipDeviceSynchronize());
  return MPM_OK;
}

EngineArgs Engine::args(double* sums, int max_steps, int sel) const {
  EngineArgs a{};
  a.s = st[cur];
  a.src = st[cur ^ 1];
  a.perm = nullptr;
  a.scenes = d_scenes;
  a.mats = d_mats;
  a.nmat = nmat;
  a.acc = d_acc;
  a.vel = d_vel;
  a.flags = d_flags;
  a.cand_list = d_cand_list;
  a.cand_count = d_counters + 4;
  a.cand_src = d_cand_src;
  a.cand_flag = d_cand_flag;
  a.cand_gen = d_counters + 5;
  a.list = d_list + (size_t)sel * (size_t)total_blocks;
  a.list_count = d_counters + sel;
  a.other_count = d_counters + (sel ^ 1);
  a.counters = d_counters;
  a.chunks = d_chunks;
  a.keys = d_sorted_keys;
  a.nchunks = d_nchunks;
  a.block_scene = d_block_scene;
  a.nscenes = (int)scenes_host.size();
  a.list_cap = (uint32_t)total_blocks;
  a.oob = d_oob;
  a.hard_list = d_hard;
  a.store_vc = store_vc ? 1 : 0;
  a.hard_count = d_counters + (sel ? 6 : 3);
  a.other_hard = d_counters + (sel ? 3 : 6);
  a.defer_ticket = d_counters + 7;
  a.fallback = d_fallback;
  a.sums = sums;
  a.max_steps = max_steps;
  a.stage = d_stage;
  a.stage_cap = stage_cap;
  a.bfirst = d_bfirst;
  a.bcount = d_bcount;
  a.pm = corr_on ? d_pm : nullptr;
  a.cellf = corr_on ? d_cellf : nullptr;
  a.planes = d_planes;
  a.nplanes = (int)planes_host.size();
  return a;
}

int Engine::upload(const mpm_particles* p, const int32_t* scene_ids) {
  // an empty state (n = 0) needs no arrays
  if (!p || p->dtype != MPM_F32 || p->n < 0 || p->n > n_max ||
      (p->n > 0 && (!p->x || !p->v || !p->F || !p->C || !p->Jp || !p->mass || !p->mu0 || !p->lambda0))) {
    set_error("mpm_state_upload: needs FP32 arrays x, v, F, C, Jp, mass, mu0, lambda0 and n <= n_max");
    return MPM_ERR_ARG;
  }
  if (nmat > 1 && p->n > 0 && !p->material) {
    set_error("mpm_state_upload: material ids are required when the table has more than one entry");
    return MPM_ERR_ARG;
  }
  if (p->n != n) drop_graphs();
  n = p->n;
  cur = 0;
  since_bin = -1;
  if (n > 0)
    hipLaunchKernelGGL(upload_kernel, dim3(nblk(n, 256)), dim3(256), 0, own_stream, st[0], n,
                       (const float*)p->x, (const float*)p->v, (const float*)p->F, (const float*)p->C,
                       (const float*)p->Jp, (const float*)p->mass, (const float*)p->mu0,
                       (const float*)p->lambda0, p->mat_scalar, p->material, scene_ids,
                       (int)scenes_host.size());
  HIPCHK(hipGetLastError());
  return MPM_OK;
}

int Engine::download(mpm_particles* p) {
  if (!p || p->dtype != MPM_F32 || p->n != n || (n > 0 && (!p->x || !p->v || !p->F || !p->C || !p->Jp))) {
    set_error("mpm_state_download: needs FP32 arrays x, v, F, C, Jp with the uploaded n");
    return MPM_ERR_ARG;
  }
  if (n > 0)
    hipLaunchKernelGGL(download_kernel, dim3(nblk(n, 256)), dim3(256), 0, own_stream, st[cur], n,
                       (float*)p->x, (float*)p->v, (float*)p->F, (float*)p->C, (float*)p->Jp);
  HIPCHK(hipGetLastError());
  return MPM_OK;
}

// Keys -> stable radix sort -> block ranges -> chunk list -> candidate blocks
// and their neighbour table.  The particles themselves move into the new
// order in the segment's first P2G (through pending_perm).
void Engine::enqueue_bin(hipStream_t s, int len, int with_sums) {
  if (n <= 0) {
    hipLaunchKernelGGL(segment_begin_kernel, dim3(1), dim3(64), 0, s, d_counters, len, with_sums, 1);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, d_nchunks, (int64_t)1, 0u);
    return;
  }
  EngineArgs a = args(nullptr, 0, 0);
  const int64_t tiles = radix_tiles(n);
  hipLaunchKernelGGL(engine_keys_kernel, dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, a, n, d_keys, d_bocc,
                     (int64_t)total_blocks, len, with_sums, radix_first_hist(d_ws, n));
  uint32_t *ks, *vs;
  const uint32_t max_key = (uint32_t)(total_blocks * 64 - 1);
  radix_sort(d_keys, d_vals, n, max_key, d_ws, true, &ks, &vs, s, true);
  // the particles move into binned order in the segment's first P2G, which
  // reads slot perm[p] of the current buffer and writes slot p of the other
  cur ^= 1;
  pending_perm = vs;
  hipLaunchKernelGGL(block_bounds_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, ks, n, d_bstart, d_bend, d_bocc);
  chunk_list(d_bocc, d_bstart, d_bend, d_block_scene, (int64_t)total_blocks, chunk, d_part, d_chunks, d_bfirst,
             d_bcount, d_nchunks, s);
  d_sorted_keys = ks;
  hipLaunchKernelGGL(cand_build_kernel, dim3(num_cus * 2), dim3(256), 0, s, args(nullptr, 0, 0), d_cand_list,
                     d_counters + 4, d_cand_flag);
  hipLaunchKernelGGL(cand_src_kernel, dim3(num_cus * 2), dim3(256), 0, s, args(nullptr, 0, 0), d_cand_src);
}

void Engine::launch_transfer(int mode, hipStream_t s, const EngineArgs& a, bool separate_deferred) {
  const int g = num_cus * grid_mult;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // timing mode: the events take the kernel's own start and end timestamps
  // (hipExtLaunchKernel), not those of markers around it
  if (timing) {
    e0 = tevent();
    e1 = tevent();
  }
#define MPM_LAUNCH_K(K)                                                                              \
  do {                                                                                               \
    if (timing)                                                                                      \
      hipExtLaunchKernelGGL(K, dim3(g), dim3(ENG_THREADS), 0, s, e0, e1, 0, a);                      \
    else                                                                                             \
      hipLaunchKernelGGL(K, dim3(g), dim3(ENG_THREADS), 0, s, a);                                    \
  } while (0)
#define MPM_LAUNCH_T(M, PV)                                                                          \
  do {                                                                                               \
    if (M != 2 && spread == 9)                                                                       \
      MPM_LAUNCH_K((transfer_kernel<M, PV, (M == 2 ? 3 : 9)>));                                      \
    else                                                                                             \
      MPM_LAUNCH_K((transfer_kernel<M, PV, 3>));                                                     \
  } while (0)
  if (dyn_sched && mode == 1 && chunk == 512 && spread == 9) {
    MPM_LAUNCH_K((transfer_kernel<1, 2, 9, true>));
  } else if (chunk == 256) {
    if (mode == 0) MPM_LAUNCH_T(0, 1);
    else if (mode == 1) MPM_LAUNCH_T(1, 1);
    else if (mode == 2) MPM_LAUNCH_T(2, 1);
    else MPM_LAUNCH_T(3, 1);
  } else if (chunk == 512) {
    if (mode == 0) MPM_LAUNCH_T(0, 2);
    else if (mode == 1) MPM_LAUNCH_T(1, 2);
    else if (mode == 2) MPM_LAUNCH_T(2, 2);
    else MPM_LAUNCH_T(3, 2);
  } else {
    if (mode == 0) MPM_LAUNCH_T(0, 4);
    else if (mode == 1) MPM_LAUNCH_T(1, 4);
    else if (mode == 2) MPM_LAUNCH_T(2, 4);
    else MPM_LAUNCH_T(3, 4);
  }
#undef MPM_LAUNCH_T
#undef MPM_LAUNCH_K
  // timing slots: 0 the segment-start pass (P2G, or the moving fused pass),
  // 1 fused, 2 G2P (3 is the grid update's)
  if (timing) tpairs.push_back({mode == 3 ? 0 : mode, e0, e1});
  // one workgroup per CU: the usual empty list costs each workgroup one load,
  // and a violent step that defers tens of thousands of particles to FP64
  // (seen with dt several times the stable step) is spread over the chip
  if (separate_deferred || mode == 2)
    hipLaunchKernelGGL(deferred_kernel, dim3(num_cus), dim3(ENG_THREADS), 0, s, a);
}

void Engine::launch_grid(hipStream_t s, const EngineArgs& a, int with_sums, int len, int k, int deferred) {
  // with sums every workgroup ends with 7 FP64 atomics on the same 7 words
  const int g = with_sums ? num_cus : num_cus * 8;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timing) {
    e0 = tevent();
    e1 = tevent();
    hipExtLaunchKernelGGL(grid_update_kernel, dim3(g), dim3(ENG_THREADS), 0, s, e0, e1, 0, a, with_sums, len, k,
                          deferred);
    tpairs.push_back({3, e0, e1});
  } else {
    hipLaunchKernelGGL(grid_update_kernel, dim3(g), dim3(ENG_THREADS), 0, s, a, with_sums, len, k, deferred);
  }
}

hipEvent_t Engine::tevent() {
  if (tnext < tevents.size()) return tevents[tnext++];
  hipEvent_t e;
  hipEventCreate(&e);
  tevents.push_back(e);
  ++tnext;
  return e;
}

// One segment of `len` substeps: bin, P2G, grid, (G2P2G, grid) x (len-1), G2P.
void Engine::enqueue_frame(hipStream_t s, float* frames, int64_t step) {
  if (!frames || n <= 0) return;
  hipLaunchKernelGGL(positions_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, st[cur], n,
                     frames + (size_t)step * 3 * (size_t)n);
}

void Engine::fill_u32(hipStream_t s, uint32_t* p, int64_t count, uint32_t v) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(nblk(count, 256)), dim3(256), 0, s, p, count, v);
}

void Engine::enqueue_state(hipStream_t s, const StateSink* sink, int64_t step) {
  if (!sink || n <= 0) return;
  const size_t o = (size_t)step * (size_t)n;
  hipLaunchKernelGGL(download_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, st[cur], n, sink->x + 3 * o,
                     sink->v + 3 * o, sink->F + 9 * o, sink->C + 9 * o, sink->Jp + o);
}

// cont: the run's previous segment ended after its last grid update (no
// G2P), and this one rebins; its first pass is then the fused G2P of that
// substep plus this segment's first P2G, moving the particles into the new
// order (MODE 3) instead of a G2P, the rebin and a moving P2G.  fin: the
// run's last segment, which ends with the G2P.
void Engine::enqueue_segment(int len, bool bin, hipStream_t s, double* sums, int max_steps, float* frames,
                             int64_t step0, const StateSink* sink, bool cont, bool fin) {
  const int ws = sums ? 1 : 0;
  if (bin)
    enqueue_bin(s, len, ws);
  else
    hipLaunchKernelGGL(segment_begin_kernel, dim3(1), dim3(64), 0, s, d_counters, len, ws, 0);
  // substep k appends to active list k % 2; its grid update clears the other
  const size_t rec_nn = (size_t)(scenes_host[0].res + 1) * (scenes_host[0].res + 1) * (scenes_host[0].res + 1);
  auto grid_args = [&](int k) {
    EngineArgs a = args(sums, max_steps, k & 1);
    if (sink && sink->grids) a.grid_rec = sink->grids + (size_t)(step0 + k) * 7 * rec_nn;
    return a;
  };
  // the deferred FP64 particles are finished by the grid update unless the
  // particles are read in between (cell-F sums, per-substep state output);
  // per-substep state output also needs each substep's v and C in memory
  const bool sep = corr_on || sink != nullptr;
  store_vc = sink != nullptr;
  {
    EngineArgs a0 = args(sums, max_steps, 0);
    a0.perm = bin ? pending_perm : nullptr;
    pending_perm = nullptr;
    launch_transfer(cont ? 3 : 0, s, a0, sep);
  }
  if (cont) {  // the previous segment's last substep finished its G2P
    enqueue_frame(s, frames, step0 - 1);
    enqueue_state(s, sink, step0 - 1);
  }
  if (corr_on) launch_cellf(s, args(sums, max_steps, 0));
  launch_grid(s, grid_args(0), ws, len, 0, sep ? 0 : 1);
  if (corr_on) launch_corrector(s, args(sums, max_steps, 0), nullptr);
  for (int k = 1; k < len; ++k) {
    launch_transfer(1, s, args(sums, max_steps, k & 1), sep);
    enqueue_frame(s, frames, step0 + k - 1);  // substep k-1 finished its G2P
    enqueue_state(s, sink, step0 + k - 1);
    if (corr_on) launch_cellf(s, args(sums, max_steps, k & 1));
    launch_grid(s, grid_args(k), ws, len, k, sep ? 0 : 1);
    if (corr_on) launch_corrector(s, args(sums, max_steps, k & 1), nullptr);
  }
  if (fin) {
    launch_transfer(2, s, args(sums, max_steps, len & 1));
    enqueue_frame(s, frames, step0 + len - 1);
    enqueue_state(s, sink, step0 + len - 1);
  }
  store_vc = false;
}

int Engine::run(int nsteps, double* sums, float* frames, const StateSink* sink) {
  if (nsteps < 0) {
    set_error("mpm_run: nsteps must be >= 0");
    return MPM_ERR_ARG;
  }
  if (nsteps == 0) return MPM_OK;
  HIPCHK(hipEventRecord(ev_in, user_stream));
  HIPCHK(hipStreamWaitEvent(own_stream, ev_in, 0));
  if (sums) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3(nblk((int64_t)nsteps * 8, 256)), dim3(256), 0, own_stream,
                       sums, (int64_t)nsteps * 8, 0.0);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, own_stream, d_counters + 2, (int64_t)1, 0u);
  }
  if (sink && sink->grids) {  // nodes no particle reaches stay zero
    const int64_t r1 = scenes_host[0].res + 1;
    const int64_t cnt = (int64_t)nsteps * 7 * r1 * r1 * r1;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(nblk(cnt, 256)), dim3(256), 0, own_stream,
                       reinterpret_cast<uint32_t*>(sink->grids), cnt, 0u);
  }
  // Segments: the particles are rebinned when they never were since the
  // upload (or a rebin request) or after sort_every substeps; the schedule
  // carries over from one call to the next, so callers that advance a few
  // substeps per call do not pay a rebin per call.  Every call ends with
  // complete particle state (G2P), so a call costs one split P2G / G2P pair
  // beyond the fused substeps.  Inside a call a segment ends after its last
  // grid update and the next one (which always rebins: segments are cut at
  // the schedule's rebins) starts with the fused G2P of that substep, which
  // also moves the particles into the new order.
  int done = 0;
  while (done < nsteps) {
    const bool bin = since_bin < 0 || since_bin >= sort_every || n <= 0;
    const int room = bin ? sort_every : sort_every - since_bin;
    const int len = (nsteps - done) < room ? (nsteps - done) : room;
    const bool cont = n > 0 && done > 0;
    const bool fin = n <= 0 || done + len == nsteps;
    if (cont && !bin) {
      set_error("mpm_run: internal error: a continued segment must rebin");
      return MPM_ERR_HIP;
    }
    const bool graphed = use_graph && !timing && n > 0 && !frames && !sink;
    if (graphed) {
      const int parity = cur;
      const GraphKey key{n, len, parity, bin ? 1 : 0, (uintptr_t)sums, sums ? nsteps : 0, cont ? 1 : 0,
                         fin ? 1 : 0};
      auto it = graphs.find(key);
      hipGraphExec_t exec = nullptr;
      if (it == graphs.end()) {
        if (graphs.size() >= MAX_GRAPHS) drop_graphs();
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(own_stream, hipStreamCaptureModeThreadLocal));
        enqueue_segment(len, bin, own_stream, sums, nsteps, nullptr, 0, nullptr, cont, fin);
        HIPCHK(hipStreamEndCapture(own_stream, &graph));
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        graphs[key] = exec;
        cur = parity;  // capture toggled the host-side buffer index; replay does it below
      } else {
        exec = it->second;
      }
      HIPCHK(hipGraphLaunch(exec, own_stream));
      cur = bin ? parity ^ 1 : parity;
    } else {
      enqueue_segment(len, bin, own_stream, sums, nsteps, frames, done, sink, cont, fin);
    }
    HIPCHK(hipGetLastError());
    since_bin = bin ? len : since_bin + len;
    last_list = (len - 1) & 1;
    done += len;
  }
  HIPCHK(hipEventRecord(ev_out, own_stream));
  HIPCHK(hipStreamWaitEvent(user_stream, ev_out, 0));
  return MPM_OK;
}

void Engine::drop_graphs() {
  if (graphs.empty()) return;
  hipStreamSynchronize(own_stream);  // a launched graph may still be running
  for (auto& kv : graphs)
    if (kv.second) hipGraphExecDestroy(kv.second);
  graphs.clear();
}

// Blocks the last substep's P2G touched (the grid update's work list).
int Engine::active_blocks(int64_t* out) {
  HIPCHK(hipStreamSynchronize(own_stream));
  uint32_t v[5] = {0, 0, 0, 0, 0};
  HIPCHK(hipMemcpy(v, d_counters, sizeof(v), hipMemcpyDeviceToHost));
  *out = (int64_t)v[4] + (int64_t)v[last_list];
  return MPM_OK;
}

int Engine::positions(float* x) {
  if (n == 0) return MPM_OK;
  if (!x) {
    set_error("mpm_state_positions: null output");
    return MPM_ERR_ARG;
  }
  enqueue_frame(own_stream, x, 0);
  HIPCHK(hipGetLastError());
  return MPM_OK;
}

int Engine::take_counter(int which, int64_t* out) {
  HIPCHK(hipStreamSynchronize(own_stream));
  unsigned long long v = 0;
  unsigned long long* p = which == 0 ? d_oob : d_fallback;
  HIPCHK(hipMemcpy(&v, p, sizeof(v), hipMemcpyDeviceToHost));
  // cleared on the engine's own stream, ahead of its later work
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, own_stream, reinterpret_cast<uint32_t*>(p),
                     (int64_t)2, 0u);
  HIPCHK(hipStreamSynchronize(own_stream));
  *out = (int64_t)v;
  return MPM_OK;
}

int Engine::particle_sums(double* out) {
  double* d = nullptr;
  int rc;
  if ((rc = alloc((void**)&d, sizeof(double) * 8))) return rc;
  // cleared on the engine's own stream (a null-stream memset would not be
  // ordered with it)
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(64), 0, own_stream, d, (int64_t)8, 0.0);
  if (n > 0)
    hipLaunchKernelGGL(particle_sums_kernel, dim3(256), dim3(ENG_THREADS), 0, own_stream, st[cur], n, d);
  HIPCHK(hipStreamSynchronize(own_stream));
  HIPCHK(hipMemcpy(out, d, sizeof(double) * 7, hipMemcpyDeviceToHost));
  allocs.pop_back();
  HIPCHK(hipFree(d));
  return MPM_OK;
}

int Engine::kernel_times(double* ms, int64_t* counts) {
  HIPCHK(hipStreamSynchronize(own_stream));
  for (int i = 0; i < 4; ++i) {
    ms[i] = 0;
    counts[i] = 0;
  }
  for (auto& tp : tpairs) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, tp.e0, tp.e1));
    ms[tp.kind] += t;
    counts[tp.kind] += 1;
  }
  tpairs.clear();
  tnext = 0;
  return MPM_OK;
}

}  // namespace mpm

namespace mpm {

// Replace the sticky half-space colliders of one scene (count 0 removes them).
// Captured graphs hold the old table size, so they are dropped.
int Engine::set_planes(int scene, int count, const double* points, const double* normals, int quirks) {
  if (scene < 0 || scene >= (int)scenes_host.size() || count < 0 || (count > 0 && (!points || !normals)) ||
      (quirks != MPM_QUIRKS_REFERENCE && quirks != MPM_QUIRKS_TEXTBOOK)) {
    set_error("mpm_ctx_set_planes: bad scene, count, arrays or quirks mode");
    return MPM_ERR_ARG;
  }
  std::vector<PlaneDev> next;
  for (const PlaneDev& p : planes_host)
    if (p.scene != scene) next.push_back(p);
  const double dx = 1.0 / scenes_host[scene].res;
  for (int i = 0; i < count; ++i) {
    PlaneDev p{};
    if (!make_plane(points + 3 * i, normals + 3 * i, dx, quirks, scene, p)) {
      set_error("mpm_ctx_set_planes: a normal is zero or not finite");
      return MPM_ERR_ARG;
    }
    next.push_back(p);
  }
  if (next.size() > (size_t)MAX_PLANES) {
    set_error("mpm_ctx_set_planes: more than 16 planes in the context");
    return MPM_ERR_ARG;
  }
  int rc;
  if (!d_planes && (rc = alloc((void**)&d_planes, sizeof(PlaneDev) * MAX_PLANES))) return rc;
  if (hipStreamSynchronize(own_stream) != hipSuccess) {
    set_error("mpm_ctx_set_planes: synchronising the engine stream failed");
    return MPM_ERR_HIP;
  }
  drop_graphs();
  if (!next.empty() &&
      (hipMemcpyAsync(d_planes, next.data(), sizeof(PlaneDev) * next.size(), hipMemcpyHostToDevice, own_stream) !=
           hipSuccess ||
       hipStreamSynchronize(own_stream) != hipSuccess)) {
    set_error("mpm_ctx_set_planes: copying the plane table failed");
    return MPM_ERR_HIP;
  }
  planes_host = next;
  return MPM_OK;
}

}  // namespace mpm
