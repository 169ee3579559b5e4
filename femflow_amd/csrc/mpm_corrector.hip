// Learned grid corrector inside the engine's substep.
//
// With the corrector on, every substep (taken at dt_big) runs
//   transfer (G2P(n-1) + P2G(n)) -> cellf_kernel -> grid update (also
//   writing each node's P2G momentum/mass) -> corrector_kernel
// and corrector_kernel, one workgroup per active block, gathers the
// 3x3x3-node stencil features of the block's 64 nodes into LDS, runs the
// 351 -> 128 -> 128 -> 3 MLP on MFMA (BF16 operands, FP32 accumulate) and
// adds the predicted correction, scaled down on nodes with little mass, to
// the nodes' updated velocities (re-clamped, walls re-applied), then clears
// the block's cell sums for the next substep.
// In sampling mode the same kernel writes the features (and the updated
// velocity) of every node with mass instead: the training data.
//
// Features of node n, for each neighbour o = (d0, d1, d2) in {-1, 0, 1}^3 in
// lexicographic order, 13 values: m / m_ref, p / (m_ref v_ref) (3), and the
// mass-weighted mean deformation gradient of the particles whose stencil base
// is node n + o, minus I (9; zero when that cell is empty).  Output: the
// velocity correction divided by v_ref.
#include <hip/hip_runtime.h>

#include "mpm_engine.h"
#include "mpm_engine_kernels.h"
#include "mpm_math.h"
#include "mpm_mlp.h"

namespace mpm {

struct CorrArgs {
  const uint16_t* packed;  // BF16 operand copies
  const float* params;     // FP32 master (biases read from here)
  float inv_mref, inv_pref, vref;
  CorrExtract ex;          // ex.X != null: sampling mode
  uint32_t* count;
};

// ------------------------------------------------------------ cell sums ----
// Particles in binned order: the lanes of a wave mostly share cells, so the
// wave sums runs of equal cells (segmented scan over lanes) and only the
// last lane of each run adds to global memory.
__global__ __launch_bounds__(256) void cellf_kernel(EngineArgs a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t key = 0xffffffffu;
  float val[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) val[c] = 0.f;
  if (i < n) {
    const SceneDev sc = a.scenes[a.s.attr[i] >> 8];
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = a.s.x[tix(i, 3, d)];
    int base[3];
    float fx[3];
    if (locate3(x, sc, base, fx)) {
      key = (uint32_t)node_addr(sc, base[0], base[1], base[2]);
      const float m = a.s.mass[i];
      val[0] = m;
#pragma unroll
      for (int e = 0; e < 9; ++e) val[1 + e] = m * a.s.F[tix(i, 9, e)];
    }
  }
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const int seg = __popcll(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull)));
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int sup = __shfl_up(seg, off);
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const float up = __shfl_up(val[c], off);
      if (lane >= off && sup == seg) val[c] += up;
    }
  }
  const int snext = __shfl_down(seg, 1);
  const bool tail = lane == 63 || snext != seg;
  if (tail && key != 0xffffffffu) {
    float* dst = reinterpret_cast<float*>(a.cellf + (size_t)key * 3);
#pragma unroll
    for (int c = 0; c < 10; ++c) atomicAdd(dst + c, val[c]);
    mark_active(a, key >> 6);
  }
}

// ----------------------------------------------------------- corrector ----
constexpr int XS = 360;  // LDS row stride of the feature tile (BF16)
// The correction applied to a node is scaled by min(1, m / (CORR_GATE m_ref)):
// nodes with little mass (surfaces) keep the plain big step's velocity.
constexpr float CORR_GATE = 0.25f;
constexpr int HS = 136;  // LDS row stride of the hidden tiles (BF16)

__global__ __launch_bounds__(256) void corrector_kernel(EngineArgs a, CorrArgs c) {
  __shared__ float reg_s[216][13];
  __shared__ __align__(16) uint16_t x_s[64 * XS];  // features, then the second hidden layer
  __shared__ __align__(16) uint16_t h_s[64 * HS];  // first hidden layer
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const uint32_t ncand = *a.cand_count;
  const uint32_t count = ncand + *a.list_count;
  for (uint32_t b = blockIdx.x; b < count; b += gridDim.x) {
    const uint32_t g = active_block(a, b, ncand);
    const SceneDev sc = a.scenes[a.block_scene[g]];
    const int local = (int)g - sc.block_base;
    const int nb = sc.nb;
    const int bi = local / (nb * nb), bj = (local / nb) % nb, bk = local % nb;
    // 6x6x6 node neighbourhood of the block -> normalised node features
    if (t < 216) {
      const int gi = bi * 4 - 1 + t / 36, gj = bj * 4 - 1 + (t / 6) % 6, gk = bk * 4 - 1 + t % 6;
      float f[13];
#pragma unroll
      for (int e = 0; e < 13; ++e) f[e] = 0.f;
      if (gi >= 0 && gj >= 0 && gk >= 0 && gi <= sc.res && gj <= sc.res && gk <= sc.res) {
        const size_t na = node_addr(sc, gi, gj, gk);
        const float4 q = a.pm[na];  // zero outside this substep's active blocks
        if (q.w > 0.f) {
          f[0] = q.w * c.inv_mref;
          f[1] = q.x * c.inv_pref;
          f[2] = q.y * c.inv_pref;
          f[3] = q.z * c.inv_pref;
        }
        const float4 c0 = a.cellf[na * 3], c1 = a.cellf[na * 3 + 1], c2 = a.cellf[na * 3 + 2];
        if (c0.x > 0.f) {
          const float im = 1.f / c0.x;
          const float F[9] = {c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y};
#pragma unroll
          for (int e = 0; e < 9; ++e) f[4 + e] = F[e] * im - ((e % 4 == 0) ? 1.f : 0.f);
        }
      }
#pragma unroll
      for (int e = 0; e < 13; ++e) reg_s[t][e] = f[e];
    }
    __syncthreads();
    // feature rows: node r of the block, neighbour o, value e
    for (int idx = t; idx < 64 * 27; idx += 256) {
      const int r = idx / 27, o = idx % 27;
      const int o0 = r >> 4, o1 = (r >> 2) & 3, o2 = r & 3;
      const int reg = ((o0 + o / 9) * 6 + (o1 + (o / 3) % 3)) * 6 + (o2 + o % 3);
      uint16_t* row = x_s + r * XS + o * 13;
#pragma unroll
      for (int e = 0; e < 13; ++e) row[e] = f2bf(reg_s[reg][e]);
    }
    if (t < 64) x_s[t * XS + MLP_IN] = 0;
    __syncthreads();
    if (c.ex.X) {
      // sampling: every node with mass becomes one training row
      if (t < 64) {
        const size_t na = (size_t)g * 64 + t;
        const float4 q = a.pm[na];
        if (q.w > 0.f) {
          const uint32_t row = atomicAdd(c.count, 1u);
          if ((int64_t)row < c.ex.max_rows) {
            const float4 v = a.vel[na];
            c.ex.nodes[row] = (uint32_t)na;
            c.ex.vel[3 * (size_t)row] = v.x;
            c.ex.vel[3 * (size_t)row + 1] = v.y;
            c.ex.vel[3 * (size_t)row + 2] = v.z;
            const uint4* src = reinterpret_cast<const uint4*>(x_s + t * XS);
            uint4* dst = reinterpret_cast<uint4*>(c.ex.X + (size_t)row * MLP_INP);
            for (int e = 0; e < MLP_INP / 8; ++e) dst[e] = src[e];
          }
        }
      }
      __syncthreads();
      continue;
    }
    // layer 1: 64 x 352 -> 128 (ReLU)
    {
      floatx4 acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
      const uint16_t* W1 = c.packed + K_W1;
      for (int k0 = 0; k0 < MLP_INP; k0 += 32) {
        const uint4 av = *reinterpret_cast<const uint4*>(x_s + (wv * 16 + li) * XS + k0 + lk);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint4 bv = *reinterpret_cast<const uint4*>(W1 + (size_t)(16 * u + li) * MLP_INP + k0 + lk);
          acc[u] = mfma16x16x32(av, bv, acc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int col = 16 * u + li;
        const float bb = c.params[P_B1 + col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc[u][r] + bb;
          h_s[(wv * 16 + (lane >> 4) * 4 + r) * HS + col] = f2bf(v > 0.f ? v : 0.f);
        }
      }
    }
    __syncthreads();
    // layer 2: 128 -> 128 (ReLU), into x_s
    {
      floatx4 acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
      const uint16_t* W2 = c.packed + K_W2;
      for (int k0 = 0; k0 < MLP_H; k0 += 32) {
        const uint4 av = *reinterpret_cast<const uint4*>(h_s + (wv * 16 + li) * HS + k0 + lk);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint4 bv = *reinterpret_cast<const uint4*>(W2 + (size_t)(16 * u + li) * MLP_H + k0 + lk);
          acc[u] = mfma16x16x32(av, bv, acc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int col = 16 * u + li;
        const float bb = c.params[P_B2 + col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc[u][r] + bb;
          x_s[(wv * 16 + (lane >> 4) * 4 + r) * HS + col] = f2bf(v > 0.f ? v : 0.f);
        }
      }
    }
    __syncthreads();
    // layer 3: 128 -> 3, add to the node velocities
    {
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      const uint16_t* W3 = c.packed + K_W3;
      for (int k0 = 0; k0 < MLP_H; k0 += 32) {
        const uint4 av = *reinterpret_cast<const uint4*>(x_s + (wv * 16 + li) * HS + k0 + lk);
        const uint4 bv = *reinterpret_cast<const uint4*>(W3 + (size_t)li * MLP_H + k0 + lk);
        acc = mfma16x16x32(av, bv, acc);
      }
      if (li < MLP_OUT) {
        const float bb = c.params[P_B3 + li];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int nr = wv * 16 + (lane >> 4) * 4 + r;
          const size_t na = (size_t)g * 64 + nr;
          const float mn = a.pm[na].w;
          if (!(mn > 0.f)) continue;
          const int I[3] = {bi * 4 + (nr >> 4), bj * 4 + ((nr >> 2) & 3), bk * 4 + (nr & 3)};
          float* vp = reinterpret_cast<float*>(a.vel + na) + li;
          const float gate = tmin(1.f, mn * c.inv_mref * (1.f / CORR_GATE));
          float v = *vp + gate * (acc[r] + bb) * c.vref;
          v = tmin(tmax(v, -sc.v_allowed), sc.v_allowed);
          if (sc.boundary == MPM_BC_REFERENCE) {
            if (I[li] < 1 || I[li] >= sc.res - 1) v = 0.f;
          } else {
#pragma unroll
            for (int d = 0; d < 3; ++d)
              if (I[d] < sc.bound || I[d] > sc.res - sc.bound) v = 0.f;
          }
          if (I[0] > sc.res || I[1] > sc.res || I[2] > sc.res) v = 0.f;
          *vp = v;
        }
      }
    }
    __syncthreads();
  }
}

// Clear the cell sums and P2G records of this substep's active blocks
// (every cell that received a particle lies in one: cellf_kernel marks it),
// so that blocks outside the next substep's active set read as empty.
__global__ __launch_bounds__(256) void cellf_clear_kernel(EngineArgs a) {
  const uint32_t ncand = *a.cand_count;
  const uint32_t count = ncand + *a.list_count;
  const int t = threadIdx.x;
  for (uint32_t b = blockIdx.x; b < count; b += gridDim.x) {
    const uint32_t g = active_block(a, b, ncand);
    float4* p = a.cellf + (size_t)g * 64 * 3;
    if (t < 192) p[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < 64) a.pm[(size_t)g * 64 + t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__global__ void gather_vel_kernel(const float4* vel, const uint32_t* nodes, int64_t count, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float4 v = vel[nodes[i]];
  out[3 * i] = v.x;
  out[3 * i + 1] = v.y;
  out[3 * i + 2] = v.z;
}

static inline unsigned cblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void Engine::launch_cellf(hipStream_t s, const EngineArgs& a) {
  if (n > 0) hipLaunchKernelGGL(cellf_kernel, dim3(cblk(n, 256)), dim3(256), 0, s, a, n);
}

void Engine::launch_corrector(hipStream_t s, const EngineArgs& a, const CorrExtract* ex) {
  CorrArgs c{};
  c.packed = d_corr_packed;
  c.params = d_corr_params;
  c.inv_mref = 1.f / corr_mref;
  c.inv_pref = 1.f / (corr_mref * corr_vref);
  c.vref = corr_vref;
  if (ex) c.ex = *ex;
  c.count = d_corr_count;
  const int g = num_cus * 2;
  hipLaunchKernelGGL(corrector_kernel, dim3(g), dim3(256), 0, s, a, c);
  hipLaunchKernelGGL(cellf_clear_kernel, dim3(g), dim3(256), 0, s, a);
}

int Engine::set_corrector(const float* params, float m_ref, float v_ref, int enable) {
  if ((enable && !params) || (params && (!(m_ref > 0.f) || !(v_ref > 0.f)))) {
    set_error("mpm_ctx_set_corrector: needs parameters and positive m_ref, v_ref");
    return MPM_ERR_ARG;
  }
  hipStreamSynchronize(own_stream);
  drop_graphs();
  int rc;
  if ((enable || params) && !d_pm) {
    const size_t nodes = (size_t)total_blocks * 64;
    if ((rc = alloc((void**)&d_pm, sizeof(float4) * nodes))) return rc;
    if ((rc = alloc((void**)&d_cellf, sizeof(float4) * 3 * nodes))) return rc;
    if ((rc = alloc((void**)&d_corr_params, sizeof(float) * P_COUNT))) return rc;
    if ((rc = alloc((void**)&d_corr_packed, sizeof(uint16_t) * K_COUNT))) return rc;
    if ((rc = alloc((void**)&d_corr_count, sizeof(uint32_t) * 4))) return rc;
    // zero on the engine's own (non-blocking) stream, before any of its work
    fill_u32(own_stream, reinterpret_cast<uint32_t*>(d_cellf), (int64_t)nodes * 12, 0u);
    fill_u32(own_stream, reinterpret_cast<uint32_t*>(d_pm), (int64_t)nodes * 4, 0u);
    if (hipStreamSynchronize(own_stream) != hipSuccess) {
      set_error("mpm_ctx_set_corrector: clearing the feature buffers failed");
      return MPM_ERR_HIP;
    }
  }
  corr_on = enable != 0;
  if (params) {
    corr_mref = m_ref;
    corr_vref = v_ref;
    // the caller's parameters are ready on its stream: order the copy after
    // it, on the engine's stream (a device-to-device hipMemcpy would run on
    // the null stream, unordered with the engine's non-blocking stream)
    if (hipEventRecord(ev_in, user_stream) != hipSuccess || hipStreamWaitEvent(own_stream, ev_in, 0) != hipSuccess ||
        hipMemcpyAsync(d_corr_params, params, sizeof(float) * P_COUNT, hipMemcpyDeviceToDevice, own_stream) !=
            hipSuccess) {
      set_error("mpm_ctx_set_corrector: parameter copy failed");
      return MPM_ERR_HIP;
    }
    mlp_pack(own_stream, d_corr_params, d_corr_packed);
    if (hipStreamSynchronize(own_stream) != hipSuccess) {
      set_error("mpm_ctx_set_corrector: pack failed");
      return MPM_ERR_HIP;
    }
  }
  return MPM_OK;
}

// One P2G + grid update from the current state (particles unchanged), then
// the corrector's features of every node with mass.  X == null: stop after
// the grid update, applying the corrector if it is enabled
// (grid_velocity() then reads node velocities).
int Engine::corrector_sample(uint16_t* X, uint32_t* nodes, float* vel, int64_t max_rows, int64_t* rows) {
  if (!d_pm) {
    set_error("mpm_corrector_sample: set a corrector (mpm_ctx_set_corrector) first");
    return MPM_ERR_ARG;
  }
  const bool was_on = corr_on;
  corr_on = true;
  hipStream_t s = own_stream;
  if (hipEventRecord(ev_in, user_stream) != hipSuccess || hipStreamWaitEvent(s, ev_in, 0) != hipSuccess) {
    corr_on = was_on;
    set_error("mpm_corrector_sample: stream setup failed");
    return MPM_ERR_HIP;
  }
  enqueue_bin(s, 1, 0);  // also starts the one-substep segment
  since_bin = 0;  // freshly binned: the next run continues from this order
  const EngineArgs a = args(nullptr, 0, 0);
  {
    EngineArgs a0 = a;
    a0.perm = pending_perm;  // this P2G moves the particles into binned order
    pending_perm = nullptr;
    launch_transfer(0, s, a0);
  }
  launch_cellf(s, a);
  launch_grid(s, a, 0, 1, 0);
  fill_u32(s, d_corr_count, 1, 0u);
  CorrExtract ex{X, nodes, vel, max_rows};
  if (X) {
    launch_corrector(s, a, &ex);  // features only
  } else if (was_on) {
    launch_corrector(s, a, nullptr);  // the substep's corrected grid velocity
  } else {
    hipLaunchKernelGGL(cellf_clear_kernel, dim3(num_cus * 2), dim3(256), 0, s, a);
  }
  corr_on = was_on;
  uint32_t cnt = 0;
  if (hipMemcpyAsync(&cnt, d_corr_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("mpm_corrector_sample: kernel failure");
    return MPM_ERR_HIP;
  }
  if (rows) *rows = (int64_t)cnt;
  return MPM_OK;
}

int Engine::grid_velocity(const uint32_t* nodes, int64_t count, float* out) {
  if (count <= 0) return MPM_OK;
  hipLaunchKernelGGL(gather_vel_kernel, dim3(cblk(count, 256)), dim3(256), 0, own_stream, d_vel, nodes, count, out);
  if (hipStreamSynchronize(own_stream) != hipSuccess) {
    set_error("mpm_grid_velocity: kernel failure");
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

}  // namespace mpm
