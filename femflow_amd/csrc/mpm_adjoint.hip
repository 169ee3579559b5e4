// Reverse mode of the explicit 3D MLS-MPM substep for the elastic
// (fixed-corotated) material on a dense (res+1)^3 grid, FP32.
//
// The forward it differentiates is the dense phase sequence of
// mls_mpm.py:40-79 (P2G of three_d/p2g.py:14-80 with the stress of
// utils.py:95-135, the grid update of three_d/grid_op.py:5-47, G2P of
// three_d/g2p.py:9-59 without plasticity), exactly as the phase kernels run
// it.  One call reverses one substep: it recomputes that substep's grids from
// the saved particle state, then
//   1. adj_g2p:  particle -> grid scatter of the velocity adjoint, and the
//                G2P parts of the x and F adjoints;
//   2. adj_grid: node-local reverse of normalise / gravity / clamp / walls;
//   3. adj_p2g:  grid -> particle gather giving the adjoints of x, v, C, F
//                and of mu0, lambda0 (stress derivative including the
//                polar-factor derivative).
// Parameters that are discontinuous in x (the stencil base, the velocity
// clamp and the wall masks) are differentiated piecewise: the base is held
// fixed and a clamped or masked component passes no gradient.
#include <hip/hip_runtime.h>

#include <climits>

#include "mpm355.h"
#include "mpm_math.h"
#include "mpm_phase.h"

namespace mpm {

struct AdjArgs {
  int64_t n;
  int res, boundary, bound, textbook, mat_scalar;
  float inv_dx, dx, dt, vol, gravity, hardening, v_allowed;
  const float* x;
  const float* v;
  const float* F;
  const float* C;
  const float* mass;
  const float* mu0;
  const float* lam0;
  const float* gp;  // P2G momentum, 3 per node
  const float* gm;  // P2G mass
  const float* gv;  // updated grid velocity, 3 per node
  float* gbar;      // 4 per node: velocity adjoint, then (momentum, mass) adjoint
  float* xb;
  float* vb;
  float* Fb;
  float* Cb;
  float* mub;
  float* lamb;
  unsigned long long* oob;
};

__device__ __forceinline__ bool adj_locate(const float* xp, float inv_dx, int res, int base[3], float fx[3]) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float s = xp[d] * inv_dx;
    const float t = s - 0.5f;
    ok = ok && (t > -1.f) && (t < (float)(res - 1));
    base[d] = ok ? (int)t : 0;
    fx[d] = s - (float)base[d];
  }
  return ok;
}

// w[i] and dw[i]/dfx for the quadratic B-spline (three_d/p2g.py:55)
__device__ __forceinline__ void bspline_d(float fx, float w[3], float dw[3]) {
  bspline(fx, w);
  dw[0] = fx - 1.5f;
  dw[1] = -2.f * (fx - 1.f);
  dw[2] = fx - 0.5f;
}

// ---------------------------------------------------------------- G2P^T ----
// The node-velocity adjoint is scattered in two levels: a wave sums the lanes
// of each run of particles in one cell (seg_reduce), and the workgroup sums
// those into an LDS arena over the bounding box of its particles' stencils,
// which it then adds to the grid with one device atomic per nonzero node
// component.  Particles in cell order give a box of a few cells, so most of
// the adds stay on the chip; device-scope atomics, which serialise across
// the XCDs, were the kernel's cost.  A box larger than the arena falls back
// to one device atomic per segment.
constexpr int AG_THREADS = 256;
constexpr int AG_CAP = 1024;  // arena nodes (x 3 components)

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__global__ __launch_bounds__(AG_THREADS) void adj_g2p_kernel(AdjArgs a) {
  __shared__ float arena[AG_CAP * 3];
  __shared__ int box_w[AG_THREADS / 64][6];
  __shared__ int box[6];  // low corner (3), extent (3) of the stencil box
  // lanes past the end or out of the grid stay in the wave (its segmented
  // atomics need every lane) and write nothing
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool inrange = p0 < a.n;
  const int64_t p = inrange ? p0 : a.n - 1;
  int base[3];
  float fx[3];
  const bool located = adj_locate(a.x + 3 * p, a.inv_dx, a.res, base, fx);
  if (inrange && !located && a.oob) atomicAdd(a.oob, 1ull);
  const bool valid = inrange && located;
  if (!located) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      base[d] = 0;
      fx[d] = 1.f;
    }
  }
  float w[3][3], dw[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline_d(fx[d], w[d], dw[d]);
  const int64_t r1 = a.res + 1;
  const float c4 = 4.f * a.inv_dx;
  // forward gather (v', C') of this particle
  float vn[3] = {0.f, 0.f, 0.f}, Cn[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) Cn[e] = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float wt = w[0][i] * w[1][j] * w[2][k];
        const int64_t node = ((int64_t)(base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
        const float* g = a.gv + 3 * node;
        const float dp[3] = {(float)i - fx[0], (float)j - fx[1], (float)k - fx[2]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float wg = wt * g[d];
          vn[d] += wg;
#pragma unroll
          for (int e = 0; e < 3; ++e) Cn[3 * d + e] += c4 * wg * dp[e];
        }
      }
  // adjoints of the outputs x', v', C', F'
  float xo[3], vo[3], Co[9], Fo[9], F[9];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    xo[d] = a.xb[3 * p + d];
    vo[d] = a.vb[3 * p + d];
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    Co[e] = a.Cb[9 * p + e];
    Fo[e] = a.Fb[9 * p + e];
    F[e] = a.F[9 * p + e];
  }
  // F' = (I + dt C') F:  Fbar = (I + dt C')^T Fo,  Cbar += dt Fo F^T
  float Fbar[9], Ct[9], vt[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = Fo[3 * r + c];
      float u = Co[3 * r + c];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        s += a.dt * Cn[3 * m + r] * Fo[3 * m + c];
        u += a.dt * Fo[3 * r + m] * F[3 * c + m];
      }
      Fbar[3 * r + c] = s;
      Ct[3 * r + c] = u;
    }
  // x' = x + dt v'
#pragma unroll
  for (int d = 0; d < 3; ++d) vt[d] = vo[d] + a.dt * xo[d];
  // C' depends on fx through (n - fx):  dC'_ab/dfx_c = -c4 v'_a delta_bc
  float gx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) gx[c] = -c4 * (Ct[c] * vn[0] + Ct[3 + c] * vn[1] + Ct[6 + c] * vn[2]);
  const WaveSeg sg = wave_segments(valid, (int)((base[0] * r1 + base[1]) * r1 + base[2]));
  {
    const int t = threadIdx.x, wv = t >> 6;
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = wave_min(valid ? base[d] : INT_MAX);
      hi[d] = wave_max(valid ? base[d] : INT_MIN);
    }
    if ((t & 63) == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        box_w[wv][d] = lo[d];
        box_w[wv][3 + d] = hi[d];
      }
    }
    __syncthreads();
    if (t == 0) {
      int l[3] = {INT_MAX, INT_MAX, INT_MAX}, h[3] = {INT_MIN, INT_MIN, INT_MIN};
      for (int w2 = 0; w2 < AG_THREADS / 64; ++w2)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          l[d] = min(l[d], box_w[w2][d]);
          h[d] = max(h[d], box_w[w2][3 + d]);
        }
      const bool any = l[0] <= h[0];
      const int64_t e0 = any ? h[0] - l[0] + 3 : 0, e1 = any ? h[1] - l[1] + 3 : 0, e2 = any ? h[2] - l[2] + 3 : 0;
      const bool fits = any && e0 * e1 * e2 <= AG_CAP;
#pragma unroll
      for (int d = 0; d < 3; ++d) box[d] = l[d];
      box[3] = fits ? (int)e0 : 0;
      box[4] = fits ? (int)e1 : 0;
      box[5] = fits ? (int)e2 : 0;
    }
    __syncthreads();
    const int vol = box[3] * box[4] * box[5];
    for (int e = t; e < 3 * vol; e += AG_THREADS) arena[e] = 0.f;
    __syncthreads();
  }
  const int bl0 = box[0], bl1 = box[1], bl2 = box[2], be1 = box[4], be2 = box[5];
  const bool use_arena = box[3] > 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float wt = w[0][i] * w[1][j] * w[2][k];
        const int64_t node = ((int64_t)(base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
        const float* g = a.gv + 3 * node;
        const float dp[3] = {(float)i - fx[0], (float)j - fx[1], (float)k - fx[2]};
        // node velocity adjoint: vt + c4 Ct (n - fx), weighted
        float gb[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          gb[d] = vt[d] + c4 * (Ct[3 * d] * dp[0] + Ct[3 * d + 1] * dp[1] + Ct[3 * d + 2] * dp[2]);
        float* dst = a.gbar + 4 * node;
        const int slot = ((base[0] + i - bl0) * be1 + (base[1] + j - bl1)) * be2 + (base[2] + k - bl2);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float r = seg_reduce(wt * gb[d], sg);
          if (sg.tail) {
            if (use_arena) atomicAdd(&arena[3 * slot + d], r);
            else atomicAdd(dst + d, r);
          }
        }
        // weight derivative
        const float s = g[0] * gb[0] + g[1] * gb[1] + g[2] * gb[2];
        gx[0] += dw[0][i] * w[1][j] * w[2][k] * s;
        gx[1] += w[0][i] * dw[1][j] * w[2][k] * s;
        gx[2] += w[0][i] * w[1][j] * dw[2][k] * s;
      }
  if (valid) {
#pragma unroll
    for (int d = 0; d < 3; ++d) a.xb[3 * p + d] = xo[d] + a.inv_dx * gx[d];
#pragma unroll
    for (int e = 0; e < 9; ++e) a.Fb[9 * p + e] = Fbar[e];
  }
  if (!use_arena) return;  // workgroup-uniform
  __syncthreads();
  const int e1 = box[4], e2 = box[5], vol = box[3] * e1 * e2;
  for (int e = threadIdx.x; e < 3 * vol; e += AG_THREADS) {
    const float val = arena[e];
    if (val != 0.f) {
      const int sl = e / 3, d = e - 3 * sl;
      const int64_t I = bl0 + sl / (e1 * e2), J = bl1 + (sl / e2) % e1, K = bl2 + sl % e2;
      atomicAdd(a.gbar + 4 * ((I * r1 + J) * r1 + K) + d, val);
    }
  }
}

// --------------------------------------------------------------- grid^T ----
__global__ void adj_grid_kernel(AdjArgs a) {
  const int64_t r1 = a.res + 1;
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= r1 * r1 * r1) return;
  float* gb = a.gbar + 4 * node;
  const int64_t I[3] = {node / (r1 * r1), (node / r1) % r1, node % r1};
  float ub[3] = {gb[0], gb[1], gb[2]};
  if (a.boundary == MPM_BC_REFERENCE) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (I[d] < 1 || I[d] >= a.res - 1) ub[d] = 0.f;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (I[d] < a.bound || I[d] > a.res - a.bound) ub[0] = ub[1] = ub[2] = 0.f;
  }
  const float m = a.gm[node];
  float mb = 0.f;
  if (m > 0.f) {
    const float* pg = a.gp + 3 * node;
    const float im = 1.f / m;
    float u[3] = {pg[0] * im, pg[1] * im, pg[2] * im};
    const float uy = u[1] + a.dt * a.gravity;
    const float uu[3] = {u[0], uy, u[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (!(uu[d] > -a.v_allowed && uu[d] < a.v_allowed)) ub[d] = 0.f;  // clamped: no gradient
    mb = -(ub[0] * u[0] + ub[1] * u[1] + ub[2] * u[2]) * im;
#pragma unroll
    for (int d = 0; d < 3; ++d) ub[d] *= im;
  }
  gb[0] = ub[0];
  gb[1] = ub[1];
  gb[2] = ub[2];
  gb[3] = mb;
}

// ---------------------------------------------------------------- P2G^T ----
// dR for F = R S: R^T dF - dF^T R = K w with K = tr(S) I - S, dR = R [w]x.
// Reverse: Fbar += R [y]x with y = K^{-1} vee(B - B^T), B = R^T Rbar.
__device__ __forceinline__ void polar_adjoint(const float F[9], const float R[9], const float Rb[9],
                                              float Fb[9]) {
  float S[9], B[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = 0.f, b = 0.f;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        s += R[3 * m + r] * F[3 * m + c];
        b += R[3 * m + r] * Rb[3 * m + c];
      }
      S[3 * r + c] = s;
      B[3 * r + c] = b;
    }
  // symmetrise S (exact in exact arithmetic)
  const float s01 = 0.5f * (S[1] + S[3]), s02 = 0.5f * (S[2] + S[6]), s12 = 0.5f * (S[5] + S[7]);
  const float tr = S[0] + S[4] + S[8];
  float K[9] = {tr - S[0], -s01, -s02, -s01, tr - S[4], -s12, -s02, -s12, tr - S[8]};
  const float b[3] = {B[7] - B[5], B[2] - B[6], B[3] - B[1]};
  float Kc[9];
  cofactor3(K, Kc);  // K symmetric: K^{-1} = cof(K) / det(K)
  const float det = K[0] * Kc[0] + K[1] * Kc[1] + K[2] * Kc[2];
  if (!(tabs(det) > 1e-30f)) return;
  const float id = 1.f / det;
  float y[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) y[r] = (Kc[3 * r] * b[0] + Kc[3 * r + 1] * b[1] + Kc[3 * r + 2] * b[2]) * id;
  const float Y[9] = {0.f, -y[2], y[1], y[2], 0.f, -y[0], -y[1], y[0], 0.f};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Fb[3 * r + c] += R[3 * r] * Y[c] + R[3 * r + 1] * Y[3 + c] + R[3 * r + 2] * Y[6 + c];
}

__global__ __launch_bounds__(128) void adj_p2g_kernel(AdjArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  int base[3];
  float fx[3];
  if (!adj_locate(a.x + 3 * p, a.inv_dx, a.res, base, fx)) return;  // counted by adj_g2p
  float w[3][3], dw[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline_d(fx[d], w[d], dw[d]);
  const int64_t mi = a.mat_scalar ? 0 : p;
  const float m = a.mass[mi];
  const float mu = a.mu0[mi] * a.hardening, lam = a.lam0[mi] * a.hardening;
  float F[9], C[9], v[3], R[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    F[e] = a.F[9 * p + e];
    C[e] = a.C[9 * p + e];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] = a.v[3 * p + d];
  if (mu != 0.f) {
    polar3<float>(F, R);
  } else {
#pragma unroll
    for (int e = 0; e < 9; ++e) R[e] = F[e];
  }
  const float J = det3(F);
  const float scale = -(a.dt * a.vol) * (4.f * a.inv_dx * a.inv_dx);
  float A[9];
  {
    const float vt = lam * (J - 1.f) * J;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        float pf = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) pf += 2.f * mu * (F[3 * i + k] - R[3 * i + k]) * F[3 * j + k];
        pf += (a.textbook && i != j) ? 0.f : vt;
        A[3 * i + j] = scale * pf + m * C[3 * i + j];
      }
  }
  const float mv[3] = {m * v[0], m * v[1], m * v[2]};
  const int64_t r1 = a.res + 1;
  // gather
  float P[3] = {0.f, 0.f, 0.f}, Ab[9], gx[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int e = 0; e < 9; ++e) Ab[e] = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float wt = w[0][i] * w[1][j] * w[2][k];
        const int64_t node = ((int64_t)(base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
        const float* gb = a.gbar + 4 * node;
        const float pb[3] = {gb[0], gb[1], gb[2]};
        const float mb = gb[3];
        const float dp[3] = {((float)i - fx[0]) * a.dx, ((float)j - fx[1]) * a.dx, ((float)k - fx[2]) * a.dx};
        float s = mb * m;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          P[d] += wt * pb[d];
          const float c = mv[d] + A[3 * d] * dp[0] + A[3 * d + 1] * dp[1] + A[3 * d + 2] * dp[2];
          s += pb[d] * c;
#pragma unroll
          for (int e = 0; e < 3; ++e) Ab[3 * d + e] += wt * pb[d] * dp[e];
        }
        gx[0] += dw[0][i] * w[1][j] * w[2][k] * s;
        gx[1] += w[0][i] * dw[1][j] * w[2][k] * s;
        gx[2] += w[0][i] * w[1][j] * dw[2][k] * s;
      }
  // x: weights (inv_dx dW/dfx) and the (n - fx) dx offsets (-A^T P)
  float xb[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    xb[c] = a.xb[3 * p + c] + a.inv_dx * gx[c] - (A[c] * P[0] + A[3 + c] * P[1] + A[6 + c] * P[2]);
  // v, C
#pragma unroll
  for (int d = 0; d < 3; ++d) a.vb[3 * p + d] = m * P[d];
#pragma unroll
  for (int e = 0; e < 9; ++e) a.Cb[9 * p + e] = m * Ab[e];
  // stress: G = scale * Ab is the adjoint of PF
  float G[9], Fb[9], Rb[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    G[e] = scale * Ab[e];
    Fb[e] = a.Fb[9 * p + e];
  }
  float gm = 0.f;  // <G, M>: all entries (reference) or the trace (textbook)
#pragma unroll
  for (int e = 0; e < 9; ++e) gm += (a.textbook && (e % 4 != 0)) ? 0.f : G[e];
  float mub = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float fr = 0.f;  // ((F - R) F^T)_ij
      float gf = 0.f;  // (G F)_ij
      float gtd = 0.f; // (G^T (F - R))_ij
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fr += (F[3 * i + k] - R[3 * i + k]) * F[3 * j + k];
        gf += G[3 * i + k] * F[3 * k + j];
        gtd += G[3 * k + i] * (F[3 * k + j] - R[3 * k + j]);
      }
      mub += 2.f * G[3 * i + j] * fr;
      Fb[3 * i + j] += 2.f * mu * (gf + gtd);
      Rb[3 * i + j] = -2.f * mu * gf;
    }
  const float lamb = (J - 1.f) * J * gm;
  float cof[9];
  cofactor3(F, cof);
  const float jf = lam * (2.f * J - 1.f) * gm;
#pragma unroll
  for (int e = 0; e < 9; ++e) Fb[e] += jf * cof[e];
  if (mu != 0.f) polar_adjoint(F, R, Rb, Fb);
#pragma unroll
  for (int d = 0; d < 3; ++d) a.xb[3 * p + d] = xb[d];
#pragma unroll
  for (int e = 0; e < 9; ++e) a.Fb[9 * p + e] = Fb[e];
  if (a.mat_scalar) {
    atomicAdd(a.mub, mub * a.hardening);
    atomicAdd(a.lamb, lamb * a.hardening);
  } else {
    a.mub[p] += mub * a.hardening;
    a.lamb[p] += lamb * a.hardening;
  }
}

__global__ void adj_fill_kernel(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void adj_copy_kernel(float* dst, const float* src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

size_t adjoint_workspace_floats(int res) {
  const size_t r1 = (size_t)res + 1;
  return 11 * r1 * r1 * r1;
}

// Forward substep on the dense workspace: gp, gm = P2G; gv = grid update of gp.
void dense_forward_grids(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                         float* work, unsigned long long* oob) {
  const int64_t r1 = prm->res + 1, nn = r1 * r1 * r1;
  float* gp = work;
  float* gm = gp + 3 * nn;
  float* gv = gm + nn;
  hipLaunchKernelGGL(adj_fill_kernel, dim3(nblk(4 * nn, 256)), dim3(256), 0, st, gp, 4 * nn, 0.f);
  phase_p2g(st, prm, p, MPM_ELASTIC, hardening, nullptr, 0, gp, gm, oob);
  hipLaunchKernelGGL(adj_copy_kernel, dim3(nblk(3 * nn, 256)), dim3(256), 0, st, gv, gp, 3 * nn);
  phase_grid_op(st, prm, MPM_F32, gv, gm);
}

// Reverse substep from the forward grids `grids` (gp, gm, gv as laid out at
// the start of the workspace); the gradient grid lives at work + 7 nn.
static void adjoint_from_grids(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                               const float* grids, float* work, float* xb, float* vb, float* Fb, float* Cb,
                               float* mub, float* lamb, unsigned long long* oob) {
  const int64_t r1 = prm->res + 1, nn = r1 * r1 * r1;
  AdjArgs a{};
  a.n = p->n;
  a.res = prm->res;
  a.boundary = prm->boundary;
  a.bound = prm->bound;
  a.textbook = prm->quirks == MPM_QUIRKS_TEXTBOOK;
  a.mat_scalar = p->mat_scalar;
  a.inv_dx = (float)prm->inv_dx;
  a.dx = (float)prm->dx;
  a.dt = (float)prm->dt;
  a.vol = (float)prm->volume;
  a.gravity = (float)prm->gravity;
  a.hardening = (float)hardening;
  a.v_allowed = (float)prm->dx * 0.9f / (float)prm->dt;
  a.x = (const float*)p->x;
  a.v = (const float*)p->v;
  a.F = (const float*)p->F;
  a.C = (const float*)p->C;
  a.mass = (const float*)p->mass;
  a.mu0 = (const float*)p->mu0;
  a.lam0 = (const float*)p->lambda0;
  a.gp = grids;
  a.gm = grids + 3 * nn;
  a.gv = grids + 4 * nn;
  a.gbar = work + 7 * nn;
  a.xb = xb;
  a.vb = vb;
  a.Fb = Fb;
  a.Cb = Cb;
  a.mub = mub;
  a.lamb = lamb;
  a.oob = oob;
  hipLaunchKernelGGL(adj_fill_kernel, dim3(nblk(4 * nn, 256)), dim3(256), 0, st, a.gbar, 4 * nn, 0.f);
  if (a.n > 0) hipLaunchKernelGGL(adj_g2p_kernel, dim3(nblk(a.n, AG_THREADS)), dim3(AG_THREADS), 0, st, a);
  hipLaunchKernelGGL(adj_grid_kernel, dim3(nblk(nn, 256)), dim3(256), 0, st, a);
  if (a.n > 0) hipLaunchKernelGGL(adj_p2g_kernel, dim3(nblk(a.n, 128)), dim3(128), 0, st, a);
}

void adjoint_step(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                  float* work, float* xb, float* vb, float* Fb, float* Cb, float* mub, float* lamb,
                  unsigned long long* oob) {
  dense_forward_grids(st, prm, p, hardening, work, oob);
  adjoint_from_grids(st, prm, p, hardening, work, work, xb, vb, Fb, Cb, mub, lamb, oob);
}

void adjoint_step_grids(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                        const float* grids, float* work, float* xb, float* vb, float* Fb, float* Cb, float* mub,
                        float* lamb, unsigned long long* oob) {
  adjoint_from_grids(st, prm, p, hardening, grids, work, xb, vb, Fb, Cb, mub, lamb, oob);
}

}  // namespace mpm
