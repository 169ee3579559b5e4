// Phase kernels on caller-owned dense grids in the reference's layouts:
// grid_v (R+1)^d x d and grid_m (R+1)^d, row-major [i, j(, k)].  These back
// the drop-in phase functions three_d.{p2g,grid_op,g2p} and
// two_d.{p2g,grid_op,g2p} (FP32 or FP64, whichever the caller's arrays are).
// One thread per particle (P2G, G2P) or per node (grid update).
#include "mpm_math.h"
#include "mpm_phase.h"

namespace mpm {

template <typename T> struct PhaseArgs {
  int64_t n;
  int res, quirks, boundary, bound, mat_scalar, kind_override;
  T inv_dx, dx, dt, vol, gravity, hardening, theta_c, theta_s;
  T* x;
  T* v;
  T* F;
  T* C;
  T* Jp;
  const T* mass;
  const T* mu0;
  const T* lam0;
  const int32_t* material;
  T* grid_v;
  T* grid_m;
  unsigned long long* oob;
  int nmat;
  MatParams mats[MPM_PHASE_MAX_MATERIALS];
};

template <typename T>
__device__ __forceinline__ void material_of(const PhaseArgs<T>& a, int64_t p, int& kind, T& h,
                                            T& tc, T& ts) {
  if (a.kind_override >= 0) {
    kind = a.kind_override;
    h = a.hardening;
    tc = a.theta_c;
    ts = a.theta_s;
    return;
  }
  int id = a.material ? a.material[p] : 0;
  id = id < 0 ? 0 : (id >= a.nmat ? a.nmat - 1 : id);
  const MatParams m = a.mats[id];
  kind = m.kind;
  h = (T)m.hardening;
  tc = (T)m.theta_c;
  ts = (T)m.theta_s;
}

// base = trunc(x inv_dx - 0.5) (three_d/p2g.py:50) with the stencil inside
// [0, R] (utils.py:138-150): valid iff -1 < x inv_dx - 0.5 < R - 1.
template <typename T, int DIM>
__device__ __forceinline__ bool locate(const T* xp, T inv_dx, int res, int base[DIM], T fx[DIM]) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T s = xp[d] * inv_dx;
    const T t = s - T(0.5);
    ok = ok && (t > T(-1)) && (t < T(res - 1));
    base[d] = ok ? (int)t : 0;
    fx[d] = s - (T)base[d];
  }
  return ok;
}

template <typename T, int DIM>
__global__ void phase_p2g_kernel(PhaseArgs<T> a) {
  // lanes past the end or out of the grid stay in the wave (its segmented
  // atomics need every lane) with a zero contribution
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool inrange = p0 < a.n;
  const int64_t p = inrange ? p0 : a.n - 1;
  int base[DIM];
  T fx[DIM];
  const bool located = locate<T, DIM>(a.x + p * DIM, a.inv_dx, a.res, base, fx);
  if (inrange && !located && a.oob) atomicAdd(a.oob, 1ull);
  const bool valid = inrange && located;
  if (!located) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      base[d] = 0;
      fx[d] = T(1);
    }
  }
  T w[DIM][3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) bspline(fx[d], w[d]);
  int kind;
  T h, tc, ts;
  material_of(a, p, kind, h, tc, ts);
  const int64_t mi = a.mat_scalar ? 0 : p;
  const T m = valid ? a.mass[mi] : T(0);
  T mu, lam;
  lame<T>(kind, h, a.mu0[mi], a.lam0[mi], a.Jp[p], mu, lam);
  const bool textbook = a.quirks == MPM_QUIRKS_TEXTBOOK;
  const int64_t r1 = a.res + 1;
  if constexpr (DIM == 3) {
    const WaveSeg sg = wave_segments(valid, (int)((base[0] * r1 + base[1]) * r1 + base[2]));
    T A[9];
    affine3<T>(a.F + 9 * p, a.C + 9 * p, m, mu, lam, kind, a.inv_dx, a.dt, a.vol, textbook, A);
    if (!valid) {
#pragma unroll
      for (int e = 0; e < 9; ++e) A[e] = T(0);
    }
    const T mv[3] = {a.v[3 * p] * m, a.v[3 * p + 1] * m, a.v[3 * p + 2] * m};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const T dp[3] = {(T(i) - fx[0]) * a.dx, (T(j) - fx[1]) * a.dx, (T(k) - fx[2]) * a.dx};
          const T weight = w[0][i] * w[1][j] * w[2][k];
          const int64_t node = ((int64_t)(base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const T ad = A[3 * d] * dp[0] + A[3 * d + 1] * dp[1] + A[3 * d + 2] * dp[2];
            const T r = seg_reduce(weight * (mv[d] + ad), sg);
            if (sg.tail) atomicAdd(a.grid_v + 3 * node + d, r);
          }
          const T rm = seg_reduce(weight * m, sg);
          if (sg.tail) atomicAdd(a.grid_m + node, rm);
        }
  } else {
    const WaveSeg sg = wave_segments(valid, (int)(base[0] * r1 + base[1]));
    T A[4];
    affine2<T>(a.F + 4 * p, a.C + 4 * p, m, mu, lam, kind, a.inv_dx, a.dt, a.vol, textbook, A);
    if (!valid) {
#pragma unroll
      for (int e = 0; e < 4; ++e) A[e] = T(0);
    }
    const T mv[2] = {a.v[2 * p] * m, a.v[2 * p + 1] * m};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T dp[2] = {(T(i) - fx[0]) * a.dx, (T(j) - fx[1]) * a.dx};
        const T weight = w[0][i] * w[1][j];
        const int64_t node = (int64_t)(base[0] + i) * r1 + (base[1] + j);
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const T ad = A[2 * d] * dp[0] + A[2 * d + 1] * dp[1];
          const T r = seg_reduce(weight * (mv[d] + ad), sg);
          if (sg.tail) atomicAdd(a.grid_v + 2 * node + d, r);
        }
        const T rm = seg_reduce(weight * m, sg);
        if (sg.tail) atomicAdd(a.grid_m + node, rm);
      }
  }
}

// three_d/grid_op.py:25-47 and two_d/grid_op.py:13-24, one thread per node.
template <typename T, int DIM>
__global__ void phase_grid_kernel(PhaseArgs<T> a) {
  const int64_t r1 = a.res + 1;
  int64_t total = r1 * r1;
  if (DIM == 3) total *= r1;
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= total) return;
  T* g = a.grid_v + DIM * node;
  const T m = a.grid_m[node];
  if constexpr (DIM == 3) {
    const int64_t i = node / (r1 * r1), j = (node / r1) % r1, k = node % r1;
    T v[3] = {g[0], g[1], g[2]};
    if (m > T(0)) {
      const T lim = a.dx * T(0.9) / a.dt;
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] /= m;
      v[1] += a.dt * a.gravity;
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = tmin(tmax(v[d], -lim), lim);
    }
    const int64_t I[3] = {i, j, k};
    if (a.boundary == MPM_BC_REFERENCE) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (I[d] < 1 || I[d] >= a.res - 1) v[d] = T(0);
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (I[d] < a.bound || I[d] > a.res - a.bound) v[0] = v[1] = v[2] = T(0);
    }
    g[0] = v[0];
    g[1] = v[1];
    g[2] = v[2];
  } else {
    if (!(m > T(0))) return;
    const int64_t i = node / r1, j = node % r1;
    T v[2] = {g[0] / m, g[1] / m};
    v[1] += a.dt * a.gravity;
    if (a.boundary == MPM_BC_REFERENCE) {
      const T xs = T(i) / T(a.res), ys = T(j) / T(a.res);
      if (xs < T(0.05) || xs > T(1) - T(0.05) || ys > T(1) - T(0.05)) v[0] = v[1] = T(0);
      if (ys < T(0.05)) v[1] = tmax(T(0), v[1]);
    } else if (i < a.bound || i > a.res - a.bound || j < a.bound || j > a.res - a.bound) {
      v[0] = v[1] = T(0);
    }
    g[0] = v[0];
    g[1] = v[1];
  }
}

// three_d/grid_op.py:50-67 (check_collision_points), one thread per node of
// the dense (R+1)^3 grid; every plane is tested in turn.
struct PlaneSet {
  PlaneDev p[MAX_PLANES];
  int n;
};

template <typename T>
__global__ void phase_planes_kernel(T* __restrict__ gv, int res, PlaneSet ps) {
  const int64_t r1 = res + 1;
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= r1 * r1 * r1) return;
  const int i = (int)(node / (r1 * r1)), j = (int)((node / r1) % r1), k = (int)(node % r1);
  bool hit = false;
  for (int q = 0; q < ps.n; ++q) hit = hit || plane_blocks(ps.p[q], i, j, k);
  if (hit) {
    gv[3 * node] = T(0);
    gv[3 * node + 1] = T(0);
    gv[3 * node + 2] = T(0);
  }
}

template <typename T, int DIM>
__global__ void phase_g2p_kernel(PhaseArgs<T> a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  int base[DIM];
  T fx[DIM];
  if (!locate<T, DIM>(a.x + p * DIM, a.inv_dx, a.res, base, fx)) {
    if (a.oob) atomicAdd(a.oob, 1ull);
    return;
  }
  T w[DIM][3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) bspline(fx[d], w[d]);
  int kind;
  T h, tc, ts;
  material_of(a, p, kind, h, tc, ts);
  const int64_t r1 = a.res + 1;
  const T c4 = T(4) * a.inv_dx;
  if constexpr (DIM == 3) {
    T v[3] = {0, 0, 0}, C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int64_t node = ((int64_t)(base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
          const T weight = w[0][i] * w[1][j] * w[2][k];
          const T dp[3] = {T(i) - fx[0], T(j) - fx[1], T(k) - fx[2]};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const T wg = weight * a.grid_v[3 * node + d];
            v[d] += wg;
#pragma unroll
            for (int e = 0; e < 3; ++e) C[3 * d + e] += c4 * (wg * dp[e]);
          }
        }
    T* xp = a.x + 3 * p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a.v[3 * p + d] = v[d];
      xp[d] += a.dt * v[d];
    }
    T M[9], Fn[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      M[i] = a.dt * C[i] + ((i % 4 == 0) ? T(1) : T(0));
      a.C[9 * p + i] = C[i];
    }
    matmul3(M, a.F + 9 * p, Fn);
    if (kind == MPM_SNOW) {
      T U[9], s[3], V[9];
      svd3_checked<T>(Fn, U, s, V);
      const T oldJ = det3(Fn);
      const T lo = T(1) - tc, hi = T(1) + ts;
      // rebuild only when a singular value is actually clamped (otherwise
      // U diag(s) V^T is F itself)
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
      const T det = det3(Fn) + T(1e-10);
      a.Jp[p] = tmin(tmax(a.Jp[p] * oldJ / det, T(0.6)), T(20));
    } else if (kind == MPM_WATER) {
      const T c = tcbrt(det3(Fn));
#pragma unroll
      for (int i = 0; i < 9; ++i) Fn[i] = (i % 4 == 0) ? c : T(0);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) a.F[9 * p + i] = Fn[i];
  } else {
    T v[2] = {0, 0}, C[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int64_t node = (int64_t)(base[0] + i) * r1 + (base[1] + j);
        const T weight = w[0][i] * w[1][j];
        const T dp[2] = {T(i) - fx[0], T(j) - fx[1]};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const T wg = weight * a.grid_v[2 * node + d];
          v[d] += wg;
#pragma unroll
          for (int e = 0; e < 2; ++e) C[2 * d + e] += c4 * (wg * dp[e]);
        }
      }
    T* xp = a.x + 2 * p;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      a.v[2 * p + d] = v[d];
      xp[d] += a.dt * v[d];
    }
    const T* Fp = a.F + 4 * p;
    const T M[4] = {T(1) + a.dt * C[0], a.dt * C[1], a.dt * C[2], T(1) + a.dt * C[3]};
    T Fn[4] = {M[0] * Fp[0] + M[1] * Fp[2], M[0] * Fp[1] + M[1] * Fp[3], M[2] * Fp[0] + M[3] * Fp[2],
               M[2] * Fp[1] + M[3] * Fp[3]};
#pragma unroll
    for (int i = 0; i < 4; ++i) a.C[4 * p + i] = C[i];
    // two_d/g2p.py:321-330: SVD rebuild and Jp update for every particle.
    // U diag(s) V^T equals F exactly unless a singular value is clamped, so
    // the rebuild (whose rounding would seed spurious stress) only runs then.
    const T oldJ = Fn[0] * Fn[3] - Fn[1] * Fn[2];
    if (kind == MPM_SNOW) {
      T U[4], s[2], V[4];
      svd2<T>(Fn, U, s, V);
      const T lo = T(1) - tc, hi = T(1) + ts;
      if (s[0] < lo || s[0] > hi || s[1] < lo || s[1] > hi) {
        s[0] = tmin(tmax(s[0], lo), hi);
        s[1] = tmin(tmax(s[1], lo), hi);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            Fn[2 * i + j] = U[2 * i] * s[0] * V[2 * j] + U[2 * i + 1] * s[1] * V[2 * j + 1];
      }
    }
    const T det = Fn[0] * Fn[3] - Fn[1] * Fn[2] + T(1e-10);
    a.Jp[p] = tmin(tmax(a.Jp[p] * oldJ / det, T(0.6)), T(20));
#pragma unroll
    for (int i = 0; i < 4; ++i) a.F[4 * p + i] = Fn[i];
  }
}

// ---------------------------------------------------------------- host ----
template <typename T>
static PhaseArgs<T> make_args(const mpm_params* prm, const mpm_particles* p, int kind_override,
                              double hardening, double theta_c, double theta_s,
                              const mpm_material* mats, int nmat, void* gv, void* gm,
                              unsigned long long* oob) {
  PhaseArgs<T> a{};
  a.n = p ? p->n : 0;
  a.res = prm->res;
  a.quirks = prm->quirks;
  a.boundary = prm->boundary;
  a.bound = prm->bound;
  a.inv_dx = (T)prm->inv_dx;
  a.dx = (T)prm->dx;
  a.dt = (T)prm->dt;
  a.vol = (T)prm->volume;
  a.gravity = (T)prm->gravity;
  a.kind_override = kind_override;
  a.hardening = (T)hardening;
  a.theta_c = (T)theta_c;
  a.theta_s = (T)theta_s;
  if (p) {
    a.mat_scalar = p->mat_scalar;
    a.x = (T*)p->x;
    a.v = (T*)p->v;
    a.F = (T*)p->F;
    a.C = (T*)p->C;
    a.Jp = (T*)p->Jp;
    a.mass = (const T*)p->mass;
    a.mu0 = (const T*)p->mu0;
    a.lam0 = (const T*)p->lambda0;
    a.material = p->material;
  }
  a.grid_v = (T*)gv;
  a.grid_m = (T*)gm;
  a.oob = oob;
  a.nmat = nmat > 0 ? nmat : 1;
  for (int i = 0; i < MPM_PHASE_MAX_MATERIALS; ++i) {
    if (mats && i < nmat) {
      a.mats[i].kind = mats[i].kind;
      a.mats[i].hardening = mats[i].hardening;
      a.mats[i].theta_c = mats[i].theta_c;
      a.mats[i].theta_s = mats[i].theta_s;
    } else {
      a.mats[i].kind = MPM_ELASTIC;
      a.mats[i].hardening = 1.0f;
      a.mats[i].theta_c = 2.5e-2f;
      a.mats[i].theta_s = 7.5e-3f;
    }
  }
  return a;
}

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void phase_p2g(hipStream_t st, const mpm_params* prm, const mpm_particles* p, int kind_override,
               double hardening, const mpm_material* mats, int nmat, void* gv, void* gm,
               unsigned long long* oob) {
  if (p->n <= 0) return;
  const unsigned nb = blocks_for(p->n, 256);
  if (p->dtype == MPM_F64) {
    auto a = make_args<double>(prm, p, kind_override, hardening, 0, 0, mats, nmat, gv, gm, oob);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_p2g_kernel<double, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_p2g_kernel<double, 2>), dim3(nb), dim3(256), 0, st, a);
  } else {
    auto a = make_args<float>(prm, p, kind_override, hardening, 0, 0, mats, nmat, gv, gm, oob);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_p2g_kernel<float, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_p2g_kernel<float, 2>), dim3(nb), dim3(256), 0, st, a);
  }
}

void phase_grid_op(hipStream_t st, const mpm_params* prm, int dtype, void* gv, const void* gm) {
  const int64_t r1 = prm->res + 1;
  const int64_t total = prm->dim == 3 ? r1 * r1 * r1 : r1 * r1;
  const unsigned nb = blocks_for(total, 256);
  if (dtype == MPM_F64) {
    auto a = make_args<double>(prm, nullptr, -1, 1, 0, 0, nullptr, 0, gv, (void*)gm, nullptr);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_grid_kernel<double, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_grid_kernel<double, 2>), dim3(nb), dim3(256), 0, st, a);
  } else {
    auto a = make_args<float>(prm, nullptr, -1, 1, 0, 0, nullptr, 0, gv, (void*)gm, nullptr);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_grid_kernel<float, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_grid_kernel<float, 2>), dim3(nb), dim3(256), 0, st, a);
  }
}

void phase_planes(hipStream_t st, int dtype, int res, const PlaneDev* planes, int n, void* gv) {
  PlaneSet ps{};
  for (int q = 0; q < n; ++q) ps.p[q] = planes[q];
  ps.n = n;
  const int64_t r1 = res + 1;
  const unsigned nb = blocks_for(r1 * r1 * r1, 256);
  if (dtype == MPM_F64) hipLaunchKernelGGL(phase_planes_kernel<double>, dim3(nb), dim3(256), 0, st, (double*)gv, res, ps);
  else hipLaunchKernelGGL(phase_planes_kernel<float>, dim3(nb), dim3(256), 0, st, (float*)gv, res, ps);
}

void phase_g2p(hipStream_t st, const mpm_params* prm, mpm_particles* p, int kind_override,
               double theta_c, double theta_s, const mpm_material* mats, int nmat, const void* gv,
               unsigned long long* oob) {
  if (p->n <= 0) return;
  const unsigned nb = blocks_for(p->n, 256);
  if (p->dtype == MPM_F64) {
    auto a = make_args<double>(prm, p, kind_override, 1, theta_c, theta_s, mats, nmat, (void*)gv,
                               nullptr, oob);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_g2p_kernel<double, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_g2p_kernel<double, 2>), dim3(nb), dim3(256), 0, st, a);
  } else {
    auto a = make_args<float>(prm, p, kind_override, 1, theta_c, theta_s, mats, nmat, (void*)gv,
                              nullptr, oob);
    if (prm->dim == 3) hipLaunchKernelGGL((phase_g2p_kernel<float, 3>), dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((phase_g2p_kernel<float, 2>), dim3(nb), dim3(256), 0, st, a);
  }
}

}  // namespace mpm
