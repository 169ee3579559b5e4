// Device math shared by every MLS-MPM kernel: quadratic B-spline weights,
// hardening laws, 3x3 / 2x2 polar decompositions and SVDs, and the
// fixed-corotated stress fused with the APIC affine momentum.
//
// Reference semantics (paths relative to the FEMFlow source tree):
//   weights            femflow/solvers/mpm/three_d/p2g.py:55
//   hardening          femflow/solvers/mpm/utils.py:7-49
//   stress + m C       femflow/solvers/mpm/utils.py:52-135
//   polar (R = U Vh)   femflow/numerics/linear_algebra.py:96-135
//
// Everything is templated on the scalar type T (float or double).  The FP32
// 3x3 polar factor comes from a scaled Newton iteration; when F is close to
// singular (or the iteration does not settle) the same factor is recomputed
// in FP64 through the Jacobi SVD, which is the one place FP32 work falls back
// to FP64 in the stress path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm355.h"

#define MPM_DEV __device__ __forceinline__

namespace mpm {

template <typename T> MPM_DEV T tsqrt(T x);
template <> MPM_DEV float tsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <> MPM_DEV double tsqrt<double>(double x) { return __builtin_sqrt(x); }
template <typename T> MPM_DEV T tabs(T x) { return x < T(0) ? -x : x; }
// Reciprocal and reciprocal square root: FP32 uses the hardware
// approximations (v_rcp_f32, v_rsq_f32; ~1 ulp), enough for the rotation
// and scaling factors below, whose exactness does not move the converged
// results; FP64 stays exact.
template <typename T> MPM_DEV T trcp(T x);
template <> MPM_DEV float trcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> MPM_DEV double trcp<double>(double x) { return 1.0 / x; }
template <typename T> MPM_DEV T trsq(T x);
template <> MPM_DEV float trsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> MPM_DEV double trsq<double>(double x) { return 1.0 / __builtin_sqrt(x); }
template <typename T> MPM_DEV T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> MPM_DEV T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> MPM_DEV T texp(T x);
template <> MPM_DEV float texp<float>(float x) { return __expf(x); }
template <> MPM_DEV double texp<double>(double x) { return exp(x); }
template <typename T> MPM_DEV T tcbrt(T x);
template <> MPM_DEV float tcbrt<float>(float x) { return cbrtf(x); }
template <> MPM_DEV double tcbrt<double>(double x) { return cbrt(x); }
template <typename T> MPM_DEV T trcbrt(T x);
template <> MPM_DEV float trcbrt<float>(float x) { return rcbrtf(x); }
template <> MPM_DEV double trcbrt<double>(double x) { return rcbrt(x); }
// Finite test on the exponent bits (independent of how the compiler treats
// NaN in floating-point comparisons).
MPM_DEV bool tfinite(float x) { return (__float_as_uint(x) & 0x7f800000u) != 0x7f800000u; }
MPM_DEV bool tfinite(double x) {
  return ((unsigned long long)__double_as_longlong(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}

// Quadratic B-spline weights in the reference's per-index polynomial form
// (three_d/p2g.py:55).  They equal N(i - fx) for fx in [0.5, 1.5); for
// fx < 0.5 (base truncated toward zero next to the low wall) w1 goes negative
// exactly as the reference's does.
template <typename T> MPM_DEV void bspline(T fx, T w[3]) {
  const T a = T(1.5) - fx, b = fx - T(1), c = fx - T(0.5);
  w[0] = T(0.5) * a * a;
  w[1] = T(0.75) - b * b;
  w[2] = T(0.5) * c * c;
}

// -------------------------------------------------------------------------
// 3x3 helpers (row-major a[3*i+j])
// -------------------------------------------------------------------------
template <typename T> MPM_DEV T det3(const T* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

template <typename T> MPM_DEV void cofactor3(const T* a, T* c) {
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[5] * a[6] - a[3] * a[8];
  c[2] = a[3] * a[7] - a[4] * a[6];
  c[3] = a[2] * a[7] - a[1] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[1] * a[6] - a[0] * a[7];
  c[6] = a[1] * a[5] - a[2] * a[4];
  c[7] = a[2] * a[3] - a[0] * a[5];
  c[8] = a[0] * a[4] - a[1] * a[3];
}

template <typename T> MPM_DEV void matmul3(const T* a, const T* b, T* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// Symmetric 3x3 eigen-solve by cyclic Jacobi with exact Givens rotations.
// A is overwritten (diagonal = eigenvalues), V accumulates eigenvectors as
// columns.  `sweeps` full sweeps are run unless the off-diagonal vanishes.
template <typename T> MPM_DEV void jacobi_eig3(T A[9], T V[9], int sweeps) {
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? T(1) : T(0);
  for (int s = 0; s < sweeps; ++s) {
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      const T apq = A[3 * p + q];
      if (apq == T(0)) continue;
      const T theta = (A[3 * q + q] - A[3 * p + p]) * trcp(T(2) * apq);
      const T t = (theta >= T(0) ? T(1) : T(-1)) * trcp(tabs(theta) + tsqrt(theta * theta + T(1)));
      const T c = trsq(t * t + T(1));
      const T sn = t * c;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T akp = A[3 * k + p], akq = A[3 * k + q];
        A[3 * k + p] = c * akp - sn * akq;
        A[3 * k + q] = sn * akp + c * akq;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T apk = A[3 * p + k], aqk = A[3 * q + k];
        A[3 * p + k] = c * apk - sn * aqk;
        A[3 * q + k] = sn * apk + c * aqk;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T vkp = V[3 * k + p], vkq = V[3 * k + q];
        V[3 * k + p] = c * vkp - sn * vkq;
        V[3 * k + q] = sn * vkp + c * vkq;
      }
    }
    if constexpr (sizeof(T) == 4) {  // FP32: stop once the off-diagonal is at rounding level
      const T off = tabs(A[1]) + tabs(A[2]) + tabs(A[5]);
      if (off <= T(1e-8) * (tabs(A[0]) + tabs(A[4]) + tabs(A[8]))) break;
    }
  }
}

template <typename T> MPM_DEV void swap_cols(T* M, int a, int b) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    T t = M[3 * r + a];
    M[3 * r + a] = M[3 * r + b];
    M[3 * r + b] = t;
  }
}

// F = U diag(s) V^T with s sorted descending and s >= 0 (LAPACK's sign
// convention; U V^T is then the orthogonal polar factor, a reflection when
// det F < 0, as linear_algebra.py:131-132 returns it).
template <typename T> MPM_DEV void svd3(const T* F, T U[9], T s[3], T V[9], int sweeps) {
  T A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      A[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
  jacobi_eig3(A, V, sweeps);
  T ev[3] = {A[0], A[4], A[8]};
  // sort eigenvalues descending (3-element network), permuting V's columns
  if (ev[0] < ev[1]) { T t = ev[0]; ev[0] = ev[1]; ev[1] = t; swap_cols(V, 0, 1); }
  if (ev[0] < ev[2]) { T t = ev[0]; ev[0] = ev[2]; ev[2] = t; swap_cols(V, 0, 2); }
  if (ev[1] < ev[2]) { T t = ev[1]; ev[1] = ev[2]; ev[2] = t; swap_cols(V, 1, 2); }
  T B[9];
  matmul3(F, V, B);
  // Gram-Schmidt on the columns of B = U diag(s)
  T n0 = tsqrt(B[0] * B[0] + B[3] * B[3] + B[6] * B[6]);
  T i0 = n0 > T(0) ? T(1) / n0 : T(0);
  U[0] = B[0] * i0; U[3] = B[3] * i0; U[6] = B[6] * i0;
  if (n0 == T(0)) { U[0] = T(1); }
  T d01 = U[0] * B[1] + U[3] * B[4] + U[6] * B[7];
  T c1[3] = {B[1] - d01 * U[0], B[4] - d01 * U[3], B[7] - d01 * U[6]};
  T n1 = tsqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
  if (n1 > T(0)) {
    T i1 = T(1) / n1;
    U[1] = c1[0] * i1; U[4] = c1[1] * i1; U[7] = c1[2] * i1;
  } else {  // any unit vector orthogonal to u0
    T ax = tabs(U[0]) < T(0.5) ? T(1) : T(0), ay = ax == T(0) ? T(1) : T(0);
    T c[3] = {U[3] * T(0) - U[6] * ay, U[6] * ax - U[0] * T(0), U[0] * ay - U[3] * ax};
    T nc = T(1) / tsqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    U[1] = c[0] * nc; U[4] = c[1] * nc; U[7] = c[2] * nc;
  }
  // u2 = u0 x u1, sign chosen so that s2 = u2 . b2 >= 0
  T u2[3] = {U[3] * U[7] - U[6] * U[4], U[6] * U[1] - U[0] * U[7], U[0] * U[4] - U[3] * U[1]};
  T d2 = u2[0] * B[2] + u2[1] * B[5] + u2[2] * B[8];
  T sg = d2 < T(0) ? T(-1) : T(1);
  U[2] = sg * u2[0]; U[5] = sg * u2[1]; U[8] = sg * u2[2];
  s[0] = n0;
  s[1] = U[1] * B[1] + U[4] * B[4] + U[7] * B[7];
  s[2] = tabs(d2);
}

// Orthogonal polar factor R of F by a scaled Newton iteration
// X <- (g X + X^{-T} / g) / 2, g = |det X|^{-1/3}.  Returns false when F is
// near-singular or the iteration has not settled after `iters` steps.
template <typename T> MPM_DEV bool polar_newton(const T* F, T R[9], int iters, T tol) {
  T X[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) X[i] = F[i];
  T fn = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) fn += F[i] * F[i];
  const T det0 = det3(F);
  // conditioning guard: |det F| small against (||F||_F / sqrt 3)^3
  const T scale3 = fn * tsqrt(fn) * T(0.19245008972987526);  // (1/sqrt(3))^3
  if (!(tabs(det0) > T(1e-3) * scale3)) return false;
  for (int it = 0; it < iters; ++it) {
    T c[9];
    cofactor3(X, c);
    const T det = X[0] * c[0] + X[1] * c[1] + X[2] * c[2];
    T g;
    if constexpr (sizeof(T) == 4) {
      // any positive scale converges to the same factor; exp2(-log2|det|/3)
      // on the hardware transcendental units is ample
      g = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(tabs(det)) * (-1.f / 3.f));
    } else {
      g = trcbrt(tabs(det));
    }
    const T a = T(0.5) * g, b = T(0.5) * trcp(g * det);
    T delta = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const T nx = a * X[i] + b * c[i];
      delta = tmax(delta, tabs(nx - X[i]));
      X[i] = nx;
    }
    if (delta <= tol) {
#pragma unroll
      for (int i = 0; i < 9; ++i) R[i] = X[i];
      return true;
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = X[i];
  return false;
}

// det(F) - 1 from D = F - I (exact in floating point for F near I):
// det(I + D) = 1 + tr D + (sum of the principal 2x2 minors of D) + det D.
// Near a rotation-free F this keeps J - 1 to the relative precision of D,
// where det3(F) - 1 would cancel to the rounding of an O(1) determinant.
template <typename T> MPM_DEV T det3_minus_one(const T* F) {
  const T d0 = F[0] - T(1), d4 = F[4] - T(1), d8 = F[8] - T(1);
  const T tr = d0 + d4 + d8;
  const T m2 = (d0 * d4 - F[1] * F[3]) + (d0 * d8 - F[2] * F[6]) + (d4 * d8 - F[5] * F[7]);
  const T dd = d0 * (d4 * d8 - F[5] * F[7]) - F[1] * (F[3] * d8 - F[5] * F[6]) + F[2] * (F[3] * F[7] - d4 * F[6]);
  return tr + m2 + dd;
}

// (F - R) F^T of the fixed-corotated stress without the polar factor R.
// With R = U V^T from the SVD F = U S V^T (linear_algebra.py:119-135),
// R F^T = U S U^T = sqrt(B), B = F F^T, for either sign of det F, so
// (F - R) F^T = B - sqrt(B).  With E = B - I = D + D^T + D D^T (D = F - I),
//   B - sqrt(B) = E/2 + E^2/8 - E^3/16 + 5 E^4/128 - 7 E^5/256 + ...
// (the binomial series of sqrt(I + E), which converges for ||E|| < 1).  Five
// terms leave a relative error below 4e-8 while ||E||_F <= 0.06 (FP64
// check against the SVD form: 1e-8 at 0.05, 4e-8 at 0.06); beyond that the
// caller takes the polar iteration.  Out: the symmetric result as
// (00, 01, 02, 11, 12, 22).  All products are of commuting symmetric
// matrices (polynomials in E), so each stays symmetric.
template <typename T> MPM_DEV bool stretch_series(const T* F, T out[6]) {
  const T D[9] = {F[0] - T(1), F[1], F[2], F[3], F[4] - T(1), F[5], F[6], F[7], F[8] - T(1)};
  T E[6];
  E[0] = T(2) * D[0] + (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
  E[1] = (D[1] + D[3]) + (D[0] * D[3] + D[1] * D[4] + D[2] * D[5]);
  E[2] = (D[2] + D[6]) + (D[0] * D[6] + D[1] * D[7] + D[2] * D[8]);
  E[3] = T(2) * D[4] + (D[3] * D[3] + D[4] * D[4] + D[5] * D[5]);
  E[4] = (D[5] + D[7]) + (D[3] * D[6] + D[4] * D[7] + D[5] * D[8]);
  E[5] = T(2) * D[8] + (D[6] * D[6] + D[7] * D[7] + D[8] * D[8]);
  const T fro2 = E[0] * E[0] + E[3] * E[3] + E[5] * E[5] + T(2) * (E[1] * E[1] + E[2] * E[2] + E[4] * E[4]);
  if (!(fro2 <= T(0.0036))) return false;  // also NaN
  // Horner: S <- E S + c I from S = c5 I, then out = E S
  const T c[5] = {T(0.5), T(0.125), T(-0.0625), T(0.0390625), T(-0.02734375)};
  T S[6] = {c[4], T(0), T(0), c[4], T(0), c[4]};
#pragma unroll
  for (int k = 3; k >= -1; --k) {
    T P[6];
    P[0] = E[0] * S[0] + E[1] * S[1] + E[2] * S[2];
    P[1] = E[0] * S[1] + E[1] * S[3] + E[2] * S[4];
    P[2] = E[0] * S[2] + E[1] * S[4] + E[2] * S[5];
    P[3] = E[1] * S[1] + E[3] * S[3] + E[4] * S[4];
    P[4] = E[1] * S[2] + E[3] * S[4] + E[4] * S[5];
    P[5] = E[2] * S[2] + E[4] * S[4] + E[5] * S[5];
    if (k >= 0) {
      P[0] += c[k];
      P[3] += c[k];
      P[5] += c[k];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) S[i] = P[i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = S[i];
  return true;
}

// R = U V^T through the FP64 Jacobi SVD: the fallback for any F the FP32
// iteration declines.
MPM_DEV void polar_svd_f64(const double* F, double R[9]) {
  double U[9], s[3], V[9];
  svd3<double>(F, U, s, V, 10);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = U[3 * i] * V[3 * j] + U[3 * i + 1] * V[3 * j + 1] + U[3 * i + 2] * V[3 * j + 2];
}

// 3x3 matrices passed by value to the out-of-line FP64 fallbacks, which keeps
// the rarely taken FP64 code (and its registers) out of the hot kernels.
struct Mat3f {
  float a[9];
};
struct Svd3f {
  float U[9], s[3], V[9];
};

__device__ __noinline__ Mat3f polar_fallback_f64(Mat3f F) {
  double Fd[9], Rd[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fd[i] = F.a[i];
  polar_svd_f64(Fd, Rd);
  Mat3f R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.a[i] = (float)Rd[i];
  return R;
}

__device__ __noinline__ Svd3f svd_fallback_f64(Mat3f F) {
  double Fd[9], Ud[9], sd[3], Vd[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fd[i] = F.a[i];
  svd3<double>(Fd, Ud, sd, Vd, 10);
  Svd3f r;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    r.U[i] = (float)Ud[i];
    r.V[i] = (float)Vd[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) r.s[i] = (float)sd[i];
  return r;
}

// Polar factor in precision T with FP64 fallback.  Returns 1 when the
// fallback ran (counted by callers for the conditioning report).  Newton
// converges quadratically: a step change of 1e-5 (FP32) / 1e-9 (FP64) leaves
// an error near 1e-10 / 1e-18, below the rounding floor of each type, while a
// tolerance at the rounding floor itself would never be met reliably.
template <typename T> MPM_DEV int polar3(const T* F, T R[9]);
template <> MPM_DEV int polar3<float>(const float* F, float R[9]) {
  if (polar_newton<float>(F, R, 12, 1e-5f)) return 0;
  Mat3f Fm;
#pragma unroll
  for (int i = 0; i < 9; ++i) Fm.a[i] = F[i];
  const Mat3f Rm = polar_fallback_f64(Fm);
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i] = Rm.a[i];
  return 1;
}
template <> MPM_DEV int polar3<double>(const double* F, double R[9]) {
  if (polar_newton<double>(F, R, 40, 1e-9)) return 0;
  polar_svd_f64(F, R);
  return 1;
}

// SVD in precision T; FP32 falls back to FP64 when s_min < 1e-2 s_max (the
// F^T F route loses the small singular value in FP32 there).
template <typename T> MPM_DEV int svd3_checked(const T* F, T U[9], T s[3], T V[9]);
template <> MPM_DEV int svd3_checked<float>(const float* F, float U[9], float s[3], float V[9]) {
  svd3<float>(F, U, s, V, 5);
  if (s[2] >= 1e-2f * s[0] && tfinite(s[0] + s[1] + s[2])) return 0;
  Mat3f Fm;
#pragma unroll
  for (int i = 0; i < 9; ++i) Fm.a[i] = F[i];
  const Svd3f r = svd_fallback_f64(Fm);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    U[i] = r.U[i];
    V[i] = r.V[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) s[i] = r.s[i];
  return 1;
}
template <> MPM_DEV int svd3_checked<double>(const double* F, double U[9], double s[3], double V[9]) {
  svd3<double>(F, U, s, V, 10);
  return 0;
}

// -------------------------------------------------------------------------
// Materials (utils.py:7-49 plus the two build-only kinds)
// -------------------------------------------------------------------------
struct MatParams {
  int kind;
  float hardening, theta_c, theta_s;
};

template <typename T>
MPM_DEV void lame(int kind, T h, T mu0, T lam0, T Jp, T& mu, T& lam) {
  if (kind == MPM_WATER) {
    mu = T(0);
    lam = lam0;
    return;
  }
  const T e = kind == MPM_SNOW ? texp(h * (T(1) - Jp)) : h;
  mu = mu0 * e;
  lam = lam0 * e;
}

// A = -(dt vol) 4 inv_dx^2 PF + m C, PF = 2 mu (F - R) F^T + lam (J-1) J * S
// where S is all-ones (reference quirk, utils.py:129) or I (textbook, water).
// Returns 1 if the polar factor took the FP64 fallback.
template <typename T>
MPM_DEV int affine3(const T* F, const T* C, T mass, T mu, T lam, int kind, T inv_dx, T dt, T vol,
                    bool textbook, T A[9]) {
  T R[9];
  int fb = 0;
  const T J = det3(F);
  const T scale = -(dt * vol) * (T(4) * inv_dx * inv_dx);
  const T vt = lam * (J - T(1)) * J;
  const bool diag_only = textbook || kind == MPM_WATER;
  if (mu != T(0)) {
    fb = polar3<T>(F, R);
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = F[i];
  }
  const T m2 = T(2) * mu;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T d0 = m2 * (F[3 * i] - R[3 * i]), d1 = m2 * (F[3 * i + 1] - R[3 * i + 1]),
            d2 = m2 * (F[3 * i + 2] - R[3 * i + 2]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T pf = d0 * F[3 * j] + d1 * F[3 * j + 1] + d2 * F[3 * j + 2];
      pf += (diag_only && i != j) ? T(0) : vt;
      A[3 * i + j] = scale * pf + mass * C[3 * i + j];
    }
  }
  return fb;
}

// 2D: closed-form rotation with the reference's eps (linear_algebra.py:96-116)
template <typename T>
MPM_DEV void affine2(const T* F, const T* C, T mass, T mu, T lam, int kind, T inv_dx, T dt, T vol,
                     bool textbook, T A[4]) {
  const T J = F[0] * F[3] - F[1] * F[2];
  const T x = F[0] + F[3], y = F[2] - F[1];
  const T denom = tsqrt(x * x + y * y) + (textbook ? T(0) : T(1e-10));
  const T sc = T(1) / denom;
  const T c = x * sc, s = y * sc;
  const T r[4] = {c, -s, s, c};
  const T scale = -(dt * vol) * (T(4) * inv_dx * inv_dx);
  const T vt = lam * (J - T(1)) * J;
  const bool diag_only = textbook || kind == MPM_WATER;
  const T m2 = T(2) * mu;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      T pf = m2 * (F[2 * i] - r[2 * i]) * F[2 * j] + m2 * (F[2 * i + 1] - r[2 * i + 1]) * F[2 * j + 1];
      pf += (diag_only && i != j) ? T(0) : vt;
      A[2 * i + j] = scale * pf + mass * C[2 * i + j];
    }
}

// 2x2 SVD F = U diag(s) V^T, s >= 0 (closed form through two half-angles).
template <typename T> MPM_DEV void svd2(const T* F, T U[4], T s[2], T V[4]) {
  const T E = T(0.5) * (F[0] + F[3]), Fq = T(0.5) * (F[0] - F[3]);
  const T G = T(0.5) * (F[2] + F[1]), H = T(0.5) * (F[2] - F[1]);
  const T Q = tsqrt(E * E + H * H), Rr = tsqrt(Fq * Fq + G * G);
  T s1 = Q + Rr, s2 = Q - Rr;
  const T a1 = atan2(G, Fq), a2 = atan2(H, E);
  const T th = T(0.5) * (a2 + a1), ph = T(0.5) * (a2 - a1);
  const T ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
  U[0] = ct; U[1] = -st; U[2] = st; U[3] = ct;
  V[0] = cp; V[1] = sp; V[2] = -sp; V[3] = cp;
  if (s2 < T(0)) {
    s2 = -s2;
    U[1] = -U[1];
    U[3] = -U[3];
  }
  s[0] = s1;
  s[1] = s2;
}

// -------------------------------------------------------------------------
// Segmented wave aggregation for particle-to-node atomic scatters.  Lanes on
// consecutive positions whose particles share a stencil base (key) form a
// segment; seg_reduce leaves the segment's sum in its last lane, which alone
// adds it to memory.  Particles in random order give one-lane segments (no
// change); particles ordered by cell cut the atomics by up to the wave width,
// which matters on this part, where same-address device atomics serialise.
// Every lane of the wave must call both (no early return before them).
struct WaveSeg {
  int start;  // first lane of this lane's segment
  int steps;  // doubling steps that cover the longest segment (wave-uniform)
  bool tail;  // last lane of its segment and holds a particle
};

MPM_DEV WaveSeg wave_segments(bool valid, int key) {
  const int lane = (int)__lane_id();
  const int kprev = __shfl_up(key, 1);
  const int vprev = __shfl_up(valid ? 1 : 0, 1);
  const bool head = lane == 0 || !valid || !vprev || kprev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & ((2ull << lane) - 1ull);  // heads at lanes <= this one
  WaveSeg sg;
  sg.start = 63 - __clzll((long long)upto);
  sg.tail = valid && (lane == 63 || ((heads >> (lane + 1)) & 1ull));
  int len = lane - sg.start + 1;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
  int steps = 0;
  while ((1 << steps) < len) ++steps;
  sg.steps = steps;
  return sg;
}

template <typename T> MPM_DEV T seg_reduce(T v, const WaveSeg& sg) {
  const int lane = (int)__lane_id();
  for (int i = 0; i < sg.steps; ++i) {
    const int off = 1 << i;
    const T u = __shfl_up(v, (unsigned)off);
    if (lane - off >= sg.start) v += u;
  }
  return v;
}

// -------------------------------------------------------------------------
// Sticky half-space collider of three_d/grid_op.py:50-67
// (check_collision_points): every node I of the plane's scene with
// dot(I dx - p, n) < 0 gets zero velocity.  n holds the reference's
// normal + 1/|normal| (the scalar added to each component) or, in textbook
// mode, the unit normal.  The test runs in FP64 with one rounding per
// operation, in the order ((o0 n0 + o1 n1) + o2 n2).
constexpr int MAX_PLANES = 16;
struct PlaneDev {
  double p[3], n[3], dx;
  int scene;
};

__device__ __forceinline__ bool plane_blocks(const PlaneDev& pl, int i, int j, int k) {
  const double o0 = __dsub_rn(__dmul_rn((double)i, pl.dx), pl.p[0]);
  const double o1 = __dsub_rn(__dmul_rn((double)j, pl.dx), pl.p[1]);
  const double o2 = __dsub_rn(__dmul_rn((double)k, pl.dx), pl.p[2]);
  const double d = __dadd_rn(__dadd_rn(__dmul_rn(o0, pl.n[0]), __dmul_rn(o1, pl.n[1])), __dmul_rn(o2, pl.n[2]));
  return d < 0.0;
}

// Host: the plane as the test above uses it (normal per quirks mode).
// Returns false for a zero or non-finite normal.
inline bool make_plane(const double* point, const double* normal, double dx, int quirks, int scene,
                       PlaneDev& out) {
#pragma clang fp contract(off)
  const double s0 = normal[0] * normal[0], s1 = normal[1] * normal[1], s2 = normal[2] * normal[2];
  const double denom = __builtin_sqrt((s0 + s1) + s2);
  if (!(denom > 0.0) || !(denom < __builtin_inf())) return false;
  for (int d = 0; d < 3; ++d) {
    out.p[d] = point[d];
    out.n[d] = quirks == MPM_QUIRKS_TEXTBOOK ? normal[d] / denom : normal[d] + 1.0 / denom;
  }
  out.dx = dx;
  out.scene = scene;
  return true;
}

}  // namespace mpm
