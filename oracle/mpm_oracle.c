/*
 * C restatement of FEMFlow's explicit MLS-MPM substep, FP64, in the
 * reference's loop order.  TEST INFRASTRUCTURE ONLY: tests/ use it as a second
 * restatement beside oracle/mpm_oracle.py, and bench.py times it as the CPU
 * baseline (single thread like the reference's @nb.njit, or OpenMP).  The
 * femflow_amd package never links or calls it.
 *
 * Follows (paths relative to the FEMFlow source tree):
 *   femflow/solvers/mpm/mls_mpm.py:40-79        substep driver
 *   femflow/solvers/mpm/three_d/p2g.py:14-80    P2G, particle-major, (i,j,k) lexicographic
 *   femflow/solvers/mpm/three_d/grid_op.py:5-47 grid update (slip walls, clamp)
 *   femflow/solvers/mpm/three_d/g2p.py:9-59     G2P, F update, snow plasticity
 *   femflow/solvers/mpm/two_d/{p2g,grid_op,g2p}.py  2D variants
 *   femflow/solvers/mpm/utils.py:7-150          hardening, stress, oob
 *   femflow/numerics/linear_algebra.py:96-135   polar decompositions
 * The 3x3 SVD is a cyclic Jacobi eigen-solve of F^T F (own code, FP64); the
 * reference calls LAPACK there, so only gauge-free quantities (R = U V^T,
 * singular values, the clamped F in the U S V^T gauge) are compared.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ELASTIC = 0, SNOW = 1, WATER = 2 };

static double det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

static void matmul3(const double *a, const double *b, double *c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = s;
    }
}

/* SVD of a 3x3 matrix, F = U diag(s) V^T with s >= 0 sorted descending. */
static void svd3(const double *F, double *U, double *s, double *V) {
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += F[3 * k + i] * F[3 * k + j];
      A[3 * i + j] = t;
    }
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double apq = A[3 * p + q];
        if (apq == 0.0) continue;
        double theta = (A[3 * q + q] - A[3 * p + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 3; ++k) { /* A <- A G */
          double akp = A[3 * k + p], akq = A[3 * k + q];
          A[3 * k + p] = c * akp - sn * akq;
          A[3 * k + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) { /* A <- G^T A */
          double apk = A[3 * p + k], aqk = A[3 * q + k];
          A[3 * p + k] = c * apk - sn * aqk;
          A[3 * q + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[3 * k + p], vkq = V[3 * k + q];
          V[3 * k + p] = c * vkp - sn * vkq;
          V[3 * k + q] = sn * vkp + c * vkq;
        }
      }
  }
  double ev[3] = {A[0], A[4], A[8]};
  int idx[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (ev[idx[j]] > ev[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
  double Vs[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) Vs[3 * r + c] = V[3 * r + idx[c]];
  memcpy(V, Vs, sizeof(Vs));
  /* B = F V, then Gram-Schmidt for U; s_i = |column| */
  double B[9];
  matmul3(F, V, B);
  for (int c = 0; c < 3; ++c) {
    double col[3] = {B[c], B[3 + c], B[6 + c]};
    for (int pc = 0; pc < c; ++pc) {
      double d = col[0] * U[pc] + col[1] * U[3 + pc] + col[2] * U[6 + pc];
      for (int r = 0; r < 3; ++r) col[r] -= d * U[3 * r + pc];
    }
    double nrm = sqrt(col[0] * col[0] + col[1] * col[1] + col[2] * col[2]);
    s[c] = sqrt(fmax(ev[idx[c]], 0.0));
    if (nrm > 1e-300) {
      for (int r = 0; r < 3; ++r) U[3 * r + c] = col[r] / nrm;
    } else { /* complete the basis */
      double u0[3] = {U[0], U[3], U[6]}, u1[3] = {U[1], U[4], U[7]};
      double cr[3] = {u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2],
                      u0[0] * u1[1] - u0[1] * u1[0]};
      for (int r = 0; r < 3; ++r) U[3 * r + c] = cr[r];
    }
  }
  /* singular values from the projection keep sign information out: s >= 0 */
  for (int c = 0; c < 3; ++c) {
    double d = 0.0;
    for (int r = 0; r < 3; ++r) d += U[3 * r + c] * B[3 * r + c];
    s[c] = fabs(d);
    if (d < 0) for (int r = 0; r < 3; ++r) U[3 * r + c] = -U[3 * r + c];
  }
}

static void lame(int kind, double h, double mu0, double lam0, double Jp, double *mu, double *lam) {
  if (kind == WATER) { *mu = 0.0; *lam = lam0; return; }
  double e = (kind == SNOW) ? exp(h * (1.0 - Jp)) : h;
  *mu = mu0 * e;
  *lam = lam0 * e;
}

/* utils.py:95-135 (3D) and :52-92 (2D) */
static void affine3(const double *F, const double *C, double mass, double mu, double lam, int kind,
                    double inv_dx, double dt, double vol, int textbook, double *A) {
  double U[9], s[3], V[9], R[9] = {0};
  svd3(F, U, s, V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) R[3 * i + j] += U[3 * i + k] * V[3 * j + k];
  double J = det3(F);
  double D_inv = 4 * inv_dx * inv_dx;
  double vt = lam * (J - 1) * J;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double pf = 0.0;
      for (int k = 0; k < 3; ++k) pf += (2 * mu * (F[3 * i + k] - R[3 * i + k])) * F[3 * j + k];
      if (kind == WATER || textbook) pf += (i == j) ? vt : 0.0;
      else pf += vt;
      A[3 * i + j] = -(dt * vol) * (D_inv * pf) + mass * C[3 * i + j];
    }
}

static void affine2(const double *F, const double *C, double mass, double mu, double lam, int kind,
                    double inv_dx, double dt, double vol, int textbook, double *A) {
  double J = F[0] * F[3] - F[1] * F[2];
  double x = F[0] + F[3], y = F[2] - F[1];
  double denom = sqrt(x * x + y * y) + (textbook ? 0.0 : 1e-10);
  double sc = 1.0 / denom, c = x * sc, s = y * sc;
  double r[4] = {c, -s, s, c};
  double D_inv = 4 * inv_dx * inv_dx;
  double vt = lam * (J - 1) * J;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      double pf = 0.0;
      for (int k = 0; k < 2; ++k) pf += (2 * mu * (F[2 * i + k] - r[2 * i + k])) * F[2 * j + k];
      if (kind == WATER || textbook) pf += (i == j) ? vt : 0.0;
      else pf += vt;
      A[2 * i + j] = -(dt * vol) * (D_inv * pf) + mass * C[2 * i + j];
    }
}

typedef struct {
  int64_t n;
  double *x, *v, *F, *C, *Jp;
  const double *mass, *mu0, *lam0, *h, *theta_c, *theta_s;
  const int32_t *kind;
} oracle_particles;

typedef struct {
  int res;
  double inv_dx, dx, dt, volume, gravity;
  int textbook;   /* 0: reference quirks */
  int sticky;     /* 0: reference boundary */
  int bound;      /* sticky thickness in nodes */
} oracle_params;

static inline void weights1(double fx, double *w) {
  w[0] = 0.5 * (1.5 - fx) * (1.5 - fx);
  w[1] = 0.75 - (fx - 1) * (fx - 1);
  w[2] = 0.5 * (fx - 0.5) * (fx - 0.5);
}

/* One particle of three_d/p2g.py:49-80, restricted to the nodes whose first
 * index lies in [ilo, ihi): weights, stress + affine, then the 27 nodes in
 * lexicographic (i, j, k) order.  Returns 1 if its stencil leaves the grid
 * (utils.py:138-150), adding nothing. */
static int p2g_particle3d(const oracle_particles *P, const oracle_params *prm, int64_t p, int64_t ilo,
                          int64_t ihi, double *grid_v, double *grid_m) {
  const int64_t r1 = prm->res + 1;
  const double *xp = P->x + 3 * p;
  int64_t base[3];
  double fx[3], w[3][3];
  int bad = 0;
  for (int d = 0; d < 3; ++d) {
    double sd = xp[d] * prm->inv_dx;
    base[d] = (int64_t)(sd - 0.5);
    fx[d] = sd - (double)base[d];
    weights1(fx[d], w[d]);
    if (base[d] < 0 || base[d] + 2 >= r1) bad = 1;
  }
  if (bad) return 1;
  double mu, lam;
  lame(P->kind[p], P->h[p], P->mu0[p], P->lam0[p], P->Jp[p], &mu, &lam);
  double A[9];
  affine3(P->F + 9 * p, P->C + 9 * p, P->mass[p], mu, lam, P->kind[p], prm->inv_dx, prm->dt,
          prm->volume, prm->textbook, A);
  const double m = P->mass[p];
  const double *vp = P->v + 3 * p;
  for (int i = 0; i < 3; ++i) {
    if (base[0] + i < ilo || base[0] + i >= ihi) continue;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double dpos[3] = {(i - fx[0]) * prm->dx, (j - fx[1]) * prm->dx, (k - fx[2]) * prm->dx};
        double weight = w[0][i] * w[1][j] * w[2][k];
        int64_t node = ((base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
        for (int d = 0; d < 3; ++d) {
          double ad = A[3 * d] * dpos[0] + A[3 * d + 1] * dpos[1] + A[3 * d + 2] * dpos[2];
          grid_v[3 * node + d] += weight * (vp[d] * m + ad);
        }
        grid_m[node] += weight * m;
      }
  }
  return 0;
}

/* three_d/p2g.py:14-80.  grid_v: (res+1)^3 x 3, grid_m: (res+1)^3; accumulated.
 * With several threads, each owns a slab of node planes (first index in
 * [ilo, ihi), slabs balanced by the particles that reach them) and walks
 * every particle in list order, adding only to its own nodes.  Each node
 * therefore receives its contributions in the reference's order (particle
 * major, then i, j, k), and the result is bit-identical to one thread's
 * whatever the thread count: a deterministic FP64 reference (finite
 * differences of two rollouts see no summation-order noise).  Particles
 * whose stencil spans two slabs are evaluated by both threads. */
int oracle_p2g3d(const oracle_particles *P, const oracle_params *prm, double *grid_v,
                 double *grid_m, int nthreads) {
  const int64_t r1 = prm->res + 1;
  int oob = 0;
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  if (nt <= 1 || P->n == 0) {
    for (int64_t p = 0; p < P->n; ++p) oob |= p2g_particle3d(P, prm, p, 0, r1, grid_v, grid_m);
    return oob ? 1 : 0;
  }
  /* work per node plane i: particles whose stencil base is i-2, i-1 or i */
  int64_t *work = (int64_t *)calloc((size_t)r1 + 1, sizeof(int64_t));
  int64_t *bnd = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nt + 1));
  if (!work || !bnd) {
    free(work);
    free(bnd);
    for (int64_t p = 0; p < P->n; ++p) oob |= p2g_particle3d(P, prm, p, 0, r1, grid_v, grid_m);
    return oob ? 1 : 0;
  }
  for (int64_t p = 0; p < P->n; ++p) {
    int64_t b = (int64_t)(P->x[3 * p] * prm->inv_dx - 0.5);
    b = b < 0 ? 0 : (b > r1 - 1 ? r1 - 1 : b);
    for (int i = 0; i < 3 && b + i < r1; ++i) work[b + i] += 1;
  }
  int64_t total = 0;
  for (int64_t i = 0; i < r1; ++i) total += work[i];
  bnd[0] = 0;
  {
    int64_t acc = 0, i = 0;
    for (int t = 1; t < nt; ++t) {
      const int64_t goal = (total * t + nt - 1) / nt;
      while (i < r1 && acc < goal) acc += work[i++];
      bnd[t] = i;
    }
    bnd[nt] = r1;
  }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(| : oob)
#endif
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const int64_t ilo = bnd[tid], ihi = bnd[tid + 1];
    if (ilo < ihi) {
      for (int64_t p = 0; p < P->n; ++p) {
        const int64_t b = (int64_t)(P->x[3 * p] * prm->inv_dx - 0.5);
        if (b + 2 < ilo || b >= ihi) continue;
        oob |= p2g_particle3d(P, prm, p, ilo, ihi, grid_v, grid_m);
      }
    }
  }
  free(work);
  free(bnd);
  return oob ? 1 : 0;
}

/* three_d/grid_op.py:5-47 */
void oracle_grid_op3d(const oracle_params *prm, double *grid_v, const double *grid_m, int nthreads) {
  const int64_t r1 = prm->res + 1;
  const double v_allowed = prm->dx * 0.9 / prm->dt;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
  for (int64_t i = 0; i < r1; ++i)
    for (int64_t j = 0; j < r1; ++j)
      for (int64_t k = 0; k < r1; ++k) {
        int64_t node = (i * r1 + j) * r1 + k;
        double *g = grid_v + 3 * node;
        double m = grid_m[node];
        if (m > 0) {
          for (int d = 0; d < 3; ++d) g[d] /= m;
          g[1] += prm->dt * prm->gravity;
          for (int d = 0; d < 3; ++d) g[d] = fmin(fmax(g[d], -v_allowed), v_allowed);
        }
        int64_t I[3] = {i, j, k};
        if (!prm->sticky) {
          for (int d = 0; d < 3; ++d)
            if (I[d] < 1 || I[d] >= prm->res - 1) g[d] = 0;
        } else {
          for (int d = 0; d < 3; ++d)
            if (I[d] < prm->bound || I[d] > prm->res - prm->bound) g[0] = g[1] = g[2] = 0;
        }
      }
}

/* three_d/g2p.py:9-59, consistent SVD gauge; WATER resets F to cbrt(J) I. */
int oracle_g2p3d(oracle_particles *P, const oracle_params *prm, const double *grid_v, int nthreads) {
  const int64_t r1 = prm->res + 1;
  int oob = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : oob) if (nthreads != 1)
#endif
  for (int64_t p = 0; p < P->n; ++p) {
    double *xp = P->x + 3 * p;
    int64_t base[3];
    double fx[3], w[3][3];
    int bad = 0;
    for (int d = 0; d < 3; ++d) {
      double sd = xp[d] * prm->inv_dx;
      base[d] = (int64_t)(sd - 0.5);
      fx[d] = sd - (double)base[d];
      weights1(fx[d], w[d]);
      if (base[d] < 0 || base[d] + 2 >= r1) bad = 1;
    }
    if (bad) { oob |= 1; continue; }
    double vn[3] = {0, 0, 0}, Cn[9] = {0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          double dpos[3] = {i - fx[0], j - fx[1], k - fx[2]};
          int64_t node = ((base[0] + i) * r1 + (base[1] + j)) * r1 + (base[2] + k);
          const double *g = grid_v + 3 * node;
          double weight = w[0][i] * w[1][j] * w[2][k];
          for (int d = 0; d < 3; ++d) vn[d] += weight * g[d];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Cn[3 * a + b] += 4 * prm->inv_dx * ((weight * g[a]) * dpos[b]);
        }
    double *vp = P->v + 3 * p, *Cp = P->C + 9 * p, *Fp = P->F + 9 * p;
    for (int d = 0; d < 3; ++d) { vp[d] = vn[d]; xp[d] += prm->dt * vn[d]; }
    memcpy(Cp, Cn, sizeof(Cn));
    double M[9], Fn[9];
    for (int i = 0; i < 9; ++i) M[i] = ((i % 4) == 0 ? 1.0 : 0.0) + prm->dt * Cn[i];
    matmul3(M, Fp, Fn);
    int kind = P->kind[p];
    if (kind == SNOW) {
      double U[9], s[3], V[9];
      svd3(Fn, U, s, V);
      double lo = 1.0 - P->theta_c[p], hi = 1.0 + P->theta_s[p];
      double old_J = det3(Fn);
      for (int c = 0; c < 3; ++c) s[c] = fmin(fmax(s[c], lo), hi);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0.0;
          for (int c = 0; c < 3; ++c) t += U[3 * i + c] * s[c] * V[3 * j + c];
          Fn[3 * i + j] = t;
        }
      double det = det3(Fn) + 1e-10;
      P->Jp[p] = fmin(fmax(P->Jp[p] * old_J / det, 0.6), 20.0);
    } else if (kind == WATER) {
      double J = det3(Fn), c = cbrt(J);
      for (int i = 0; i < 9; ++i) Fn[i] = (i % 4 == 0) ? c : 0.0;
    }
    memcpy(Fp, Fn, sizeof(Fn));
  }
  return oob ? 1 : 0;
}

/* One 3D substep (mls_mpm.py:40-79) with caller-provided scratch grids. */
int oracle_substep3d(oracle_particles *P, const oracle_params *prm, double *grid_v, double *grid_m,
                     int nthreads) {
  const int64_t r1 = prm->res + 1;
  memset(grid_v, 0, sizeof(double) * 3 * r1 * r1 * r1);
  memset(grid_m, 0, sizeof(double) * r1 * r1 * r1);
  if (oracle_p2g3d(P, prm, grid_v, grid_m, nthreads)) return 1;
  oracle_grid_op3d(prm, grid_v, grid_m, nthreads);
  return oracle_g2p3d(P, prm, grid_v, nthreads);
}

/* ---------------- 2D (two_d/*.py) ---------------- */
void oracle_p2g2d(const oracle_particles *P, const oracle_params *prm, double *grid_v,
                  double *grid_m) {
  const int64_t r1 = prm->res + 1;
  for (int64_t p = 0; p < P->n; ++p) {
    const double *xp = P->x + 2 * p;
    int64_t base[2];
    double fx[2], w[2][3];
    for (int d = 0; d < 2; ++d) {
      double sd = xp[d] * prm->inv_dx;
      base[d] = (int64_t)(sd - 0.5);
      fx[d] = sd - (double)base[d];
      weights1(fx[d], w[d]);
    }
    double mu, lam, A[4];
    lame(P->kind[p], P->h[p], P->mu0[p], P->lam0[p], P->Jp[p], &mu, &lam);
    affine2(P->F + 4 * p, P->C + 4 * p, P->mass[p], mu, lam, P->kind[p], prm->inv_dx, prm->dt,
            prm->volume, prm->textbook, A);
    const double m = P->mass[p];
    const double *vp = P->v + 2 * p;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double dpos[2] = {(i - fx[0]) * prm->dx, (j - fx[1]) * prm->dx};
        double weight = w[0][i] * w[1][j];
        int64_t node = (base[0] + i) * r1 + (base[1] + j);
        for (int d = 0; d < 2; ++d) {
          double ad = A[2 * d] * dpos[0] + A[2 * d + 1] * dpos[1];
          grid_v[2 * node + d] += weight * (vp[d] * m + ad);
        }
        grid_m[node] += weight * m;
      }
  }
}

void oracle_grid_op2d(const oracle_params *prm, double *grid_v, const double *grid_m) {
  const int64_t r1 = prm->res + 1;
  for (int64_t i = 0; i < r1; ++i)
    for (int64_t j = 0; j < r1; ++j) {
      int64_t node = i * r1 + j;
      double m = grid_m[node];
      if (!(m > 0)) continue;
      double *g = grid_v + 2 * node;
      g[0] /= m;
      g[1] /= m;
      g[1] += prm->dt * prm->gravity;
      if (!prm->sticky) {
        double x = (double)i / prm->res, y = (double)j / prm->res;
        if (x < 0.05 || x > 1 - 0.05 || y > 1 - 0.05) g[0] = g[1] = 0.0;
        if (y < 0.05) g[1] = fmax(0.0, g[1]);
      } else if (i < prm->bound || i > prm->res - prm->bound || j < prm->bound ||
                 j > prm->res - prm->bound) {
        g[0] = g[1] = 0.0;
      }
    }
}

/* two_d/g2p.py:5-47; the SVD rebuild runs for every particle (consistent gauge). */
void oracle_g2p2d(oracle_particles *P, const oracle_params *prm, const double *grid_v) {
  const int64_t r1 = prm->res + 1;
  for (int64_t p = 0; p < P->n; ++p) {
    double *xp = P->x + 2 * p;
    int64_t base[2];
    double fx[2], w[2][3];
    for (int d = 0; d < 2; ++d) {
      double sd = xp[d] * prm->inv_dx;
      base[d] = (int64_t)(sd - 0.5);
      fx[d] = sd - (double)base[d];
      weights1(fx[d], w[d]);
    }
    double vn[2] = {0, 0}, Cn[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double dpos[2] = {i - fx[0], j - fx[1]};
        int64_t node = (base[0] + i) * r1 + (base[1] + j);
        const double *g = grid_v + 2 * node;
        double weight = w[0][i] * w[1][j];
        for (int d = 0; d < 2; ++d) vn[d] += weight * g[d];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) Cn[2 * a + b] += 4 * prm->inv_dx * ((weight * g[a]) * dpos[b]);
      }
    double *vp = P->v + 2 * p, *Cp = P->C + 4 * p, *Fp = P->F + 4 * p;
    for (int d = 0; d < 2; ++d) { vp[d] = vn[d]; xp[d] += prm->dt * vn[d]; }
    memcpy(Cp, Cn, sizeof(Cn));
    double M[4] = {1 + prm->dt * Cn[0], prm->dt * Cn[1], prm->dt * Cn[2], 1 + prm->dt * Cn[3]};
    double Fn[4] = {M[0] * Fp[0] + M[1] * Fp[2], M[0] * Fp[1] + M[1] * Fp[3],
                    M[2] * Fp[0] + M[3] * Fp[2], M[2] * Fp[1] + M[3] * Fp[3]};
    /* 2x2 SVD: F = R(theta) diag(s1, s2) R(phi)^T with s1 >= |s2|, then s >= 0 */
    double E = 0.5 * (Fn[0] + Fn[3]), Fq = 0.5 * (Fn[0] - Fn[3]);
    double G = 0.5 * (Fn[2] + Fn[1]), H = 0.5 * (Fn[2] - Fn[1]);
    double Q = sqrt(E * E + H * H), Rr = sqrt(Fq * Fq + G * G);
    double s1 = Q + Rr, s2 = Q - Rr;
    double a1 = atan2(G, Fq), a2 = atan2(H, E);
    double th = 0.5 * (a2 + a1), ph = 0.5 * (a2 - a1);
    double U[4] = {cos(th), -sin(th), sin(th), cos(th)};
    double Vm[4] = {cos(ph), sin(ph), -sin(ph), cos(ph)}; /* V = rot(theta)^T */
    if (s2 < 0) { s2 = -s2; U[1] = -U[1]; U[3] = -U[3]; }
    double old_J = Fn[0] * Fn[3] - Fn[1] * Fn[2];
    if (P->kind[p] == SNOW) {
      double lo = 1.0 - P->theta_c[p], hi = 1.0 + P->theta_s[p];
      s1 = fmin(fmax(s1, lo), hi);
      s2 = fmin(fmax(s2, lo), hi);
    }
    double S[2] = {s1, s2};
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) Fn[2 * i + j] = U[2 * i] * S[0] * Vm[2 * j] + U[2 * i + 1] * S[1] * Vm[2 * j + 1];
    double det = Fn[0] * Fn[3] - Fn[1] * Fn[2] + 1e-10;
    P->Jp[p] = fmin(fmax(P->Jp[p] * old_J / det, 0.6), 20.0);
    memcpy(Fp, Fn, sizeof(Fn));
  }
}

void oracle_substep2d(oracle_particles *P, const oracle_params *prm, double *grid_v, double *grid_m) {
  const int64_t r1 = prm->res + 1;
  memset(grid_v, 0, sizeof(double) * 2 * r1 * r1);
  memset(grid_m, 0, sizeof(double) * r1 * r1);
  oracle_p2g2d(P, prm, grid_v, grid_m);
  oracle_grid_op2d(prm, grid_v, grid_m);
  oracle_g2p2d(P, prm, grid_v);
}

/* Run n substeps; returns 0 or 1 (out of bounds). */
int oracle_run(int dim, oracle_particles *P, const oracle_params *prm, double *grid_v,
               double *grid_m, int nsteps, int nthreads) {
  for (int s = 0; s < nsteps; ++s) {
    if (dim == 3) {
      if (oracle_substep3d(P, prm, grid_v, grid_m, nthreads)) return 1;
    } else {
      oracle_substep2d(P, prm, grid_v, grid_m);
    }
  }
  return 0;
}
