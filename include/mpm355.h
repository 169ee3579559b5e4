/*
 * mpm355 -- C ABI of the MI355X (gfx950) explicit MLS-MPM substep.
 *
 * Every entry point takes plain pointers and sizes; pointers named "device"
 * are HIP device addresses (for example torch.Tensor.data_ptr() of a CUDA
 * tensor on ROCm) and stream arguments are hipStream_t values (NULL = the
 * legacy default stream).  Each function returns an mpm_status; the message
 * of the most recent failure on the calling thread is mpm_last_error_string().
 *
 * The reference interface each entry point replaces is cited with paths
 * relative to the FEMFlow source tree (jparr721/FEMFlow).
 */
#ifndef MPM355_H
#define MPM355_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define MPM_OK 0
#define MPM_ERR_OOB 1       /* a particle's stencil left [0, res]: utils.py:138-150 raise */
#define MPM_ERR_NONFINITE 2 /* NaN/Inf found by a conservation check            */
#define MPM_ERR_ARG 3       /* bad argument (null pointer, size, dtype ...)       */
#define MPM_ERR_HIP 4       /* a HIP runtime call failed                          */

/* ---- enums ----------------------------------------------------------- */
#define MPM_F32 0
#define MPM_F64 1

#define MPM_ELASTIC 0 /* fixed-corotated, (mu, lam) = (mu0 h, lam0 h): utils.py:7-24      */
#define MPM_SNOW 1    /* e = exp(h (1 - Jp)) hardening + singular value clamp: utils.py:27-49, three_d/g2p.py:48-58 */
#define MPM_WATER 2   /* weakly compressible: PF = K (J-1) J I, F <- cbrt(J) I (build-only) */

#define MPM_QUIRKS_REFERENCE 0 /* lambda (J-1) J added to all d*d entries (utils.py:129) */
#define MPM_QUIRKS_TEXTBOOK 1  /* lambda (J-1) J * I                                    */

#define MPM_BC_REFERENCE 0 /* 3D: three_d/grid_op.py:37-47; 2D: two_d/grid_op.py:19-24 */
#define MPM_BC_STICKY 1    /* every component zero within `bound` nodes of a wall       */

#define MPM_KEY_BLOCKED 0  /* (4^d-cell block id, cell within block), both row-major */
#define MPM_KEY_ROWMAJOR 1 /* the grid's own [i, j(, k)] row-major node order          */

/* ---- plain data ------------------------------------------------------ */
typedef struct {
  int32_t kind;     /* MPM_ELASTIC / MPM_SNOW / MPM_WATER */
  float hardening;  /* h: the reference's `hardening` argument (p2g.py:16) */
  float theta_c;    /* snow: singular values clamped to [1 - theta_c, 1 + theta_s] */
  float theta_s;
} mpm_material;

/* Step parameters: the scalar arguments of solve_mls_mpm_3d (mls_mpm.py:40-53)
 * and of the phase functions (three_d/p2g.py:15-29, grid_op.py:6-13). */
typedef struct {
  int32_t dim;      /* 2 or 3 */
  int32_t res;      /* grid resolution R: (R+1)^dim nodes */
  double inv_dx, dx, dt, volume, gravity;
  int32_t boundary; /* MPM_BC_* */
  int32_t bound;    /* sticky wall thickness in nodes (MPM_BC_STICKY only) */
  int32_t quirks;   /* MPM_QUIRKS_* */
  int32_t check_oob;/* 1: count particles whose stencil leaves [0, R] (3D reference) */
} mpm_params;

/* Particle arrays in the reference's layouts (simulation.py:99-108):
 * x, v: (n, dim); F, C: (n, dim, dim) row-major; Jp: (n) [the reference's
 * (n, 1)]; mass, mu0, lambda0: (n), or one value each when mat_scalar != 0
 * (the 2D signature, two_d/p2g.py:12-29); material: (n) ids into a material
 * table, or NULL for id 0.  All device pointers of dtype `dtype`. */
typedef struct {
  int64_t n;
  int32_t dtype;      /* MPM_F32 / MPM_F64 */
  int32_t mat_scalar;
  void* x;
  void* v;
  void* F;
  void* C;
  void* Jp;
  const void* mass;
  const void* mu0;
  const void* lambda0;
  const int32_t* material;
} mpm_particles;

/* One scene of the batched engine (a grid with its own resolution, time
 * step, gravity and walls). */
typedef struct {
  int32_t res;
  int32_t boundary;
  int32_t bound;
  int32_t quirks;
  double dt, gravity, volume;
} mpm_scene;

/* ---- library ----------------------------------------------------------- */
const char* mpm_last_error_string(void);
int32_t mpm_version(void);
/* Number of visible HIP devices (0 when none); never fails for lack of a GPU. */
int32_t mpm_device_count(void);

/* ---- phase entry points on caller grids (dense reference layouts) ------
 * grid_v: (R+1)^dim x dim, grid_m: (R+1)^dim, row-major [i, j(, k)], dtype of
 * the particles.  oob_count (device int64, may be NULL) is incremented by the
 * number of particles whose stencil leaves the grid (those are skipped). */

/* Replaces three_d.p2g (three_d/p2g.py:14-80) and two_d.p2g
 * (two_d/p2g.py:11-76).  kind_override >= 0 forces one material kind for all
 * particles with hardening `hardening` (the reference's model string:
 * "neo_hookean" -> MPM_ELASTIC, anything else -> MPM_SNOW); -1 uses the
 * per-particle material ids with `materials` (host array of nmat). */
int32_t mpm_p2g(void* stream, const mpm_params* prm, const mpm_particles* p,
                int32_t kind_override, double hardening, const mpm_material* materials,
                int32_t nmat, void* grid_v, void* grid_m, int64_t* oob_count);

/* Replaces three_d.grid_op (three_d/grid_op.py:5-47) and two_d.grid_op
 * (two_d/grid_op.py:5-24). In place on grid_v. */
int32_t mpm_grid_op(void* stream, const mpm_params* prm, int32_t dtype, void* grid_v,
                    const void* grid_m);

/* Replaces three_d.grid_op.check_collision_points (three_d/grid_op.py:50-67):
 * sticky half-space colliders on a dense 3D grid_v ((R+1)^3 x 3, dtype
 * MPM_F32/MPM_F64, device).  points, normals: host (nplanes x 3) doubles,
 * nplanes <= 16.  Every node I with dot(I dx - point, n) < 0 gets v = 0,
 * where n = normal + 1/|normal| (MPM_QUIRKS_REFERENCE: the reference adds the
 * scalar to each component) or normal/|normal| (MPM_QUIRKS_TEXTBOOK).  The
 * test is evaluated in FP64. */
int32_t mpm_collide_planes(void* stream, int32_t dtype, int32_t res, double dx, int32_t nplanes,
                           const double* points, const double* normals, int32_t quirks, void* grid_v);

/* Replaces three_d.g2p (three_d/g2p.py:9-59) and two_d.g2p (two_d/g2p.py:5-47).
 * kind_override >= 0: MPM_SNOW clamps with theta_c/theta_s for every
 * particle (model == "snow"), MPM_ELASTIC never clamps. */
int32_t mpm_g2p(void* stream, const mpm_params* prm, mpm_particles* p, int32_t kind_override,
                double theta_c, double theta_s, const mpm_material* materials, int32_t nmat,
                const void* grid_v, int64_t* oob_count);

/* ---- engine: particle-resident rollout on block-sparse grids -----------
 * No reference counterpart as a single function: it runs the loop of
 * simulation.py:125-139 (solve_mls_mpm_3d per substep, mls_mpm.py:40-79)
 * for many substeps and many scenes without leaving the device.  FP32, 3D. */
typedef struct mpm_ctx mpm_ctx;

/* flags: reserved, pass 0.  res of every scene <= 1024. */
int32_t mpm_ctx_create(mpm_ctx** out, int64_t n_max, int32_t nscenes, const mpm_scene* scenes,
                       int32_t nmat, const mpm_material* materials, int32_t flags);
int32_t mpm_ctx_destroy(mpm_ctx* ctx);
int32_t mpm_ctx_set_stream(mpm_ctx* ctx, void* stream);
/* Tunables: chunk = particles per work item (256, 512 or 1024); sort_every =
 * substeps between rebinning (>= 1); grid_mult = persistent workgroups per CU;
 * use_graph = capture and replay the substep sequence as a hipGraph.  A new
 * context starts at chunk 512, sort_every 48, grid_mult 8, graphs on. */
int32_t mpm_ctx_tune(mpm_ctx* ctx, int32_t chunk, int32_t sort_every, int32_t grid_mult,
                     int32_t use_graph);
/* Scatter spread (3 or 9, default 9): the number of different starting nodes
 * that particles sharing a cell use in the P2G scatter; 9 keeps cells with
 * hundreds of particles as fast as sparse ones.  Results agree to rounding.
 * No reference counterpart (a tunable of this implementation). */
int32_t mpm_ctx_set_spread(mpm_ctx* ctx, int32_t spread);

/* Per-kernel timing with HIP events on the engine's stream (eager launches
 * only; graphs are not used while enabled).  mpm_ctx_kernel_times returns
 * total milliseconds and launch counts for [segment-start pass (a run's
 * first P2G, or the fused pass that moves the particles after a rebin inside
 * a run), fused G2P2G, G2P, grid update] since the previous call. */
int32_t mpm_ctx_kernel_timing(mpm_ctx* ctx, int32_t enable);
int32_t mpm_ctx_kernel_times(mpm_ctx* ctx, double* ms, int64_t* counts);

/* Load particles (device FP32 arrays, reference layouts, material ids
 * required unless nmat == 1) with an optional per-particle scene id (device
 * int32, NULL = scene 0).  The engine keeps its own binned copy. */
int32_t mpm_state_upload(mpm_ctx* ctx, const mpm_particles* p, const int32_t* scene_ids);
/* Write the engine's state back to caller arrays in the caller's order
 * (x, v, F, C, Jp). */
int32_t mpm_state_download(mpm_ctx* ctx, mpm_particles* p);
/* Advance every scene by nsteps substeps.  If sums != NULL (device double,
 * nsteps * 8 entries) each substep writes FP64 [grid mass, grid momentum x3
 * after P2G, grid momentum x3 after the update, number of nodes with mass]. */
int32_t mpm_run(mpm_ctx* ctx, int32_t nsteps, double* sums);
/* Like mpm_run (no sums), also writing each substep's positions in the
 * caller's order to frames (device float, nsteps * n * 3): the per-step frame
 * record of simulation.py:140-142 (map_particles_to_pos, particle.py:30-33). */
int32_t mpm_run_frames(mpm_ctx* ctx, int32_t nsteps, float* frames);
/* nsteps substeps writing the full particle state after substep k (caller
 * order) to slot k: x, v (nsteps x n x 3), F, C (nsteps x n x 3 x 3), Jp
 * (nsteps x n), FP32 device arrays.  No reference counterpart: it serves
 * differentiable rollouts, which replay the states of a checkpoint interval. */
int32_t mpm_run_states(mpm_ctx* ctx, int32_t nsteps, float* x, float* v, float* F, float* C, float* Jp);
/* mpm_run_states that also records each substep's dense grids: slot k of
 * grids (device float, nsteps x 7 (res+1)^3) receives substep k's P2G
 * momentum ((res+1)^3 x 3), P2G mass ((res+1)^3) and updated velocity
 * ((res+1)^3 x 3), zero where no particle reached, in the layout
 * mpm_adjoint_step_grids reads.  Single-scene contexts without a corrector.
 * No reference counterpart: the reverse pass of a differentiable rollout
 * takes its forward grids from here instead of recomputing them. */
int32_t mpm_run_states_grids(mpm_ctx* ctx, int32_t nsteps, float* x, float* v, float* F, float* C, float* Jp,
                             float* grids);
/* Current positions in the caller's order (device float, n * 3). */
int32_t mpm_state_positions(mpm_ctx* ctx, float* x);
/* Convenience for the reference's per-call pattern: upload, run, download. */
int32_t mpm_substep(mpm_ctx* ctx, mpm_particles* p, const int32_t* scene_ids, int32_t nsteps);
/* Particles whose stencil left their grid since the context was created or
 * this counter was last read (reading resets it).  Synchronises the stream.
 * A value of 2^40 or more means a grid update's bounded wait for the FP64
 * pass over deferred particles ran out (an internal error, not particles). */
int32_t mpm_ctx_oob_count(mpm_ctx* ctx, int64_t* out);
/* FP32 polar/SVD evaluations that fell back to FP64 (reading resets it). */
int32_t mpm_ctx_fallback_count(mpm_ctx* ctx, int64_t* out);
/* Grid blocks (4x4x4 nodes) the last substep's P2G reached: the work list of
 * its grid update.  Synchronises the stream.  No reference counterpart (the
 * reference sweeps the dense grid, three_d/grid_op.py:25-47); it sizes the
 * per-node term of the algorithmic byte count. */
int32_t mpm_ctx_active_blocks(mpm_ctx* ctx, int64_t* out);
/* Rebin the particles at the start of the next mpm_run.  The engine rebins
 * after every upload and then every sort_every substeps, counting across
 * calls; this restarts that count.  No reference counterpart. */
int32_t mpm_ctx_request_rebin(mpm_ctx* ctx);
/* FP64 sums over the engine's particles: out[0] = sum m, out[1..3] = sum m v,
 * out[4..6] = sum m x (host double[7]). Synchronises the stream. */
int32_t mpm_particle_sums(mpm_ctx* ctx, double* out);

/* ---- binning: cell keys and the stable radix sort ----------------------
 * keys[i] = cell key of trunc(fl32(fl32(x*inv_dx) - 0.5)) clamped to [0, res];
 * perm = the stable ordering of keys (equal to numpy.argsort(keys,
 * kind="stable")).  x: device float (n, dim).  keys, perm: device uint32 (n).
 * No reference counterpart (the reference walks particles in list order,
 * three_d/p2g.py:49); the key derives from its base, three_d/p2g.py:50. */
int32_t mpm_bin(mpm_ctx* ctx, int32_t dim, int32_t res, double inv_dx, int32_t layout,
                const float* x, int64_t n, uint32_t* keys, uint32_t* perm);
/* mpm_bin that allocates nothing: the same arguments and bit-identical keys
 * and perm, with the sort's workspace in a fixed device array of the library
 * (about 20 MiB per GPU, held until the process ends) instead of per-call
 * hipMalloc/hipFree.  n is limited to 2^20 (larger n: MPM_ERR_ARG).  Calls
 * that overlap on one GPU share that array and race.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, allocation-free harnesses). */
int32_t mpm_bin_noalloc(mpm_ctx* ctx, int32_t dim, int32_t res, double inv_dx, int32_t layout,
                        const float* x, int64_t n, uint32_t* keys, uint32_t* perm);

/* ---- differentiable rollouts: dense-grid substep and its adjoint ------
 * Elastic (fixed-corotated, utils.py:95-135 with constant_hardening,
 * utils.py:7-24) 3D substeps on a dense (res+1)^3 grid held in a caller
 * workspace of mpm_dense_workspace_bytes() bytes (device, 16-byte aligned).
 * FP32 particles (p->dtype == MPM_F32, per-particle or scalar mass/mu0/lambda0).
 * No reference counterpart: the reference has no gradients.  The substep is
 * the one mls_mpm.py:40-79 takes (three_d/p2g.py:14-80,
 * three_d/grid_op.py:5-47, three_d/g2p.py:9-59) and is what the adjoint
 * reverses. */
int32_t mpm_dense_workspace_bytes(const mpm_params* prm, int64_t* bytes);
/* One forward substep in place on p (x, v, F, C), using the workspace grids. */
int32_t mpm_dense_substep(void* stream, const mpm_params* prm, mpm_particles* p, double hardening,
                          void* work, int64_t* oob_count);
/* Reverse of one substep taken from state p (unchanged).  On entry x_bar,
 * v_bar, F_bar, C_bar (device float, n*3 / n*3 / n*9 / n*9) hold the adjoints
 * of the substep's outputs; on return, the adjoints of its inputs.  mu0_bar
 * and lambda0_bar (n floats, or 1 when p->mat_scalar) are accumulated into.
 * The stencil base, the velocity clamp and the wall masks are differentiated
 * piecewise (held fixed / no gradient where active). */
int32_t mpm_adjoint_step(void* stream, const mpm_params* prm, const mpm_particles* p, double hardening,
                         void* work, float* x_bar, float* v_bar, float* F_bar, float* C_bar,
                         float* mu0_bar, float* lambda0_bar, int64_t* oob_count);
/* mpm_adjoint_step with the substep's forward grids given (one slot of
 * mpm_run_states_grids: P2G momentum, P2G mass, updated velocity) instead of
 * recomputed from p; work is the same workspace (its gradient grid is used). */
int32_t mpm_adjoint_step_grids(void* stream, const mpm_params* prm, const mpm_particles* p, double hardening,
                               const float* grids, void* work, float* x_bar, float* v_bar, float* F_bar,
                               float* C_bar, float* mu0_bar, float* lambda0_bar, int64_t* oob_count);

/* ---- learned corrector inside an engine context --------------------------
 * No reference counterpart.  params: device FP32 MLP parameters (layout of
 * mpm_mlp_sizes / mpm_mlp_pack); m_ref, v_ref: feature scales (node mass and
 * velocity).  enable = 1: every substep of mpm_run adds the network's
 * correction to the updated grid velocity (the context's scene dt is the big
 * step); 0: plain substeps. */
int32_t mpm_ctx_set_corrector(mpm_ctx* ctx, const float* params, float m_ref, float v_ref,
                              int32_t enable);
/* Sticky half-space colliders of one scene of the engine, applied by every
 * grid update after the walls (the rule of mpm_collide_planes; the engine
 * grid spans [0, 1]^3 with dx = 1/res).  Replaces the scene's previous
 * planes; nplanes = 0 removes them; at most 16 planes per context.  The
 * reference defines the collider (three_d/grid_op.py:50-67) but its step
 * never calls it. */
int32_t mpm_ctx_set_planes(mpm_ctx* ctx, int32_t scene, int32_t nplanes, const double* points,
                           const double* normals, int32_t quirks);
/* Training data: one P2G + grid update from the current state (particles
 * unchanged), then for every node with mass (up to max_rows) its features
 * (X, rows x 352 BF16), global node index and updated velocity (rows x 3).
 * X == NULL: only the P2G + grid update and, if enabled, the correction
 * (then mpm_grid_velocity).  Needs a
 * corrector set once (enable may be 0).  Synchronises; *rows = nodes found. */
int32_t mpm_corrector_sample(mpm_ctx* ctx, uint16_t* X, uint32_t* nodes, float* vel, int64_t max_rows,
                             int64_t* rows);
/* Updated grid velocity (count x 3) at global node indices from the last
 * grid update. */
int32_t mpm_grid_velocity(mpm_ctx* ctx, const uint32_t* nodes, int64_t count, float* out);

/* ---- grid-corrector MLP (BF16 MFMA, FP32 accumulate) -------------------
 * 351 node-stencil features -> 128 -> 128 -> 3, ReLU hidden layers.  No
 * reference counterpart (the reference has no learned component).  Device
 * buffers, row-major: X rows x 352 BF16 (column 351 zero), H1/H2/dH1/dH2
 * rows x 128 BF16, Y rows x 16 FP32 (3 used), targets T rows x 3 FP32,
 * dY rows x 32 BF16.  params: FP32 master copy (n_params floats: W1 [128][351],
 * b1, W2 [128][128], b2, W3 [3][128], b3); packed: BF16 operand copies
 * (n_packed uint16) made by mpm_mlp_pack. */
int32_t mpm_mlp_sizes(int64_t* n_params, int64_t* n_packed);
int32_t mpm_mlp_pack(void* stream, const float* params, uint16_t* packed);
/* Y = MLP(X); H1, H2 receive the hidden activations (needed by backward). */
int32_t mpm_mlp_forward(void* stream, const uint16_t* packed, const float* params, const uint16_t* X,
                        int64_t rows, uint16_t* H1, uint16_t* H2, float* Y);
/* loss += inv_norm * sum_i w_i |Y_i - T_i|^2 (w NULL = 1); dY = dloss/dY. */
int32_t mpm_mlp_loss(void* stream, const float* Y, const float* T, const float* w, int64_t rows,
                     float inv_norm, uint16_t* dY, double* loss);
/* grads (n_params floats, accumulated) of the loss whose dY is given. */
int32_t mpm_mlp_backward(void* stream, const uint16_t* packed, const uint16_t* X, int64_t rows,
                         const uint16_t* H1, const uint16_t* H2, const uint16_t* dY, uint16_t* dH2,
                         uint16_t* dH1, float* grads);
/* One Adam step (FP32) on n parameters; step counts from 1. */
int32_t mpm_mlp_adam(void* stream, float* params, const float* grads, float* m, float* v, int64_t n,
                     float lr, float beta1, float beta2, float eps, int32_t step);

#ifdef __cplusplus
}
#endif

#endif /* MPM355_H */
