// extern "C" entry points declared in include/mpm355.h.
#include <string.h>

#include <string>

#include "mpm355.h"
#include "mpm_engine.h"
#include "mpm_mlp.h"
#include "mpm_phase.h"
#include "mpm_sort.h"

namespace mpm {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace mpm

using mpm::set_error;

struct mpm_ctx {
  mpm::Engine eng;
};

#define CHECK_HIP(expr)                                                   \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      return MPM_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

static int check_params(const mpm_params* prm, const char* fn) {
  if (!prm || (prm->dim != 2 && prm->dim != 3) || prm->res < 3 || prm->res > 4096) {
    set_error(std::string(fn) + ": params must have dim in {2,3} and 3 <= res <= 4096");
    return MPM_ERR_ARG;
  }
  return MPM_OK;
}

static int check_particles(const mpm_particles* p, const char* fn, bool need_material) {
  if (!p || p->n < 0 || (p->dtype != MPM_F32 && p->dtype != MPM_F64)) {
    set_error(std::string(fn) + ": particles must have n >= 0 and dtype MPM_F32/MPM_F64");
    return MPM_ERR_ARG;
  }
  if (p->n > 0 && (!p->x || !p->v || !p->F || !p->C || !p->Jp ||
                   (need_material && (!p->mass || !p->mu0 || !p->lambda0)))) {
    set_error(std::string(fn) + ": null particle array");
    return MPM_ERR_ARG;
  }
  return MPM_OK;
}

extern "C" {

const char* mpm_last_error_string(void) { return mpm::g_last_error.c_str(); }

int32_t mpm_version(void) { return 1; }

int32_t mpm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int32_t mpm_p2g(void* stream, const mpm_params* prm, const mpm_particles* p, int32_t kind_override,
                double hardening, const mpm_material* materials, int32_t nmat, void* grid_v,
                void* grid_m, int64_t* oob_count) {
  int rc;
  if ((rc = check_params(prm, "mpm_p2g")) || (rc = check_particles(p, "mpm_p2g", true))) return rc;
  if (!grid_v || !grid_m) {
    set_error("mpm_p2g: null grid");
    return MPM_ERR_ARG;
  }
  if (kind_override < 0 && (nmat <= 0 || nmat > MPM_PHASE_MAX_MATERIALS || !materials)) {
    set_error("mpm_p2g: per-particle materials need 1 <= nmat <= 16 and a table");
    return MPM_ERR_ARG;
  }
  mpm::phase_p2g((hipStream_t)stream, prm, p, kind_override, hardening, materials, nmat, grid_v,
                 grid_m, (unsigned long long*)oob_count);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

static int check_dense(const mpm_params* prm, const mpm_particles* p, const void* work, const char* fn) {
  int rc;
  if ((rc = check_params(prm, fn)) || (rc = check_particles(p, fn, true))) return rc;
  if (prm->dim != 3 || p->dtype != MPM_F32 || !work) {
    set_error(std::string(fn) + ": needs dim 3, FP32 particles and a workspace");
    return MPM_ERR_ARG;
  }
  return MPM_OK;
}

int32_t mpm_dense_workspace_bytes(const mpm_params* prm, int64_t* bytes) {
  int rc;
  if ((rc = check_params(prm, "mpm_dense_workspace_bytes"))) return rc;
  if (!bytes) {
    set_error("mpm_dense_workspace_bytes: null output");
    return MPM_ERR_ARG;
  }
  *bytes = (int64_t)(sizeof(float) * mpm::adjoint_workspace_floats(prm->res));
  return MPM_OK;
}

int32_t mpm_dense_substep(void* stream, const mpm_params* prm, mpm_particles* p, double hardening,
                          void* work, int64_t* oob_count) {
  int rc;
  if ((rc = check_dense(prm, p, work, "mpm_dense_substep"))) return rc;
  const int64_t r1 = prm->res + 1, nn = r1 * r1 * r1;
  float* gv = (float*)work + 4 * nn;
  mpm::dense_forward_grids((hipStream_t)stream, prm, p, hardening, (float*)work,
                           (unsigned long long*)oob_count);
  mpm::phase_g2p((hipStream_t)stream, prm, p, MPM_ELASTIC, 0.0, 0.0, nullptr, 0, gv,
                 (unsigned long long*)oob_count);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_adjoint_step(void* stream, const mpm_params* prm, const mpm_particles* p, double hardening,
                         void* work, float* x_bar, float* v_bar, float* F_bar, float* C_bar,
                         float* mu0_bar, float* lambda0_bar, int64_t* oob_count) {
  int rc;
  if ((rc = check_dense(prm, p, work, "mpm_adjoint_step"))) return rc;
  if (p->n > 0 && (!x_bar || !v_bar || !F_bar || !C_bar || !mu0_bar || !lambda0_bar)) {
    set_error("mpm_adjoint_step: null adjoint array");
    return MPM_ERR_ARG;
  }
  mpm::adjoint_step((hipStream_t)stream, prm, p, hardening, (float*)work, x_bar, v_bar, F_bar, C_bar,
                    mu0_bar, lambda0_bar, (unsigned long long*)oob_count);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_adjoint_step_grids(void* stream, const mpm_params* prm, const mpm_particles* p, double hardening,
                               const float* grids, void* work, float* x_bar, float* v_bar, float* F_bar,
                               float* C_bar, float* mu0_bar, float* lambda0_bar, int64_t* oob_count) {
  int rc;
  if ((rc = check_dense(prm, p, work, "mpm_adjoint_step_grids"))) return rc;
  if (!grids || (p->n > 0 && (!x_bar || !v_bar || !F_bar || !C_bar || !mu0_bar || !lambda0_bar))) {
    set_error("mpm_adjoint_step_grids: null grid or adjoint array");
    return MPM_ERR_ARG;
  }
  mpm::adjoint_step_grids((hipStream_t)stream, prm, p, hardening, grids, (float*)work, x_bar, v_bar, F_bar,
                          C_bar, mu0_bar, lambda0_bar, (unsigned long long*)oob_count);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_ctx_set_corrector(mpm_ctx* ctx, const float* params, float m_ref, float v_ref,
                              int32_t enable) {
  if (!ctx) {
    set_error("mpm_ctx_set_corrector: null context");
    return MPM_ERR_ARG;
  }
  return ctx->eng.set_corrector(params, m_ref, v_ref, enable);
}

int32_t mpm_corrector_sample(mpm_ctx* ctx, uint16_t* X, uint32_t* nodes, float* vel, int64_t max_rows,
                             int64_t* rows) {
  if (!ctx || (X && (!nodes || !vel || max_rows < 0))) {
    set_error("mpm_corrector_sample: bad argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.corrector_sample(X, nodes, vel, max_rows, rows);
}

int32_t mpm_grid_velocity(mpm_ctx* ctx, const uint32_t* nodes, int64_t count, float* out) {
  if (!ctx || count < 0 || (count > 0 && (!nodes || !out))) {
    set_error("mpm_grid_velocity: bad argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.grid_velocity(nodes, count, out);
}

int32_t mpm_mlp_sizes(int64_t* n_params, int64_t* n_packed) {
  if (!n_params || !n_packed) {
    set_error("mpm_mlp_sizes: null output");
    return MPM_ERR_ARG;
  }
  *n_params = mpm::P_COUNT;
  *n_packed = mpm::K_COUNT;
  return MPM_OK;
}

int32_t mpm_mlp_pack(void* stream, const float* params, uint16_t* packed) {
  if (!params || !packed) {
    set_error("mpm_mlp_pack: null buffer");
    return MPM_ERR_ARG;
  }
  mpm::mlp_pack((hipStream_t)stream, params, packed);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_mlp_forward(void* stream, const uint16_t* packed, const float* params, const uint16_t* X,
                        int64_t rows, uint16_t* H1, uint16_t* H2, float* Y) {
  if (rows < 0 || (rows > 0 && (!packed || !params || !X || !H1 || !H2 || !Y))) {
    set_error("mpm_mlp_forward: bad argument");
    return MPM_ERR_ARG;
  }
  mpm::mlp_forward((hipStream_t)stream, packed, params, X, rows, H1, H2, Y);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_mlp_loss(void* stream, const float* Y, const float* T, const float* w, int64_t rows,
                     float inv_norm, uint16_t* dY, double* loss) {
  if (rows < 0 || (rows > 0 && (!Y || !T || !dY || !loss))) {
    set_error("mpm_mlp_loss: bad argument");
    return MPM_ERR_ARG;
  }
  mpm::mlp_loss((hipStream_t)stream, Y, T, w, rows, inv_norm, dY, loss);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_mlp_backward(void* stream, const uint16_t* packed, const uint16_t* X, int64_t rows,
                         const uint16_t* H1, const uint16_t* H2, const uint16_t* dY, uint16_t* dH2,
                         uint16_t* dH1, float* grads) {
  if (rows < 0 || (rows > 0 && (!packed || !X || !H1 || !H2 || !dY || !dH2 || !dH1 || !grads))) {
    set_error("mpm_mlp_backward: bad argument");
    return MPM_ERR_ARG;
  }
  mpm::mlp_backward((hipStream_t)stream, packed, X, rows, H1, H2, dY, dH2, dH1, grads);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_mlp_adam(void* stream, float* params, const float* grads, float* m, float* v, int64_t n,
                     float lr, float beta1, float beta2, float eps, int32_t step) {
  if (!params || !grads || !m || !v || n <= 0 || n > (int64_t)1 << 30 || step < 1) {
    set_error("mpm_mlp_adam: bad argument");
    return MPM_ERR_ARG;
  }
  mpm::mlp_adam((hipStream_t)stream, params, grads, m, v, (int)n, lr, beta1, beta2, eps, step);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_grid_op(void* stream, const mpm_params* prm, int32_t dtype, void* grid_v,
                    const void* grid_m) {
  int rc;
  if ((rc = check_params(prm, "mpm_grid_op"))) return rc;
  if (!grid_v || !grid_m || (dtype != MPM_F32 && dtype != MPM_F64)) {
    set_error("mpm_grid_op: null grid or bad dtype");
    return MPM_ERR_ARG;
  }
  mpm::phase_grid_op((hipStream_t)stream, prm, dtype, grid_v, grid_m);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_collide_planes(void* stream, int32_t dtype, int32_t res, double dx, int32_t nplanes,
                           const double* points, const double* normals, int32_t quirks, void* grid_v) {
  if (!grid_v || (dtype != MPM_F32 && dtype != MPM_F64) || res < 1 || nplanes < 0 || nplanes > mpm::MAX_PLANES ||
      (nplanes > 0 && (!points || !normals)) || (quirks != MPM_QUIRKS_REFERENCE && quirks != MPM_QUIRKS_TEXTBOOK)) {
    set_error("mpm_collide_planes: null grid, bad dtype, res < 1, more than 16 planes or bad quirks mode");
    return MPM_ERR_ARG;
  }
  if (nplanes == 0) return MPM_OK;
  mpm::PlaneDev planes[mpm::MAX_PLANES];
  for (int q = 0; q < nplanes; ++q) {
    if (!mpm::make_plane(points + 3 * q, normals + 3 * q, dx, quirks, 0, planes[q])) {
      set_error("mpm_collide_planes: a normal is zero or not finite");
      return MPM_ERR_ARG;
    }
  }
  mpm::phase_planes((hipStream_t)stream, dtype, res, planes, nplanes, grid_v);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_ctx_set_planes(mpm_ctx* ctx, int32_t scene, int32_t nplanes, const double* points,
                           const double* normals, int32_t quirks) {
  if (!ctx) {
    set_error("mpm_ctx_set_planes: null context");
    return MPM_ERR_ARG;
  }
  return ctx->eng.set_planes(scene, nplanes, points, normals, quirks);
}

int32_t mpm_g2p(void* stream, const mpm_params* prm, mpm_particles* p, int32_t kind_override,
                double theta_c, double theta_s, const mpm_material* materials, int32_t nmat,
                const void* grid_v, int64_t* oob_count) {
  int rc;
  if ((rc = check_params(prm, "mpm_g2p")) || (rc = check_particles(p, "mpm_g2p", false))) return rc;
  if (!grid_v) {
    set_error("mpm_g2p: null grid");
    return MPM_ERR_ARG;
  }
  if (kind_override < 0 && (nmat <= 0 || nmat > MPM_PHASE_MAX_MATERIALS || !materials)) {
    set_error("mpm_g2p: per-particle materials need 1 <= nmat <= 16 and a table");
    return MPM_ERR_ARG;
  }
  mpm::phase_g2p((hipStream_t)stream, prm, p, kind_override, theta_c, theta_s, materials, nmat, grid_v,
                 (unsigned long long*)oob_count);
  CHECK_HIP(hipGetLastError());
  return MPM_OK;
}

int32_t mpm_ctx_create(mpm_ctx** out, int64_t n_max, int32_t nscenes, const mpm_scene* scenes,
                       int32_t nmat, const mpm_material* materials, int32_t flags) {
  (void)flags;
  if (!out) {
    set_error("mpm_ctx_create: null output handle");
    return MPM_ERR_ARG;
  }
  *out = nullptr;
  mpm_ctx* c = new mpm_ctx();
  const int rc = c->eng.init(n_max, nscenes, scenes, nmat, materials);
  if (rc != MPM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return MPM_OK;
}

int32_t mpm_ctx_destroy(mpm_ctx* ctx) {
  if (!ctx) return MPM_OK;
  hipStreamSynchronize(ctx->eng.own_stream);
  delete ctx;
  return MPM_OK;
}

int32_t mpm_ctx_set_stream(mpm_ctx* ctx, void* stream) {
  if (!ctx) {
    set_error("mpm_ctx_set_stream: null context");
    return MPM_ERR_ARG;
  }
  ctx->eng.user_stream = (hipStream_t)stream;
  return MPM_OK;
}

int32_t mpm_ctx_tune(mpm_ctx* ctx, int32_t chunk, int32_t sort_every, int32_t grid_mult,
                     int32_t use_graph) {
  if (!ctx || (chunk != 256 && chunk != 512 && chunk != 1024) || sort_every < 1 || grid_mult < 1 ||
      grid_mult > 64) {
    set_error("mpm_ctx_tune: chunk in {256,512,1024}, sort_every >= 1, 1 <= grid_mult <= 64");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  if (e.chunk != chunk || e.grid_mult != grid_mult) {
    e.drop_graphs();
    e.since_bin = -1;  // the chunk list was built for the old chunk size
  }
  e.chunk = chunk;
  e.sort_every = sort_every;
  e.grid_mult = grid_mult;
  e.use_graph = use_graph != 0;
  return MPM_OK;
}

int32_t mpm_ctx_set_spread(mpm_ctx* ctx, int32_t spread) {
  if (!ctx || (spread != 3 && spread != 9)) {
    set_error("mpm_ctx_set_spread: spread must be 3 or 9");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  if (e.spread != spread) e.drop_graphs();  // captured graphs hold the other kernel
  e.spread = spread;
  return MPM_OK;
}

int32_t mpm_ctx_kernel_timing(mpm_ctx* ctx, int32_t enable) {
  if (!ctx) {
    set_error("mpm_ctx_kernel_timing: null context");
    return MPM_ERR_ARG;
  }
  ctx->eng.timing = enable != 0;
  return MPM_OK;
}

int32_t mpm_ctx_kernel_times(mpm_ctx* ctx, double* ms, int64_t* counts) {
  if (!ctx || !ms || !counts) {
    set_error("mpm_ctx_kernel_times: null argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.kernel_times(ms, counts);
}

int32_t mpm_state_upload(mpm_ctx* ctx, const mpm_particles* p, const int32_t* scene_ids) {
  if (!ctx) {
    set_error("mpm_state_upload: null context");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  CHECK_HIP(hipEventRecord(e.ev_in, e.user_stream));
  CHECK_HIP(hipStreamWaitEvent(e.own_stream, e.ev_in, 0));
  const int rc = e.upload(p, scene_ids);
  if (rc) return rc;
  CHECK_HIP(hipEventRecord(e.ev_out, e.own_stream));
  CHECK_HIP(hipStreamWaitEvent(e.user_stream, e.ev_out, 0));
  return MPM_OK;
}

int32_t mpm_state_download(mpm_ctx* ctx, mpm_particles* p) {
  if (!ctx) {
    set_error("mpm_state_download: null context");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  CHECK_HIP(hipEventRecord(e.ev_in, e.user_stream));
  CHECK_HIP(hipStreamWaitEvent(e.own_stream, e.ev_in, 0));
  const int rc = e.download(p);
  if (rc) return rc;
  CHECK_HIP(hipEventRecord(e.ev_out, e.own_stream));
  CHECK_HIP(hipStreamWaitEvent(e.user_stream, e.ev_out, 0));
  return MPM_OK;
}

int32_t mpm_run(mpm_ctx* ctx, int32_t nsteps, double* sums) {
  if (!ctx) {
    set_error("mpm_run: null context");
    return MPM_ERR_ARG;
  }
  return ctx->eng.run(nsteps, sums);
}

int32_t mpm_run_frames(mpm_ctx* ctx, int32_t nsteps, float* frames) {
  if (!ctx || !frames) {
    set_error("mpm_run_frames: null argument");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  return e.run(nsteps, nullptr, frames);
}

int32_t mpm_run_states(mpm_ctx* ctx, int32_t nsteps, float* x, float* v, float* F, float* C, float* Jp) {
  if (!ctx || (nsteps > 0 && ctx->eng.n > 0 && (!x || !v || !F || !C || !Jp))) {
    set_error("mpm_run_states: null argument");
    return MPM_ERR_ARG;
  }
  const mpm::StateSink sink{x, v, F, C, Jp, nullptr};
  return ctx->eng.run(nsteps, nullptr, nullptr, &sink);
}

int32_t mpm_run_states_grids(mpm_ctx* ctx, int32_t nsteps, float* x, float* v, float* F, float* C, float* Jp,
                             float* grids) {
  if (!ctx || (nsteps > 0 && ctx->eng.n > 0 && (!x || !v || !F || !C || !Jp)) || (nsteps > 0 && !grids)) {
    set_error("mpm_run_states_grids: null argument");
    return MPM_ERR_ARG;
  }
  if (ctx->eng.scenes_host.size() != 1 || ctx->eng.corr_on) {
    set_error("mpm_run_states_grids: needs a single-scene context without a corrector");
    return MPM_ERR_ARG;
  }
  const mpm::StateSink sink{x, v, F, C, Jp, grids};
  return ctx->eng.run(nsteps, nullptr, nullptr, &sink);
}

int32_t mpm_state_positions(mpm_ctx* ctx, float* x) {
  if (!ctx) {
    set_error("mpm_state_positions: null context");
    return MPM_ERR_ARG;
  }
  mpm::Engine& e = ctx->eng;
  CHECK_HIP(hipEventRecord(e.ev_in, e.user_stream));
  CHECK_HIP(hipStreamWaitEvent(e.own_stream, e.ev_in, 0));
  const int rc = e.positions(x);
  if (rc) return rc;
  CHECK_HIP(hipEventRecord(e.ev_out, e.own_stream));
  CHECK_HIP(hipStreamWaitEvent(e.user_stream, e.ev_out, 0));
  return MPM_OK;
}

int32_t mpm_substep(mpm_ctx* ctx, mpm_particles* p, const int32_t* scene_ids, int32_t nsteps) {
  int rc;
  if ((rc = mpm_state_upload(ctx, p, scene_ids))) return rc;
  if ((rc = mpm_run(ctx, nsteps, nullptr))) return rc;
  if ((rc = mpm_state_download(ctx, p))) return rc;
  int64_t oob = 0;
  if ((rc = ctx->eng.take_counter(0, &oob))) return rc;
  if (oob >= (int64_t(1) << 40)) {
    set_error("mpm_substep: the FP64 pass for deferred particles did not finish in time");
    return MPM_ERR_HIP;
  }
  if (oob) {
    set_error("mpm_substep: " + std::to_string(oob) + " particle stencils left the grid");
    return MPM_ERR_OOB;
  }
  return MPM_OK;
}

int32_t mpm_ctx_oob_count(mpm_ctx* ctx, int64_t* out) {
  if (!ctx || !out) {
    set_error("mpm_ctx_oob_count: null argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.take_counter(0, out);
}

int32_t mpm_ctx_fallback_count(mpm_ctx* ctx, int64_t* out) {
  if (!ctx || !out) {
    set_error("mpm_ctx_fallback_count: null argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.take_counter(1, out);
}

int32_t mpm_ctx_active_blocks(mpm_ctx* ctx, int64_t* out) {
  if (!ctx || !out) {
    set_error("mpm_ctx_active_blocks: null argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.active_blocks(out);
}

int32_t mpm_ctx_request_rebin(mpm_ctx* ctx) {
  if (!ctx) {
    set_error("mpm_ctx_request_rebin: null context");
    return MPM_ERR_ARG;
  }
  ctx->eng.since_bin = -1;
  return MPM_OK;
}

int32_t mpm_particle_sums(mpm_ctx* ctx, double* out) {
  if (!ctx || !out) {
    set_error("mpm_particle_sums: null argument");
    return MPM_ERR_ARG;
  }
  return ctx->eng.particle_sums(out);
}

int32_t mpm_bin(mpm_ctx* ctx, int32_t dim, int32_t res, double inv_dx, int32_t layout, const float* x,
                int64_t n, uint32_t* keys, uint32_t* perm) {
  (void)ctx;
  if ((dim != 2 && dim != 3) || res < 1 || res > 4096 || n < 0 || (n > 0 && (!x || !keys || !perm)) ||
      (layout != MPM_KEY_BLOCKED && layout != MPM_KEY_ROWMAJOR)) {
    set_error("mpm_bin: bad argument");
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
    set_error("mpm_bin: key space exceeds 32 bits");
    return MPM_ERR_ARG;
  }
  hipStream_t st = nullptr;
  mpm::bin_keys(x, n, dim, res, (float)inv_dx, layout, keys, st);
  uint32_t* tmp = nullptr;
  void* ws = nullptr;
  CHECK_HIP(hipMalloc(&tmp, sizeof(uint32_t) * n));
  if (hipMalloc(&ws, mpm::radix_workspace_bytes(n)) != hipSuccess) {
    hipFree(tmp);
    set_error("mpm_bin: workspace allocation failed");
    return MPM_ERR_HIP;
  }
  hipMemcpyAsync(tmp, keys, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
  uint32_t *ks, *vs;
  mpm::radix_sort(tmp, perm, n, (uint32_t)max_key, ws, true, &ks, &vs, st);
  if (vs != perm) hipMemcpyAsync(perm, vs, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
  hipError_t e = hipStreamSynchronize(st);
  hipFree(tmp);
  hipFree(ws);
  if (e != hipSuccess) {
    set_error(std::string("mpm_bin: ") + hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

}  // extern "C"
