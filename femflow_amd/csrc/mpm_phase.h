// Host-side launchers of the dense-grid phase kernels (mpm_phase.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm355.h"
#include "mpm_math.h"

#define MPM_PHASE_MAX_MATERIALS 16

namespace mpm {

void phase_p2g(hipStream_t st, const mpm_params* prm, const mpm_particles* p, int kind_override,
               double hardening, const mpm_material* mats, int nmat, void* gv, void* gm,
               unsigned long long* oob);
void phase_grid_op(hipStream_t st, const mpm_params* prm, int dtype, void* gv, const void* gm);
void phase_planes(hipStream_t st, int dtype, int res, const PlaneDev* planes, int n, void* gv);
void phase_g2p(hipStream_t st, const mpm_params* prm, mpm_particles* p, int kind_override,
               double theta_c, double theta_s, const mpm_material* mats, int nmat, const void* gv,
               unsigned long long* oob);

size_t adjoint_workspace_floats(int res);
void dense_forward_grids(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                         float* work, unsigned long long* oob);
void adjoint_step(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                  float* work, float* xb, float* vb, float* Fb, float* Cb, float* mub, float* lamb,
                  unsigned long long* oob);
void adjoint_step_grids(hipStream_t st, const mpm_params* prm, const mpm_particles* p, double hardening,
                        const float* grids, float* work, float* xb, float* vb, float* Fb, float* Cb, float* mub,
                        float* lamb, unsigned long long* oob);

}  // namespace mpm
