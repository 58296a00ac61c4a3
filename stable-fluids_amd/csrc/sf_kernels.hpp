// sf_kernels.hpp -- host-side launchers of the gfx950 stage kernels (sf_kernels.hip).
// Every launcher enqueues asynchronously on `s` and returns hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sf_device.hpp"

namespace sf {

enum SmootherKind : int {
    kSmoothRBGS   = 0, // red ((x+y+z)&1 == 0) then black, fused into one ping-pong pass
    kSmoothJacobi = 1, // every cell from the previous iterate
};

enum SolveKind : int {
    kSolveDiffusion = 0, // x = (b + coef*sum) / (1 + 6*coef), coef = dt*diffusivity/h^2
    kSolvePoisson   = 1, // x = (sum - coef*b) / 6, coef = h^2
};

// One smoother pass: reads x_old (or zeros when x_old == nullptr) and b, writes x_new.
// out_scale multiplies every written value (dissipation fused into the last pass).
struct SmoothPass {
    const float* x_old;
    const float* b;
    float* x_new;
    float coef;
    float out_scale;
};

hipError_t launch_fill(hipStream_t s, float* p, float value, uint64_t n);
hipError_t launch_copy_scaled(hipStream_t s, float* dst, const float* src, float factor, uint64_t n);
hipError_t launch_add_scaled(hipStream_t s, float* dst, const float* src, float dt, uint64_t n);
hipError_t launch_add_force(hipStream_t s, float* vx, float* vy, float* vz, const float* fx, const float* fy, const float* fz, float dt, uint64_t n);
hipError_t launch_advect_velocity(hipStream_t s, float* tx, float* ty, float* tz, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsFlowBoundaryConfig& bc);
hipError_t launch_advect_scalar(hipStream_t s, float* out, const float* src, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc);
hipError_t launch_smooth(hipStream_t s, SmootherKind smoother, SolveKind solve, const SmoothPass& pass, Dims d, const StencilBC& bc);
// divergence of (vx,vy,vz); bc.face[2c+side] is the rule of component c on axis c.
hipError_t launch_divergence(hipStream_t s, float* div, const float* vx, const float* vy, const float* vz, Dims d, float h, const StencilBC& div_bc);
// v_out = v_in - grad p; div and |v| of the projected field, all in one pass.
hipError_t launch_project(hipStream_t s, float* ox, float* oy, float* oz, float* div, float* mag, const float* vx, const float* vy, const float* vz, const float* p, Dims d, float h, const StencilBC& p_bc, const StencilBC& div_bc);

} // namespace sf
