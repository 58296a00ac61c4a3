/*
 * sf_oracle.h -- C API of the host-loop oracle for the stable-fluids 3D step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * library (libstable-fluids.so) never links, calls or falls back to it.
 *
 * The oracle restates the reference step (backend-cuda.cu:58-434 kernels,
 * :608-675 step order) as plain host loops, templated on the arithmetic type:
 * precision 64 runs every operation in double (the parity oracle), precision 32 runs
 * the same operations in float (used to separate FP32 rounding from logic errors).
 * Fields cross the API as float (inputs, like the reference ABI) or double (outputs).
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>
#include "../include/stable-fluids-3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* smoother selection for sfo_create */
#define SFO_SMOOTHER_RBGS   0 /* red-black Gauss-Seidel, the reference smoother */
#define SFO_SMOOTHER_JACOBI 1 /* weighted-free Jacobi with ping-pong buffers */

/* ---- whole-simulation API (mirrors the C ABI of stable-fluids-3d.h) ---- */
void* sfo_create(const StableFluidsContextCreateDesc* desc, int precision, int smoother);
void sfo_destroy(void* sim);
int sfo_update_scalar_field(void* sim, uint32_t handle, const float* values);
int sfo_update_scalar_field_source(void* sim, uint32_t handle, const float* values);
int sfo_update_vector_field(void* sim, uint32_t handle, const float* x, const float* y, const float* z);
int sfo_set_flow_velocity(void* sim, const float* x, const float* y, const float* z);
void sfo_step(void* sim, int steps);
/* kind = StableFluidsViewKind; out1/out2 used for the SOA3 kinds. */
int sfo_get_field(void* sim, uint32_t kind, uint32_t handle, double* out0, double* out1, double* out2);
int sfo_threads(void);
void sfo_set_threads(int n); /* threads of the following parallel loops (n >= 1) */

/* ---- per-stage entry points (FP64), used by the oracle's own pinning tests ---- */
double sfo_load_scalar(const double* field, int x, int y, int z, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig* bc);
double sfo_load_pressure(const double* field, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc);
double sfo_load_velocity(const double* field, int component, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc);
double sfo_sample_scalar(const double* field, double x, double y, double z, int nx, int ny, int nz, double h, const StableFluidsScalarBoundaryConfig* bc);
double sfo_sample_velocity(const double* field, int component, double x, double y, double z, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc);
void sfo_trace_rk2(const double* vx, const double* vy, const double* vz, double x, double y, double z, double dt, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc, double* out3);
void sfo_advect_velocity(double* out, const double* src, int component, const double* vx, const double* vy, const double* vz, double dt, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc);
void sfo_advect_scalar(double* out, const double* src, const double* vx, const double* vy, const double* vz, double dt, int nx, int ny, int nz, double h, const StableFluidsScalarBoundaryConfig* sbc, const StableFluidsFlowBoundaryConfig* fbc);
/* iterations of the chosen smoother, in place on x */
void sfo_smooth_scalar(double* x, const double* b, double alpha, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig* bc);
void sfo_smooth_velocity(double* x, const double* b, double alpha, int component, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc);
void sfo_smooth_pressure(double* p, const double* div, double h2, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc);
void sfo_divergence(double* div, const double* vx, const double* vy, const double* vz, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc);
void sfo_project(double* vx, double* vy, double* vz, const double* p, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc);

#ifdef __cplusplus
}
#endif

#endif /* SF_ORACLE_H */
