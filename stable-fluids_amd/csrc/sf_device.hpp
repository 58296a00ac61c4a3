// sf_device.hpp -- device-side types and boundary rules shared by the gfx950 kernels.
//
// Two forms of the reference's ghost rules (backend-cuda.cu:62-189) live here:
//  * StencilBC: for 7-point stencils. A face neighbour outside the grid on a non-periodic
//    axis always resolves to a function of the cell itself (the interior value at the
//    clamped position is the centre cell), so each face reduces to one of three rules:
//    copy the centre, reflect (2*v_face - centre) or a constant. The host derives the rule
//    per face and per field (scalar / pressure / velocity component).
//  * The full rules with axis precedence (x, then y, then z) and clamping of every axis,
//    used by trilinear sampling where corner taps can leave the grid on several axes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/stable-fluids-3d.h"

namespace sf {

struct Dims {
    int nx, ny, nz;
};

enum GhostMode : int {
    kGhostCopy    = 0, // ghost = centre value (zero-flux scalar, wall pressure, outflow / tangential velocity)
    kGhostReflect = 1, // ghost = value - centre, value = 2 * wall velocity component
    kGhostConst   = 2, // ghost = value (fixed-value scalar, outflow pressure)
};

struct FaceRule {
    int mode;
    float value;
};

// face index = 2 * axis + (plus side ? 1 : 0); bit a of periodic set when axis a wraps
struct StencilBC {
    FaceRule face[6];
    int periodic;
};

__device__ __forceinline__ float apply_ghost(const FaceRule r, const float centre) {
    return r.mode == kGhostCopy ? centre : (r.mode == kGhostReflect ? r.value - centre : r.value);
}

__device__ __forceinline__ uint64_t lin(const int x, const int y, const int z, const Dims d) {
    return (static_cast<uint64_t>(z) * static_cast<uint64_t>(d.ny) + static_cast<uint64_t>(y)) * static_cast<uint64_t>(d.nx) + static_cast<uint64_t>(x);
}

__device__ __forceinline__ int wrap_index(int c, const int n) {
    c %= n;
    return c < 0 ? c + n : c;
}

// Moves coordinate `c` (extent n) by dir; returns false when the neighbour is a ghost
// (outside on a non-periodic axis); wraps it on a periodic axis.
__device__ __forceinline__ bool step_axis(int& c, const int dir, const int n, const bool periodic) {
    c += dir;
    if (c >= 0 && c < n) return true;
    if (!periodic) return false;
    c = wrap_index(c, n);
    return true;
}

// ---- full reference rules (axis precedence), used by the samplers ----

__device__ __forceinline__ const StableFluidsFlowBoundaryFaceDesc& flow_face(const StableFluidsFlowBoundaryConfig& b, const int axis, const bool minus) {
    return axis == 0 ? (minus ? b.x_minus : b.x_plus) : axis == 1 ? (minus ? b.y_minus : b.y_plus) : (minus ? b.z_minus : b.z_plus);
}
__device__ __forceinline__ const StableFluidsScalarBoundaryFaceDesc& scalar_face(const StableFluidsScalarBoundaryConfig& b, const int axis, const bool minus) {
    return axis == 0 ? (minus ? b.x_minus : b.x_plus) : axis == 1 ? (minus ? b.y_minus : b.y_plus) : (minus ? b.z_minus : b.z_plus);
}
__device__ __forceinline__ bool flow_periodic(const StableFluidsFlowBoundaryConfig& b, const int a) {
    return flow_face(b, a, true).type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC && flow_face(b, a, false).type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC;
}
__device__ __forceinline__ bool scalar_periodic(const StableFluidsScalarBoundaryConfig& b, const int a) {
    return scalar_face(b, a, true).type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC && scalar_face(b, a, false).type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC;
}

// Scalar tap (reference backend-cuda.cu:62-94).
__device__ inline float tap_scalar(const float* __restrict__ f, int x, int y, int z, const Dims d, const StableFluidsScalarBoundaryConfig& bc) {
    if (x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz) return f[lin(x, y, z, d)];
    int c[3]       = {x, y, z};
    const int n[3] = {d.nx, d.ny, d.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < n[a]) continue;
        const auto& face = scalar_face(bc, a, c[a] < 0);
        if (scalar_periodic(bc, a))
            c[a] = wrap_index(c[a], n[a]);
        else if (face.type == STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX)
            c[a] = c[a] < 0 ? 0 : n[a] - 1;
        else
            return face.value;
    }
    return f[lin(c[0], c[1], c[2], d)];
}

// Velocity-component tap (reference backend-cuda.cu:130-189).
__device__ inline float tap_velocity(const float* __restrict__ f, const int comp, int x, int y, int z, const Dims d, const StableFluidsFlowBoundaryConfig& bc) {
    if (x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz) return f[lin(x, y, z, d)];
    int c[3]       = {x, y, z};
    const int n[3] = {d.nx, d.ny, d.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < n[a]) continue;
        const auto& face = flow_face(bc, a, c[a] < 0);
        if (flow_periodic(bc, a)) {
            c[a] = wrap_index(c[a], n[a]);
            continue;
        }
        const int qx         = min(max(c[0], 0), n[0] - 1);
        const int qy         = min(max(c[1], 0), n[1] - 1);
        const int qz         = min(max(c[2], 0), n[2] - 1);
        const float interior = f[lin(qx, qy, qz, d)];
        if (face.type == STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW) return interior;
        if (face.type == STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL && comp != a) return interior;
        const float prescribed = comp == 0 ? face.velocity_x : (comp == 1 ? face.velocity_y : face.velocity_z);
        return 2.0f * prescribed - interior;
    }
    return f[lin(c[0], c[1], c[2], d)];
}

__device__ __forceinline__ float wrap_position(float p, const int n, const float h) {
    const float extent = static_cast<float>(n) * h;
    if (extent <= 0.0f) return 0.0f;
    p = fmodf(p, extent);
    if (p < 0.0f) p += extent;
    return p;
}

// Trilinear interpolation at world position (x,y,z) through `tap`, lerping x, y, z in
// turn (reference backend-cuda.cu:208-238).
template <class Tap>
__device__ __forceinline__ float trilinear(const float x, const float y, const float z, const float h, Tap tap) {
    const float gx = x / h - 0.5f;
    const float gy = y / h - 0.5f;
    const float gz = z / h - 0.5f;
    const int x0   = static_cast<int>(floorf(gx));
    const int y0   = static_cast<int>(floorf(gy));
    const int z0   = static_cast<int>(floorf(gz));
    const float tx = gx - static_cast<float>(x0);
    const float ty = gy - static_cast<float>(y0);
    const float tz = gz - static_cast<float>(z0);
    const float c000 = tap(x0, y0, z0), c100 = tap(x0 + 1, y0, z0);
    const float c010 = tap(x0, y0 + 1, z0), c110 = tap(x0 + 1, y0 + 1, z0);
    const float c001 = tap(x0, y0, z0 + 1), c101 = tap(x0 + 1, y0, z0 + 1);
    const float c011 = tap(x0, y0 + 1, z0 + 1), c111 = tap(x0 + 1, y0 + 1, z0 + 1);
    const float c00 = c000 + (c100 - c000) * tx;
    const float c10 = c010 + (c110 - c010) * tx;
    const float c01 = c001 + (c101 - c001) * tx;
    const float c11 = c011 + (c111 - c011) * tx;
    const float c0  = c00 + (c10 - c00) * ty;
    const float c1  = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

// reference backend-cuda.cu:241-289
__device__ inline float sample_velocity(const float* __restrict__ f, const int comp, float x, float y, float z, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc) {
    if (flow_periodic(bc, 0)) x = wrap_position(x, d.nx, h);
    if (flow_periodic(bc, 1)) y = wrap_position(y, d.ny, h);
    if (flow_periodic(bc, 2)) z = wrap_position(z, d.nz, h);
    return trilinear(x, y, z, h, [&](int i, int j, int k) { return tap_velocity(f, comp, i, j, k, d, bc); });
}

// reference backend-cuda.cu:191-239
__device__ inline float sample_scalar(const float* __restrict__ f, float x, float y, float z, const Dims d, const float h, const StableFluidsScalarBoundaryConfig& bc) {
    if (scalar_periodic(bc, 0)) x = wrap_position(x, d.nx, h);
    if (scalar_periodic(bc, 1)) y = wrap_position(y, d.ny, h);
    if (scalar_periodic(bc, 2)) z = wrap_position(z, d.nz, h);
    return trilinear(x, y, z, h, [&](int i, int j, int k) { return tap_scalar(f, i, j, k, d, bc); });
}

// A trilinear sample position after the periodic wrap: lower corner, weights, and (when
// all 8 taps lie inside the grid) the 32-bit linear index of the lower corner. Positions
// are located once and shared by every field sampled there.
struct TriPos {
    int x0, y0, z0;
    float tx, ty, tz;
    bool interior;
    uint32_t base;
};

__device__ __forceinline__ TriPos locate(const float x, const float y, const float z, const Dims d, const float h) {
    TriPos p;
    const float gx = x / h - 0.5f;
    const float gy = y / h - 0.5f;
    const float gz = z / h - 0.5f;
    p.x0           = static_cast<int>(floorf(gx));
    p.y0           = static_cast<int>(floorf(gy));
    p.z0           = static_cast<int>(floorf(gz));
    p.tx           = gx - static_cast<float>(p.x0);
    p.ty           = gy - static_cast<float>(p.y0);
    p.tz           = gz - static_cast<float>(p.z0);
    p.interior     = p.x0 >= 0 && p.x0 + 1 < d.nx && p.y0 >= 0 && p.y0 + 1 < d.ny && p.z0 >= 0 && p.z0 + 1 < d.nz;
    p.base         = p.interior ? (static_cast<uint32_t>(p.z0) * static_cast<uint32_t>(d.ny) + static_cast<uint32_t>(p.y0)) * static_cast<uint32_t>(d.nx) + static_cast<uint32_t>(p.x0) : 0u;
    return p;
}

// Same value as trilinear() at the located position: interior taps come from the base
// index and constant strides, other taps from `tap` (the boundary rules).
template <class Tap>
__device__ __forceinline__ float sample_at(const float* __restrict__ f, const TriPos& p, const Dims d, Tap tap) {
    float c000, c100, c010, c110, c001, c101, c011, c111;
    if (p.interior) {
        const uint32_t nx = static_cast<uint32_t>(d.nx), pl = nx * static_cast<uint32_t>(d.ny);
        const float* q    = f + p.base;
        c000              = q[0];
        c100              = q[1];
        c010              = q[nx];
        c110              = q[nx + 1];
        c001              = q[pl];
        c101              = q[pl + 1];
        c011              = q[pl + nx];
        c111              = q[pl + nx + 1];
    } else {
        c000 = tap(p.x0, p.y0, p.z0);
        c100 = tap(p.x0 + 1, p.y0, p.z0);
        c010 = tap(p.x0, p.y0 + 1, p.z0);
        c110 = tap(p.x0 + 1, p.y0 + 1, p.z0);
        c001 = tap(p.x0, p.y0, p.z0 + 1);
        c101 = tap(p.x0 + 1, p.y0, p.z0 + 1);
        c011 = tap(p.x0, p.y0 + 1, p.z0 + 1);
        c111 = tap(p.x0 + 1, p.y0 + 1, p.z0 + 1);
    }
    const float c00 = c000 + (c100 - c000) * p.tx;
    const float c10 = c010 + (c110 - c010) * p.tx;
    const float c01 = c001 + (c101 - c001) * p.tx;
    const float c11 = c011 + (c111 - c011) * p.tx;
    const float c0  = c00 + (c10 - c00) * p.ty;
    const float c1  = c01 + (c11 - c01) * p.ty;
    return c0 + (c1 - c0) * p.tz;
}

// Velocity at world position (x,y,z) for all three components (reference sample_velocity,
// backend-cuda.cu:291-293): one wrap and one located position serve the three fields.
__device__ __forceinline__ void sample_velocity3(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, float x, float y, float z, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc, float& u, float& v, float& w) {
    if (flow_periodic(bc, 0)) x = wrap_position(x, d.nx, h);
    if (flow_periodic(bc, 1)) y = wrap_position(y, d.ny, h);
    if (flow_periodic(bc, 2)) z = wrap_position(z, d.nz, h);
    const TriPos p = locate(x, y, z, d, h);
    u              = sample_at(vx, p, d, [&](int i, int j, int k) { return tap_velocity(vx, 0, i, j, k, d, bc); });
    v              = sample_at(vy, p, d, [&](int i, int j, int k) { return tap_velocity(vy, 1, i, j, k, d, bc); });
    w              = sample_at(vz, p, d, [&](int i, int j, int k) { return tap_velocity(vz, 2, i, j, k, d, bc); });
}

// Midpoint backtrace from the centre of cell (cx,cy,cz), world position p (reference
// trace_particle_rk2, backend-cuda.cu:295-300). When the centre maps to grid coordinates
// exactly (h a power of two), every trilinear weight of the first sample is zero and the
// sample is the cell value itself, which is loaded directly.
__device__ __forceinline__ void trace_rk2_cell(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const int cx, const int cy, const int cz, const float px, const float py, const float pz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc, float& ox, float& oy, float& oz) {
    float u0, v0, w0;
    const bool exact = px / h - 0.5f == static_cast<float>(cx) && py / h - 0.5f == static_cast<float>(cy) && pz / h - 0.5f == static_cast<float>(cz);
    if (exact) {
        const uint64_t i = lin(cx, cy, cz, d);
        u0               = vx[i];
        v0               = vy[i];
        w0               = vz[i];
    } else {
        sample_velocity3(vx, vy, vz, px, py, pz, d, h, bc, u0, v0, w0);
    }
    const float mx = px - 0.5f * dt * u0;
    const float my = py - 0.5f * dt * v0;
    const float mz = pz - 0.5f * dt * w0;
    float u1, v1, w1;
    sample_velocity3(vx, vy, vz, mx, my, mz, d, h, bc, u1, v1, w1);
    ox = px - dt * u1;
    oy = py - dt * v1;
    oz = pz - dt * w1;
}

// Midpoint backtrace from world position p (reference backend-cuda.cu:291-300).
__device__ inline void trace_rk2(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float px, const float py, const float pz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc, float& ox, float& oy, float& oz) {
    const float u0 = sample_velocity(vx, 0, px, py, pz, d, h, bc);
    const float v0 = sample_velocity(vy, 1, px, py, pz, d, h, bc);
    const float w0 = sample_velocity(vz, 2, px, py, pz, d, h, bc);
    const float mx = px - 0.5f * dt * u0;
    const float my = py - 0.5f * dt * v0;
    const float mz = pz - 0.5f * dt * w0;
    const float u1 = sample_velocity(vx, 0, mx, my, mz, d, h, bc);
    const float v1 = sample_velocity(vy, 1, mx, my, mz, d, h, bc);
    const float w1 = sample_velocity(vz, 2, mx, my, mz, d, h, bc);
    ox = px - dt * u1;
    oy = py - dt * v1;
    oz = pz - dt * w1;
}

} // namespace sf
