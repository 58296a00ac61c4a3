// sf_advect.hip -- semi-Lagrangian advection through field buffer resources (gfx950).
//
// Same arithmetic as the reference (backend-cuda.cu:191-352): RK2 midpoint trace,
// trilinear samples g = p / h - 0.5, corner floor(g), lerps in x, y, z, boundary taps by
// the reference rules with their axis precedence. What differs is the mapping onto the
// GPU:
//  * a wave covers 16 x 4 cells (four 64-byte row segments), so only the waves on the
//    outermost columns and rows of the grid ever take boundary taps (a 64 x 1 row would
//    put half of all waves on a wall);
//  * interior taps are buffer loads off one per-sample byte offset with the row and
//    plane strides as scalar offsets, for all fields sampled at that position;
//  * boundary taps resolve the reference precedence per corner (clamp / wrap per axis,
//    first non-periodic outside axis picks the face) without branches, with the
//    velocity face rules as (a, c) pairs v -> fma(a, v, c) in a small LDS table;
//  * with h a power of two, p / h is the exact product p * (1 / h).
// Fields up to 2 GiB (the byte offsets are 32-bit); larger grids use the row kernels of
// sf_kernels.hip.
#include <cmath>

#include "sf_buffer.hpp"
#include "sf_kernels.hpp"

namespace sf {

namespace {

constexpr int kAX       = 16; // block: 16 x 16 cells, four waves of 16 x 4
constexpr int kAY       = 16;
constexpr int kAThreads = kAX * kAY;

__device__ __forceinline__ float ld(const Rsrc r, const int voff, const int soff) { return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0)); }

// Lower corner and weights of a trilinear sample (reference trilinear,
// backend-cuda.cu:208-238); off is the byte offset of the lower corner when all eight
// taps lie inside the grid.
struct Tri {
    int x0, y0, z0;
    float tx, ty, tz;
    bool interior;
    int off;
};

struct Grid {
    int nx, ny, nz;
    int sy, sz; // byte strides of a row and a plane
    float h, inv_h;
};

template <bool kPow2>
__device__ __forceinline__ float grid_coord(const float p, const Grid& g) {
    if constexpr (kPow2)
        return p * g.inv_h - 0.5f; // p / h exactly
    else
        return p / g.h - 0.5f;
}

template <bool kPow2>
__device__ __forceinline__ Tri locate_b(const float x, const float y, const float z, const Grid& g) {
    Tri p;
    const float gx = grid_coord<kPow2>(x, g);
    const float gy = grid_coord<kPow2>(y, g);
    const float gz = grid_coord<kPow2>(z, g);
    p.x0           = static_cast<int>(floorf(gx));
    p.y0           = static_cast<int>(floorf(gy));
    p.z0           = static_cast<int>(floorf(gz));
    p.tx           = gx - static_cast<float>(p.x0);
    p.ty           = gy - static_cast<float>(p.y0);
    p.tz           = gz - static_cast<float>(p.z0);
    p.interior     = p.x0 >= 0 && p.x0 + 1 < g.nx && p.y0 >= 0 && p.y0 + 1 < g.ny && p.z0 >= 0 && p.z0 + 1 < g.nz;
    p.off          = p.interior ? p.z0 * g.sz + p.y0 * g.sy + 4 * p.x0 : 0;
    return p;
}

// taps in the reference order 000, 100, 010, 110, 001, 101, 011, 111
__device__ __forceinline__ float lerp8(const float (&c)[8], const Tri& p) {
    const float c00 = c[0] + (c[1] - c[0]) * p.tx;
    const float c10 = c[2] + (c[3] - c[2]) * p.tx;
    const float c01 = c[4] + (c[5] - c[4]) * p.tx;
    const float c11 = c[6] + (c[7] - c[6]) * p.tx;
    const float c0  = c00 + (c10 - c00) * p.ty;
    const float c1  = c01 + (c11 - c01) * p.ty;
    return c0 + (c1 - c0) * p.tz;
}

__device__ __forceinline__ void taps_interior(const Rsrc r, const Tri& p, const Grid& g, float (&c)[8]) {
    c[0] = ld(r, p.off, 0);
    c[1] = ld(r, p.off + 4, 0);
    c[2] = ld(r, p.off, g.sy);
    c[3] = ld(r, p.off + 4, g.sy);
    c[4] = ld(r, p.off, g.sz);
    c[5] = ld(r, p.off + 4, g.sz);
    c[6] = ld(r, p.off, g.sy + g.sz);
    c[7] = ld(r, p.off + 4, g.sy + g.sz);
}

// One axis of a boundary sample: for corners i0, i0 + 1 the index clamped into the grid
// (q) and wrapped (w; on a periodic axis the wrapped position keeps corners in [-1, n]),
// whether the corner is outside on a non-periodic axis (o) and on the minus side (m).
struct Axis {
    int q[2], w[2];
    bool o[2], m[2];
};

__device__ __forceinline__ Axis axis_of(const int i0, const int n, const bool periodic) {
    Axis a;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i    = i0 + k;
        const bool out = i < 0 || i >= n;
        a.q[k]         = min(max(i, 0), n - 1);
        a.w[k]         = i < 0 ? i + n : (i >= n ? i - n : i);
        a.o[k]         = out && !periodic;
        a.m[k]         = i < 0;
    }
    return a;
}

struct Periodic {
    bool x, y, z;
};

// Velocity face rules of component c at face f (2 axis + plus side; 6, 7 = no face) as
// (a, c): outflow and free slip across the other axes keep the interior value (1, -0),
// walls give 2 u_wall - interior (-1, 2 u_wall), reference backend-cuda.cu:170-186.
__device__ __forceinline__ void fill_velocity_rules(float2 (*rules)[8], const StableFluidsFlowBoundaryConfig& bc, const int t) {
    if (t >= 24) return;
    const int comp = t >> 3, f = t & 7;
    float2 r       = make_float2(1.0f, -0.0f);
    if (f < 6) {
        const int axis = f >> 1;
        const auto& fd = axis == 0 ? (f & 1 ? bc.x_plus : bc.x_minus) : axis == 1 ? (f & 1 ? bc.y_plus : bc.y_minus) : (f & 1 ? bc.z_plus : bc.z_minus);
        const bool keep = fd.type == STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW || (fd.type == STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL && comp != axis);
        if (!keep) {
            const float prescribed = comp == 0 ? fd.velocity_x : (comp == 1 ? fd.velocity_y : fd.velocity_z);
            r                      = make_float2(-1.0f, 2.0f * prescribed);
        }
    }
    rules[comp][f] = r;
}

struct VelocityField {
    Rsrc r[3];
};

// The three velocity components at a located position (reference sample_velocity,
// backend-cuda.cu:241-289, with tap_velocity :130-189 for boundary taps).
__device__ __forceinline__ void sample_velocity_b(const VelocityField& v, const Tri& p, const Grid& g, const Periodic& per, const float2 (*rules)[8], float (&out)[3]) {
    float c[3][8];
    if (p.interior) {
#pragma unroll
        for (int k = 0; k < 3; ++k) taps_interior(v.r[k], p, g, c[k]);
    } else {
        const Axis X = axis_of(p.x0, g.nx, per.x), Y = axis_of(p.y0, g.ny, per.y), Z = axis_of(p.z0, g.nz, per.z);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int cx = t & 1, cy = (t >> 1) & 1, cz = t >> 2;
            const bool ox = X.o[cx], oy = Y.o[cy], oz = Z.o[cz];
            // first non-periodic outside axis: its clamp of every coordinate and its face
            const int face = ox ? (X.m[cx] ? 0 : 1) : oy ? (Y.m[cy] ? 2 : 3) : oz ? (Z.m[cz] ? 4 : 5) : 6;
            const int ix   = per.x ? X.w[cx] : X.q[cx];
            const int iy   = (per.y && !ox) ? Y.w[cy] : Y.q[cy];
            const int iz   = (per.z && !ox && !oy) ? Z.w[cz] : Z.q[cz];
            const int off  = iz * g.sz + iy * g.sy + 4 * ix;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float2 r = rules[k][face];
                c[k][t]        = __builtin_fmaf(r.x, ld(v.r[k], off, 0), r.y);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = lerp8(c[k], p);
}

__device__ __forceinline__ float wrap_pos(float q, const int n, const float h) {
    const float extent = static_cast<float>(n) * h;
    if (extent <= 0.0f) return 0.0f;
    q = fmodf(q, extent);
    if (q < 0.0f) q += extent;
    return q;
}

// Velocity at world position q, flow-periodic axes wrapped first.
template <bool kPow2>
__device__ __forceinline__ void velocity_at(const VelocityField& v, float qx, float qy, float qz, const Grid& g, const Periodic& per, const float2 (*rules)[8], float (&out)[3]) {
    if (per.x) qx = wrap_pos(qx, g.nx, g.h);
    if (per.y) qy = wrap_pos(qy, g.ny, g.h);
    if (per.z) qz = wrap_pos(qz, g.nz, g.h);
    sample_velocity_b(v, locate_b<kPow2>(qx, qy, qz, g), g, per, rules, out);
}

// Midpoint backtrace from the centre of cell (x, y, z) (reference trace_particle_rk2,
// backend-cuda.cu:295-300). When the centre maps to grid coordinates exactly (always for
// h a power of two) the first sample is the cell value itself.
template <bool kPow2>
__device__ __forceinline__ void trace_cell(const VelocityField& v, const int x, const int y, const int z, const int cell_off, const float dt, const Grid& g, const Periodic& per, const float2 (*rules)[8], float& ox, float& oy, float& oz) {
    const float px = (static_cast<float>(x) + 0.5f) * g.h;
    const float py = (static_cast<float>(y) + 0.5f) * g.h;
    const float pz = (static_cast<float>(z) + 0.5f) * g.h;
    float u0[3];
    bool exact = true;
    if constexpr (!kPow2) exact = px / g.h - 0.5f == static_cast<float>(x) && py / g.h - 0.5f == static_cast<float>(y) && pz / g.h - 0.5f == static_cast<float>(z);
    if (exact) {
#pragma unroll
        for (int k = 0; k < 3; ++k) u0[k] = ld(v.r[k], cell_off, 0);
    } else {
        velocity_at<kPow2>(v, px, py, pz, g, per, rules, u0);
    }
    const float mx = px - 0.5f * dt * u0[0];
    const float my = py - 0.5f * dt * u0[1];
    const float mz = pz - 0.5f * dt * u0[2];
    float u1[3];
    velocity_at<kPow2>(v, mx, my, mz, g, per, rules, u1);
    ox = px - dt * u1[0];
    oy = py - dt * u1[1];
    oz = pz - dt * u1[2];
}

__device__ __forceinline__ Grid grid_of(const Dims d, const float h, const float inv_h) { return Grid{d.nx, d.ny, d.nz, 4 * d.nx, 4 * d.nx * d.ny, h, inv_h}; }

// Blocks run on a 1D grid renumbered per XCD: workgroups are dealt round-robin over the 8
// XCDs, so XCD k receives the contiguous block range [k n / 8, (k + 1) n / 8) in x, y, z
// order, and the tap halos shared by neighbouring blocks meet in the same L2.
// Cell of this thread in 16 x 16 block b of d (blocks in x, y, z order; each wave a 16 x 4
// patch); false past the x or y extent.
__device__ __forceinline__ bool cell_in_block(const Dims d, const int b, int& x, int& y, int& z) {
    const int bx_n = (d.nx + kAX - 1) / kAX, by_n = (d.ny + kAY - 1) / kAY;
    const int bx = b % bx_n, by = (b / bx_n) % by_n;
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    x              = bx * kAX + (lane & 15);
    y              = by * kAY + 4 * wave + (lane >> 4);
    z              = b / (bx_n * by_n);
    return x < d.nx && y < d.ny;
}

__device__ __forceinline__ bool cell_of(const Dims d, int& x, int& y, int& z) {
    const int n = static_cast<int>(gridDim.x);
    const int w = static_cast<int>(blockIdx.x);
    return cell_in_block(d, (n & 7) == 0 ? (w & 7) * (n >> 3) + (w >> 3) : w, x, y, z);
}

// One midpoint trace per cell, three component samples (reference
// advect_velocity_component_kernel backend-cuda.cu:330-340, launched per component).
template <bool kPow2>
__device__ __forceinline__ void advect_velocity_cell(float* __restrict__ tx, float* __restrict__ ty, float* __restrict__ tz, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const float inv_h, const StableFluidsFlowBoundaryConfig& bc, const float2 (*rules)[8], const int x, const int y, const int z) {
    const uint32_t bytes = 4u * static_cast<uint32_t>(d.nx) * static_cast<uint32_t>(d.ny) * static_cast<uint32_t>(d.nz);
    const VelocityField v{{plane_rsrc(vx, bytes), plane_rsrc(vy, bytes), plane_rsrc(vz, bytes)}};
    const Grid g         = grid_of(d, h, inv_h);
    const Periodic per   = {flow_periodic(bc, 0), flow_periodic(bc, 1), flow_periodic(bc, 2)};
    const int cell_off   = z * g.sz + y * g.sy + 4 * x;
    float qx, qy, qz;
    trace_cell<kPow2>(v, x, y, z, cell_off, dt, g, per, rules, qx, qy, qz);
    float u[3];
    velocity_at<kPow2>(v, qx, qy, qz, g, per, rules, u);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(u[0]), plane_rsrc(tx, bytes), cell_off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(u[1]), plane_rsrc(ty, bytes), cell_off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(u[2]), plane_rsrc(tz, bytes), cell_off, 0, 0);
}

// Six waves per SIMD (at most 80 VGPRs; 81 had allocated 88, five waves): the compiler
// spills two registers inside the boundary-tap path only, and the gather-latency-bound
// interior runs 338.8 -> 332.3 us per velocity advection at 256^3.
constexpr int kAdvectWaves = 6;

template <bool kPow2>
__global__ __launch_bounds__(kAThreads) __attribute__((amdgpu_waves_per_eu(kAdvectWaves))) void k_advect_velocity_b(float* __restrict__ tx, float* __restrict__ ty, float* __restrict__ tz, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const float inv_h, const StableFluidsFlowBoundaryConfig bc) {
    __shared__ float2 rules[3][8];
    fill_velocity_rules(rules, bc, static_cast<int>(threadIdx.x));
    __syncthreads();
    int x, y, z;
    if (!cell_of(d, x, y, z)) return;
    advect_velocity_cell<kPow2>(tx, ty, tz, vx, vy, vz, dt, d, h, inv_h, bc, rules, x, y, z);
}

// batched: workgroup i runs 16 x 16 block items[i].block of descriptor items[i].desc
__global__ __launch_bounds__(kAThreads) void k_advect_velocity_bb(const AdvectVelocityDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    __shared__ float2 rules[3][8];
    const BatchItem it = items[blockIdx.x];
    if (it.desc == kIdleItem) return;
    const AdvectVelocityDesc& a = descs[it.desc];
    fill_velocity_rules(rules, a.bc, static_cast<int>(threadIdx.x));
    __syncthreads();
    int x, y, z;
    if (!cell_in_block(a.d, static_cast<int>(it.block), x, y, z)) return;
    if (a.pow2) advect_velocity_cell<true>(a.t[0], a.t[1], a.t[2], a.v[0], a.v[1], a.v[2], a.dt, a.d, a.h, a.inv_h, a.bc, rules, x, y, z);
    else advect_velocity_cell<false>(a.t[0], a.t[1], a.t[2], a.v[0], a.v[1], a.v[2], a.dt, a.d, a.h, a.inv_h, a.bc, rules, x, y, z);
}

// Scalar face data of one axis (reference tap_scalar, backend-cuda.cu:62-94): a
// zero-flux face clamps the coordinate, any other non-periodic face returns its value.
struct ScalarAxis {
    bool per, fixed_m, fixed_p;
    float val_m, val_p;
};

__device__ __forceinline__ ScalarAxis scalar_axis(const StableFluidsScalarBoundaryConfig& b, const int a) {
    const auto& m = a == 0 ? b.x_minus : a == 1 ? b.y_minus : b.z_minus;
    const auto& p = a == 0 ? b.x_plus : a == 1 ? b.y_plus : b.z_plus;
    ScalarAxis s;
    s.per     = m.type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC && p.type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC;
    s.fixed_m = !s.per && m.type != STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX;
    s.fixed_p = !s.per && p.type != STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX;
    s.val_m   = m.value;
    s.val_p   = p.value;
    return s;
}

// reference advect_scalar_component_kernel backend-cuda.cu:342-352
template <bool kPow2>
__device__ __forceinline__ void advect_scalar_cell(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const float inv_h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc, const float2 (*rules)[8], const int x, const int y, const int z) {
    const uint32_t bytes = 4u * static_cast<uint32_t>(d.nx) * static_cast<uint32_t>(d.ny) * static_cast<uint32_t>(d.nz);
    const VelocityField v{{plane_rsrc(vx, bytes), plane_rsrc(vy, bytes), plane_rsrc(vz, bytes)}};
    const Rsrc s         = plane_rsrc(src, bytes);
    const Grid g         = grid_of(d, h, inv_h);
    const Periodic per   = {flow_periodic(fbc, 0), flow_periodic(fbc, 1), flow_periodic(fbc, 2)};
    const int cell_off   = z * g.sz + y * g.sy + 4 * x;
    float qx, qy, qz;
    trace_cell<kPow2>(v, x, y, z, cell_off, dt, g, per, rules, qx, qy, qz);
    const ScalarAxis ax = scalar_axis(sbc, 0), ay = scalar_axis(sbc, 1), az = scalar_axis(sbc, 2);
    if (ax.per) qx = wrap_pos(qx, d.nx, h);
    if (ay.per) qy = wrap_pos(qy, d.ny, h);
    if (az.per) qz = wrap_pos(qz, d.nz, h);
    const Tri p = locate_b<kPow2>(qx, qy, qz, g);
    float c[8];
    if (p.interior) {
        taps_interior(s, p, g, c);
    } else {
        const Axis X = axis_of(p.x0, g.nx, ax.per), Y = axis_of(p.y0, g.ny, ay.per), Z = axis_of(p.z0, g.nz, az.per);
        // per corner: does this coordinate hit a fixed-value face, and its value
        bool fx[2], fy[2], fz[2];
        float vx2[2], vy2[2], vz2[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            fx[k]  = X.o[k] && (X.m[k] ? ax.fixed_m : ax.fixed_p);
            fy[k]  = Y.o[k] && (Y.m[k] ? ay.fixed_m : ay.fixed_p);
            fz[k]  = Z.o[k] && (Z.m[k] ? az.fixed_m : az.fixed_p);
            vx2[k] = X.m[k] ? ax.val_m : ax.val_p;
            vy2[k] = Y.m[k] ? ay.val_m : ay.val_p;
            vz2[k] = Z.m[k] ? az.val_m : az.val_p;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int cx = t & 1, cy = (t >> 1) & 1, cz = t >> 2;
            const int ix = ax.per ? X.w[cx] : X.q[cx];
            const int iy = ay.per ? Y.w[cy] : Y.q[cy];
            const int iz = az.per ? Z.w[cz] : Z.q[cz];
            const float f = ld(s, iz * g.sz + iy * g.sy + 4 * ix, 0);
            c[t]          = fx[cx] ? vx2[cx] : fy[cy] ? vy2[cy] : fz[cz] ? vz2[cz] : f;
        }
    }
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(lerp8(c, p)), plane_rsrc(out, bytes), cell_off, 0, 0);
}

template <bool kPow2>
__global__ __launch_bounds__(kAThreads) __attribute__((amdgpu_waves_per_eu(kAdvectWaves))) void k_advect_scalar_b(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const float inv_h, const StableFluidsScalarBoundaryConfig sbc, const StableFluidsFlowBoundaryConfig fbc) {
    __shared__ float2 rules[3][8];
    fill_velocity_rules(rules, fbc, static_cast<int>(threadIdx.x));
    __syncthreads();
    int x, y, z;
    if (!cell_of(d, x, y, z)) return;
    advect_scalar_cell<kPow2>(out, src, vx, vy, vz, dt, d, h, inv_h, sbc, fbc, rules, x, y, z);
}

__global__ __launch_bounds__(kAThreads) void k_advect_scalar_bb(const AdvectScalarDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    __shared__ float2 rules[3][8];
    const BatchItem it = items[blockIdx.x];
    if (it.desc == kIdleItem) return;
    const AdvectScalarDesc& a = descs[it.desc];
    fill_velocity_rules(rules, a.fbc, static_cast<int>(threadIdx.x));
    __syncthreads();
    int x, y, z;
    if (!cell_in_block(a.d, static_cast<int>(it.block), x, y, z)) return;
    if (a.pow2) advect_scalar_cell<true>(a.out, a.src, a.v[0], a.v[1], a.v[2], a.dt, a.d, a.h, a.inv_h, a.sbc, a.fbc, rules, x, y, z);
    else advect_scalar_cell<false>(a.out, a.src, a.v[0], a.v[1], a.v[2], a.dt, a.d, a.h, a.inv_h, a.sbc, a.fbc, rules, x, y, z);
}

bool pow2_spacing(const float h) {
    int e;
    return h > 0.0f && std::isfinite(h) && std::frexp(h, &e) == 0.5f;
}

inline unsigned adv_grid(const Dims d) { return static_cast<unsigned>(((d.nx + kAX - 1) / kAX) * ((d.ny + kAY - 1) / kAY) * d.nz); }

} // namespace

bool advect_pow2(const float h) { return pow2_spacing(h); }
uint32_t advect_blocks(const Dims d) { return adv_grid(d); }

hipError_t launch_advect_velocity_batched(hipStream_t s, const AdvectVelocityDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_advect_velocity_bb<<<n_items, kAThreads, 0, s>>>(descs, items);
    return hipGetLastError();
}
hipError_t launch_advect_scalar_batched(hipStream_t s, const AdvectScalarDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_advect_scalar_bb<<<n_items, kAThreads, 0, s>>>(descs, items);
    return hipGetLastError();
}

bool advect_buffers_fit(const Dims d) { return 4ull * static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny) * static_cast<uint64_t>(d.nz) < (1ull << 31); }

hipError_t launch_advect_velocity_buffers(hipStream_t s, float* tx, float* ty, float* tz, const float* vx, const float* vy, const float* vz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc) {
    const float inv_h = 1.0f / h;
    if (pow2_spacing(h))
        k_advect_velocity_b<true><<<adv_grid(d), kAThreads, 0, s>>>(tx, ty, tz, vx, vy, vz, dt, d, h, inv_h, bc);
    else
        k_advect_velocity_b<false><<<adv_grid(d), kAThreads, 0, s>>>(tx, ty, tz, vx, vy, vz, dt, d, h, inv_h, bc);
    return hipGetLastError();
}

hipError_t launch_advect_scalar_buffers(hipStream_t s, float* out, const float* src, const float* vx, const float* vy, const float* vz, const float dt, const Dims d, const float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc) {
    const float inv_h = 1.0f / h;
    if (pow2_spacing(h))
        k_advect_scalar_b<true><<<adv_grid(d), kAThreads, 0, s>>>(out, src, vx, vy, vz, dt, d, h, inv_h, sbc, fbc);
    else
        k_advect_scalar_b<false><<<adv_grid(d), kAThreads, 0, s>>>(out, src, vx, vy, vz, dt, d, h, inv_h, sbc, fbc);
    return hipGetLastError();
}

} // namespace sf
