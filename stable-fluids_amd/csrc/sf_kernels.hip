// sf_kernels.hip -- gfx950 stage kernels of the stable-fluids 3D step.
//
// Arithmetic follows the reference kernels (backend-cuda.cu:302-434) expression by
// expression (same operand order, IEEE division, floorf/fmodf/sqrtf), so the FP32 results
// track the reference within rounding. The kernel set is fused differently:
//  * one RK2 trace per cell serves all three velocity components (reference traces 3x);
//  * smoother passes live in sf_smooth.hip (LDS z-marching tiles);
//  * projection, divergence of the projected field and |v| run in one pass.
#include <algorithm>

#include "sf_buffer.hpp"
#include "sf_kernels.hpp"

namespace sf {

namespace {

constexpr int kElemBlock = 256;
constexpr int kTileX     = 64; // one wavefront per row of a 3D tile
constexpr int kTileY     = 4;

__device__ __forceinline__ uint64_t global_tid() { return static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }

// ---------------- elementwise ----------------

__global__ __launch_bounds__(kElemBlock) void k_fill(float* __restrict__ p, const float value, const uint64_t n) {
    const uint64_t i = global_tid() * 4;
    if (i + 3 < n) {
        *reinterpret_cast<float4*>(p + i) = make_float4(value, value, value, value);
    } else {
        for (uint64_t k = i; k < n; ++k) p[k] = value;
    }
}

// dst = src * factor (reference dissipate_kernel backend-cuda.cu:377-384; factor 1 is a copy)
__global__ __launch_bounds__(kElemBlock) void k_copy_scaled(float* __restrict__ dst, const float* __restrict__ src, const float factor, const uint64_t n) {
    const uint64_t i = global_tid() * 4;
    if (i + 3 < n) {
        const float4 a = *reinterpret_cast<const float4*>(src + i);
        *reinterpret_cast<float4*>(dst + i) = make_float4(a.x * factor, a.y * factor, a.z * factor, a.w * factor);
    } else {
        for (uint64_t k = i; k < n; ++k) dst[k] = src[k] * factor;
    }
}

// dst += dt * src (reference add_field_source_kernel backend-cuda.cu:321-328)
__global__ __launch_bounds__(kElemBlock) void k_add_scaled(float* __restrict__ dst, const float* __restrict__ src, const float dt, const uint64_t n) {
    const uint64_t i = global_tid() * 4;
    if (i + 3 < n) {
        float4 a       = *reinterpret_cast<float4*>(dst + i);
        const float4 f = *reinterpret_cast<const float4*>(src + i);
        a.x += dt * f.x;
        a.y += dt * f.y;
        a.z += dt * f.z;
        a.w += dt * f.w;
        *reinterpret_cast<float4*>(dst + i) = a;
    } else {
        for (uint64_t k = i; k < n; ++k) dst[k] += dt * src[k];
    }
}

// v += dt * f for all three components (reference add_force_kernel backend-cuda.cu:310-319)
__global__ __launch_bounds__(kElemBlock) void k_add_force(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz, const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz, const float dt, const uint64_t n) {
    const uint64_t i = global_tid() * 4;
    float* v[3]       = {vx, vy, vz};
    const float* f[3] = {fx, fy, fz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (i + 3 < n) {
            float4 a        = *reinterpret_cast<float4*>(v[c] + i);
            const float4 fv = *reinterpret_cast<const float4*>(f[c] + i);
            a.x += dt * fv.x;
            a.y += dt * fv.y;
            a.z += dt * fv.z;
            a.w += dt * fv.w;
            *reinterpret_cast<float4*>(v[c] + i) = a;
        } else {
            for (uint64_t k = i; k < n; ++k) v[c][k] += dt * f[c][k];
        }
    }
}

// Animated force of the plume scene (vulkan-app/scene_plume.cpp:323-335, see
// stable_fluids_ext_set_force_animation_cuda): coefficients from the step counter, then
// x = ax * drift + s * swirl_x and z = az * drift + s * swirl_z per cell.
__global__ __launch_bounds__(kElemBlock) void k_animate_force(float* __restrict__ fx, float* __restrict__ fz, const float* __restrict__ drift, const float* __restrict__ swirl_x, const float* __restrict__ swirl_z, const ForceAnimation a, const uint32_t* __restrict__ counter, const uint64_t n) {
    const float phase = static_cast<float>(*counter) * a.phase_rate;
    const float ax    = a.drift_x_amplitude * sinf(phase);
    const float az    = a.drift_z_amplitude * cosf(phase * a.drift_z_rate);
    const float sw    = a.swirl_amplitude * cosf(phase * a.swirl_rate);
    for (uint64_t i = global_tid(); i < n; i += static_cast<uint64_t>(gridDim.x) * kElemBlock) {
        const float m = drift[i];
        fx[i]         = ax * m + sw * swirl_x[i];
        fz[i]         = az * m + sw * swirl_z[i];
    }
}

__global__ void k_count_step(uint32_t* __restrict__ counter) { *counter += 1u; }

// ---------------- advection ----------------

// One midpoint trace per cell, three component samples (reference
// advect_velocity_component_kernel backend-cuda.cu:330-340, launched per component).
__device__ __forceinline__ void advect_velocity_body(float* __restrict__ tx, float* __restrict__ ty, float* __restrict__ tz, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc, const int x, const int y, const int z) {
    if (x >= d.nx || y >= d.ny) return;
    const float px = (static_cast<float>(x) + 0.5f) * h;
    const float py = (static_cast<float>(y) + 0.5f) * h;
    const float pz = (static_cast<float>(z) + 0.5f) * h;
    float qx, qy, qz;
    trace_rk2_cell(vx, vy, vz, x, y, z, px, py, pz, dt, d, h, bc, qx, qy, qz);
    const uint64_t i = lin(x, y, z, d);
    float u, v, w;
    sample_velocity3(vx, vy, vz, qx, qy, qz, d, h, bc, u, v, w);
    tx[i] = u;
    ty[i] = v;
    tz[i] = w;
}

// reference advect_scalar_component_kernel backend-cuda.cu:342-352
__device__ __forceinline__ void advect_scalar_body(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc, const int x, const int y, const int z) {
    if (x >= d.nx || y >= d.ny) return;
    const float px = (static_cast<float>(x) + 0.5f) * h;
    const float py = (static_cast<float>(y) + 0.5f) * h;
    const float pz = (static_cast<float>(z) + 0.5f) * h;
    float qx, qy, qz;
    trace_rk2_cell(vx, vy, vz, x, y, z, px, py, pz, dt, d, h, fbc, qx, qy, qz);
    if (scalar_periodic(sbc, 0)) qx = wrap_position(qx, d.nx, h);
    if (scalar_periodic(sbc, 1)) qy = wrap_position(qy, d.ny, h);
    if (scalar_periodic(sbc, 2)) qz = wrap_position(qz, d.nz, h);
    const TriPos p       = locate(qx, qy, qz, d, h);
    out[lin(x, y, z, d)] = sample_at(src, p, d, [&](int i, int j, int k) { return tap_scalar(src, i, j, k, d, sbc); });
}

// ---------------- divergence / projection ----------------

// Component `axis` of the field at the neighbour of (x,y,z) along `axis` (dir = +-1),
// where `value_at(q)` gives the in-grid value and `centre` the value at (x,y,z).
template <class F>
__device__ __forceinline__ float along_axis(const int axis, const int dir, const int x, const int y, const int z, const Dims d, const StencilBC& bc, const float centre, F value_at) {
    const int n[3] = {d.nx, d.ny, d.nz};
    int q[3]       = {x, y, z};
    if (step_axis(q[axis], dir, n[axis], (bc.periodic >> axis) & 1)) return value_at(q[0], q[1], q[2]);
    return apply_ghost(bc.face[2 * axis + (dir > 0 ? 1 : 0)], centre);
}

// reference compute_divergence_kernel backend-cuda.cu:386-396
__device__ __forceinline__ void divergence_body(float* __restrict__ div, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const Dims d, const float h, const StencilBC& dbc, const int x, const int y, const int z) {
    if (x >= d.nx || y >= d.ny) return;
    const uint64_t i    = lin(x, y, z, d);
    const float inv_2h  = 0.5f / h;
    const float* v[3]   = {vx, vy, vz};
    float acc           = 0.0f;
    if (x > 0 && x < d.nx - 1 && y > 0 && y < d.ny - 1 && z > 0 && z < d.nz - 1) {
        // interior: every tap is the direct neighbour (the same expressions, no ghost rule)
        const uint64_t st[3] = {1, static_cast<uint64_t>(d.nx), static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny)};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float dd = (v[c][i + st[c]] - v[c][i - st[c]]) * inv_2h;
            acc            = c == 0 ? dd : acc + dd;
        }
        div[i] = acc;
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float* f     = v[c];
        const float centre = f[i];
        auto at            = [&](int a, int b2, int c2) { return f[lin(a, b2, c2, d)]; };
        const float plus   = along_axis(c, 1, x, y, z, d, dbc, centre, at);
        const float minus  = along_axis(c, -1, x, y, z, d, dbc, centre, at);
        const float dd     = (plus - minus) * inv_2h;
        acc                = c == 0 ? dd : acc + dd;
    }
    div[i] = acc;
}

// Gradient component `c` of p at (x,y,z) (reference backend-cuda.cu:416-418).
__device__ __forceinline__ float grad_p(const float* __restrict__ p, const int c, const int x, const int y, const int z, const Dims d, const StencilBC& pbc, const float inv_2h) {
    const float centre = p[lin(x, y, z, d)];
    auto at            = [&](int a, int b2, int c2) { return p[lin(a, b2, c2, d)]; };
    const float plus   = along_axis(c, 1, x, y, z, d, pbc, centre, at);
    const float minus  = along_axis(c, -1, x, y, z, d, pbc, centre, at);
    return (plus - minus) * inv_2h;
}

// reference project_velocity_kernel :409-422, then compute_divergence_kernel :386-396 and
// velocity_magnitude_kernel :424-434 on the projected field (the scalar stage between
// them does not touch the velocity). Out of place: neighbours are re-projected locally.
__device__ __forceinline__ void project_body(float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, float* __restrict__ div, float* __restrict__ mag, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ p, const Dims d, const float h, const StencilBC& pbc, const StencilBC& dbc, const int x, const int y, const int z) {
    if (x >= d.nx || y >= d.ny) return;
    const uint64_t i   = lin(x, y, z, d);
    const float inv_2h = 0.5f / h;
    const float* v[3]  = {vx, vy, vz};
    float proj[3];
    if (x > 1 && x < d.nx - 2 && y > 1 && y < d.ny - 2 && z > 1 && z < d.nz - 2) {
        // two cells from every face: every velocity and pressure tap (pressure two cells out,
        // through the projected neighbours) is a direct load; the same expressions, no ghost rule
        const uint64_t st[3] = {1, static_cast<uint64_t>(d.nx), static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny)};
        const float pc       = p[i];
        float acc            = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint64_t s1 = st[c], s2 = 2 * st[c];
            const float pp = p[i + s1], pm = p[i - s1];
            proj[c]          = v[c][i] - (pp - pm) * inv_2h;
            const float plus  = v[c][i + s1] - (p[i + s2] - pc) * inv_2h;
            const float minus = v[c][i - s1] - (pc - p[i - s2]) * inv_2h;
            const float dd    = (plus - minus) * inv_2h;
            acc               = c == 0 ? dd : acc + dd;
        }
        ox[i]  = proj[0];
        oy[i]  = proj[1];
        oz[i]  = proj[2];
        div[i] = acc;
        mag[i] = sqrtf(proj[0] * proj[0] + proj[1] * proj[1] + proj[2] * proj[2]);
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) proj[c] = v[c][i] - grad_p(p, c, x, y, z, d, pbc, inv_2h);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float* f    = v[c];
        auto at           = [&](int a, int b2, int c2) { return f[lin(a, b2, c2, d)] - grad_p(p, c, a, b2, c2, d, pbc, inv_2h); };
        const float plus  = along_axis(c, 1, x, y, z, d, dbc, proj[c], at);
        const float minus = along_axis(c, -1, x, y, z, d, dbc, proj[c], at);
        const float dd    = (plus - minus) * inv_2h;
        acc               = c == 0 ? dd : acc + dd;
    }
    ox[i]  = proj[0];
    oy[i]  = proj[1];
    oz[i]  = proj[2];
    div[i] = acc;
    mag[i] = sqrtf(proj[0] * proj[0] + proj[1] * proj[1] + proj[2] * proj[2]);
}


__global__ __launch_bounds__(kTileX* kTileY) void k_advect_velocity(float* __restrict__ tx, float* __restrict__ ty, float* __restrict__ tz, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig bc) {
    advect_velocity_body(tx, ty, tz, vx, vy, vz, dt, d, h, bc, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y), static_cast<int>(blockIdx.z));
}
__global__ __launch_bounds__(kTileX* kTileY) void k_advect_scalar(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float dt, const Dims d, const float h, const StableFluidsScalarBoundaryConfig sbc, const StableFluidsFlowBoundaryConfig fbc) {
    advect_scalar_body(out, src, vx, vy, vz, dt, d, h, sbc, fbc, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y), static_cast<int>(blockIdx.z));
}
// Cell of this thread in a 64 x 4 tile: wave w covers x in [16 w, 16 w + 16) of all four
// rows (four 64-byte row segments), so only the waves on the outermost 16 columns of a row
// meet the x faces (a 64-cell row per wave put half of the waves there).
__device__ __forceinline__ void cell_of_thread(const int bx, const int by, int& x, int& y) {
    const int t = static_cast<int>(threadIdx.y) * kTileX + static_cast<int>(threadIdx.x);
    const int w = t >> 6, l = t & 63;
    x           = bx * kTileX + 16 * w + (l & 15);
    y           = by * kTileY + (l >> 4);
}

__global__ __launch_bounds__(kTileX* kTileY) void k_divergence(float* __restrict__ div, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const Dims d, const float h, const StencilBC dbc) {
    int x, y;
    cell_of_thread(static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), x, y);
    divergence_body(div, vx, vy, vz, d, h, dbc, x, y, static_cast<int>(blockIdx.z));
}
__global__ __launch_bounds__(kTileX* kTileY) void k_project(float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, float* __restrict__ div, float* __restrict__ mag, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ p, const Dims d, const float h, const StencilBC pbc, const StencilBC dbc) {
    int x, y;
    cell_of_thread(static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), x, y);
    project_body(ox, oy, oz, div, mag, vx, vy, vz, p, d, h, pbc, dbc, x, y, static_cast<int>(blockIdx.z));
}

// ---------------- divergence / projection, z-marching ----------------
//
// The same cell arithmetic, reorganised so every input is read once: a 512-thread
// workgroup owns a 64 x 16 tile of a z chunk and marches upward. Per plane it stages the
// tile plus its halo of the fields whose x and y neighbours are needed (pressure: two
// cells, the x / y velocity components: one cell along their own axis) in LDS,
// double-buffered so one barrier per plane orders the ring; the z neighbours come from
// per-lane register windows (pressure z - 2 .. z + 2, the z component z - 1 .. z + 1).
// Loads run one plane ahead through buffer resources; halo positions on periodic axes
// are loaded from the wrapped cell, others outside the grid read 0 and are never used.
// Every cell evaluates the general rule of project_body / divergence_body on the staged
// values: a neighbour step that stays in the grid (or wraps on a periodic axis) takes the
// neighbour, any other the face's ghost rule of the centre (apply_ghost), in the same
// operand order, so the results are bit-identical. Fields below 2 GiB.
constexpr int kZX = 64, kZWaves = 8, kZRows = 2, kZY = kZWaves * kZRows; // lane: column x; wave w: rows w, w + 8
constexpr int kZThreads = 64 * kZWaves;
// Planes per chunk: 32, fewer where the grid's tiles would leave CUs idle -- about 512
// workgroups (two per CU of MI355X's 256), at least 4 planes. Host and device both take
// it from here, and results do not depend on it.
__host__ __device__ __forceinline__ int z_chunk(const Dims d) {
    const int tiles = ((d.nx + kZX - 1) / kZX) * ((d.ny + kZY - 1) / kZY);
    const int zc    = (d.nz * tiles + 511) / 512;
    return zc < 4 ? 4 : (zc > 32 ? 32 : zc);
}

struct ZTile {
    int x0, y0, z0, z1;
};

// local block w of a z-marching launch over d -> tile and chunk (chunks of a tile column
// are consecutive, so an XCD-contiguous block order gives each XCD runs of chunks of a few
// tile columns, whose register-window planes meet in its L2)
__device__ __forceinline__ ZTile z_tile_of(const Dims d, const int w) {
    const int zc = z_chunk(d);
    const int tx = (d.nx + kZX - 1) / kZX, ch = (d.nz + zc - 1) / zc;
    ZTile t;
    const int cz = w % ch, tt = w / ch;
    t.x0 = (tt % tx) * kZX;
    t.y0 = (tt / tx) * kZY;
    t.z0 = cz * zc;
    t.z1 = min(t.z0 + zc, d.nz);
    return t;
}
// the workgroup's tile in a single-context launch (blocks renumbered per XCD)
__device__ __forceinline__ ZTile z_tile(const Dims d) {
    const int n = static_cast<int>(gridDim.x), w0 = static_cast<int>(blockIdx.x);
    return z_tile_of(d, (n & 7) == 0 ? (w0 & 7) * (n >> 3) + (w0 >> 3) : w0);
}

// A staged plane of W x H positions (the tile plus halos HX, HY) spread over the
// workgroup's threads, kSlots per thread: in-plane byte offsets (wrapped on periodic
// axes, kOutOfRange outside the grid otherwise) and the values of the plane in flight.
template <int W, int H, int HX, int HY>
struct ZStage {
    static constexpr int kN     = W * H;
    static constexpr int kSlots = (kN + kZThreads - 1) / kZThreads;
    int off[kSlots];
    float v[kSlots];
    __device__ __forceinline__ void init(const ZTile& t, const Dims d, const int periodic) {
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const int e = static_cast<int>(threadIdx.x) + k * kZThreads;
            int gx = t.x0 + e % W - HX, gy = t.y0 + e / W - HY;
            bool in = e < kN;
            if (gx < 0 || gx >= d.nx) {
                in = in && (periodic & 1);
                gx = wrap_index(gx, d.nx);
            }
            if (gy < 0 || gy >= d.ny) {
                in = in && (periodic & 2);
                gy = wrap_index(gy, d.ny);
            }
            off[k] = in ? 4 * (gy * d.nx + gx) : kOutOfRange;
        }
    }
    __device__ __forceinline__ void load(const Rsrc r, const int soff) {
#pragma unroll
        for (int k = 0; k < kSlots; ++k) v[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off[k], soff, 0));
    }
    __device__ __forceinline__ void store(float* lds) const {
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const int e = static_cast<int>(threadIdx.x) + k * kZThreads;
            if (k + 1 < kSlots || e < kN) lds[e] = v[k];
        }
    }
};

// Resource over the whole field and the byte offset of plane z, wrapped on a periodic z
// axis; a 0-byte resource (reads 0) for planes outside a non-periodic one.
__device__ __forceinline__ Rsrc z_plane(const void* base, int z, const Dims d, const bool periodic, int& soff) {
    const uint32_t plane = 4u * static_cast<uint32_t>(d.nx) * static_cast<uint32_t>(d.ny);
    if (periodic) z = wrap_index(z, d.nz);
    const bool in = z >= 0 && z < d.nz;
    soff          = in ? z * static_cast<int>(plane) : 0;
    return plane_rsrc(base, in ? plane * static_cast<uint32_t>(d.nz) : 0u);
}

// One axis of the divergence of the projected field at a cell (project_body's general
// rule): c0 centre of the velocity component along the axis, v1m / v1p its neighbours,
// p2m .. p2p the pressure at -2 .. +2 along the axis; a1m, a1p, a2m, a2p: the pressure
// steps -1, +1, -2, +2 stay in the grid (pressure rules); d1m, d1p: the velocity steps do
// (divergence rules); fp = 2 axis. Returns the projected component, and the term in dd.
__device__ __forceinline__ float project_axis(const float c0, const float v1m, const float v1p, const float p2m, const float p1m, const float pc, const float p1p, const float p2p, const bool a2m, const bool a1m, const bool a1p, const bool a2p, const bool d1m, const bool d1p, const StencilBC& pbc, const StencilBC& dbc, const int fp, const float inv_2h, float& dd) {
    const float gplus  = a1p ? p1p : apply_ghost(pbc.face[fp + 1], pc);
    const float gminus = a1m ? p1m : apply_ghost(pbc.face[fp], pc);
    const float proj   = c0 - (gplus - gminus) * inv_2h;
    // the neighbours' projected values: their gradient has this cell on the inner side
    const float npp   = a2p ? p2p : apply_ghost(pbc.face[fp + 1], p1p);
    const float nmm   = a2m ? p2m : apply_ghost(pbc.face[fp], p1m);
    const float plus  = d1p ? v1p - (npp - pc) * inv_2h : apply_ghost(dbc.face[fp + 1], proj);
    const float minus = d1m ? v1m - (pc - nmm) * inv_2h : apply_ghost(dbc.face[fp], proj);
    dd                = (plus - minus) * inv_2h;
    return proj;
}

__device__ __forceinline__ void project_z_body(float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, float* __restrict__ div, float* __restrict__ mag, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ p, const Dims d, const float h, const StencilBC& pbc, const StencilBC& dbc, const ZTile t) {
    constexpr int PW = kZX + 4, PH = kZY + 4; // pressure: two-cell halo
    constexpr int XW = kZX + 2;               // x component: one-cell x halo
    constexpr int YH = kZY + 2;               // y component: one-cell y halo
    using SP = ZStage<PW, PH, 2, 2>;
    using SX = ZStage<XW, kZY, 1, 0>;
    using SY = ZStage<kZX, YH, 0, 1>;
    __shared__ float s_p[2][PW * PH], s_x[2][XW * kZY], s_y[2][kZX * YH];
    const int lane     = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int x        = t.x0 + lane;
    const float inv_2h = 0.5f / h;
    const int pp = pbc.periodic, dp = dbc.periodic;
    SP sp;
    SX sx;
    SY sy;
    sp.init(t, d, pp);
    sx.init(t, d, dp);
    sy.init(t, d, dp);
    // own columns (rows y0 + wave + 8 r): in-plane offsets and register windows
    int o_off[kZRows];
    float pw[5][kZRows], zw[3][kZRows], pn[kZRows], zn[kZRows];
#pragma unroll
    for (int r = 0; r < kZRows; ++r) {
        const int y = t.y0 + wave + kZWaves * r;
        o_off[r]    = (x < d.nx && y < d.ny) ? 4 * (y * d.nx + x) : kOutOfRange;
    }
    auto own = [&](const float* f, const int z, const bool per, float (&out)[kZRows]) {
        int so;
        const Rsrc rs = z_plane(f, z, d, per, so);
#pragma unroll
        for (int r = 0; r < kZRows; ++r) out[r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, o_off[r], so, 0));
    };
    auto stage_loads = [&](const int z) {
        int so;
        Rsrc rs = z_plane(p, z, d, false, so);
        sp.load(rs, so);
        rs = z_plane(vx, z, d, false, so);
        sx.load(rs, so);
        rs = z_plane(vy, z, d, false, so);
        sy.load(rs, so);
    };
    // prologue: windows for plane z0, the staged planes of z0 in flight
    for (int k = 0; k < 5; ++k) own(p, t.z0 - 2 + k, pp & 4, pw[k]);
    for (int k = 0; k < 3; ++k) own(vz, t.z0 - 1 + k, dp & 4, zw[k]);
    stage_loads(t.z0);
    // x step flags of the lane's column (y and z per row and plane)
    const bool ax2m = (pp & 1) || x >= 2, ax1m = (pp & 1) || x >= 1, ax1p = (pp & 1) || x + 1 < d.nx, ax2p = (pp & 1) || x + 2 < d.nx;
    const bool dx1m = (dp & 1) || x >= 1, dx1p = (dp & 1) || x + 1 < d.nx;
    for (int z = t.z0; z < t.z1; ++z) {
        const int b = z & 1;
        sp.store(s_p[b]);
        sx.store(s_x[b]);
        sy.store(s_y[b]);
        if (z + 1 < t.z1) stage_loads(z + 1);
        own(p, z + 3, pp & 4, pn);
        own(vz, z + 2, dp & 4, zn);
        __syncthreads();
        const bool az2m = (pp & 4) || z >= 2, az1m = (pp & 4) || z >= 1, az1p = (pp & 4) || z + 1 < d.nz, az2p = (pp & 4) || z + 2 < d.nz;
        const bool dz1m = (dp & 4) || z >= 1, dz1p = (dp & 4) || z + 1 < d.nz;
#pragma unroll
        for (int r = 0; r < kZRows; ++r) {
            const int ty = wave + kZWaves * r, y = t.y0 + ty;
            if (x >= d.nx || y >= d.ny) continue;
            const bool ay2m = (pp & 2) || y >= 2, ay1m = (pp & 2) || y >= 1, ay1p = (pp & 2) || y + 1 < d.ny, ay2p = (pp & 2) || y + 2 < d.ny;
            const bool dy1m = (dp & 2) || y >= 1, dy1p = (dp & 2) || y + 1 < d.ny;
            const float* P = s_p[b] + (ty + 2) * PW + lane + 2;
            const float* X = s_x[b] + ty * XW + lane + 1;
            const float* Y = s_y[b] + (ty + 1) * kZX + lane;
            const float pc = pw[2][r];
            float proj[3], acc, dd;
            proj[0] = project_axis(X[0], X[-1], X[1], P[-2], P[-1], pc, P[1], P[2], ax2m, ax1m, ax1p, ax2p, dx1m, dx1p, pbc, dbc, 0, inv_2h, dd);
            acc     = dd;
            proj[1] = project_axis(Y[0], Y[-kZX], Y[kZX], P[-2 * PW], P[-PW], pc, P[PW], P[2 * PW], ay2m, ay1m, ay1p, ay2p, dy1m, dy1p, pbc, dbc, 2, inv_2h, dd);
            acc     = acc + dd;
            proj[2] = project_axis(zw[1][r], zw[0][r], zw[2][r], pw[0][r], pw[1][r], pc, pw[3][r], pw[4][r], az2m, az1m, az1p, az2p, dz1m, dz1p, pbc, dbc, 4, inv_2h, dd);
            acc     = acc + dd;
            const uint64_t i = lin(x, y, z, d);
            ox[i]  = proj[0];
            oy[i]  = proj[1];
            oz[i]  = proj[2];
            div[i] = acc;
            mag[i] = sqrtf(proj[0] * proj[0] + proj[1] * proj[1] + proj[2] * proj[2]);
        }
#pragma unroll
        for (int r = 0; r < kZRows; ++r) {
            pw[0][r] = pw[1][r];
            pw[1][r] = pw[2][r];
            pw[2][r] = pw[3][r];
            pw[3][r] = pw[4][r];
            pw[4][r] = pn[r];
            zw[0][r] = zw[1][r];
            zw[1][r] = zw[2][r];
            zw[2][r] = zn[r];
        }
    }
}

// One axis of divergence_body's general rule: (plus - minus) / 2h with ghost rules of the
// centre at the faces.
__device__ __forceinline__ float divergence_axis(const float c0, const float v1m, const float v1p, const bool d1m, const bool d1p, const StencilBC& dbc, const int fp, const float inv_2h) {
    const float plus  = d1p ? v1p : apply_ghost(dbc.face[fp + 1], c0);
    const float minus = d1m ? v1m : apply_ghost(dbc.face[fp], c0);
    return (plus - minus) * inv_2h;
}

__device__ __forceinline__ void divergence_z_body(float* __restrict__ div, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const Dims d, const float h, const StencilBC& dbc, const ZTile t) {
    constexpr int XW = kZX + 2, YH = kZY + 2;
    using SX = ZStage<XW, kZY, 1, 0>;
    using SY = ZStage<kZX, YH, 0, 1>;
    __shared__ float s_x[2][XW * kZY], s_y[2][kZX * YH];
    const int lane     = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int x        = t.x0 + lane;
    const float inv_2h = 0.5f / h;
    const int dp       = dbc.periodic;
    SX sx;
    SY sy;
    sx.init(t, d, dp);
    sy.init(t, d, dp);
    int o_off[kZRows];
    float zw[3][kZRows], zn[kZRows];
#pragma unroll
    for (int r = 0; r < kZRows; ++r) {
        const int y = t.y0 + wave + kZWaves * r;
        o_off[r]    = (x < d.nx && y < d.ny) ? 4 * (y * d.nx + x) : kOutOfRange;
    }
    auto own = [&](const int z, float (&out)[kZRows]) {
        int so;
        const Rsrc rs = z_plane(vz, z, d, dp & 4, so);
#pragma unroll
        for (int r = 0; r < kZRows; ++r) out[r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, o_off[r], so, 0));
    };
    auto stage_loads = [&](const int z) {
        int so;
        Rsrc rs = z_plane(vx, z, d, false, so);
        sx.load(rs, so);
        rs = z_plane(vy, z, d, false, so);
        sy.load(rs, so);
    };
    for (int k = 0; k < 3; ++k) own(t.z0 - 1 + k, zw[k]);
    stage_loads(t.z0);
    const bool dx1m = (dp & 1) || x >= 1, dx1p = (dp & 1) || x + 1 < d.nx;
    for (int z = t.z0; z < t.z1; ++z) {
        const int b = z & 1;
        sx.store(s_x[b]);
        sy.store(s_y[b]);
        if (z + 1 < t.z1) stage_loads(z + 1);
        own(z + 2, zn);
        __syncthreads();
        const bool dz1m = (dp & 4) || z >= 1, dz1p = (dp & 4) || z + 1 < d.nz;
#pragma unroll
        for (int r = 0; r < kZRows; ++r) {
            const int ty = wave + kZWaves * r, y = t.y0 + ty;
            if (x >= d.nx || y >= d.ny) continue;
            const bool dy1m = (dp & 2) || y >= 1, dy1p = (dp & 2) || y + 1 < d.ny;
            const float* X = s_x[b] + ty * XW + lane + 1;
            const float* Y = s_y[b] + (ty + 1) * kZX + lane;
            float acc      = divergence_axis(X[0], X[-1], X[1], dx1m, dx1p, dbc, 0, inv_2h);
            acc            = acc + divergence_axis(Y[0], Y[-kZX], Y[kZX], dy1m, dy1p, dbc, 2, inv_2h);
            acc            = acc + divergence_axis(zw[1][r], zw[0][r], zw[2][r], dz1m, dz1p, dbc, 4, inv_2h);
            div[lin(x, y, z, d)] = acc;
        }
#pragma unroll
        for (int r = 0; r < kZRows; ++r) {
            zw[0][r] = zw[1][r];
            zw[1][r] = zw[2][r];
            zw[2][r] = zn[r];
        }
    }
}

__global__ __launch_bounds__(kZThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_project_z(float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, float* __restrict__ div, float* __restrict__ mag, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ p, const Dims d, const float h, const StencilBC pbc, const StencilBC dbc) {
    project_z_body(ox, oy, oz, div, mag, vx, vy, vz, p, d, h, pbc, dbc, z_tile(d));
}
__global__ __launch_bounds__(kZThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_divergence_z(float* __restrict__ div, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const Dims d, const float h, const StencilBC dbc) {
    divergence_z_body(div, vx, vy, vz, d, h, dbc, z_tile(d));
}
// batched: workgroup i runs z-marching block items[i].block of descriptor items[i].desc
__global__ __launch_bounds__(kZThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_project_z_batched(const ProjectDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    if (it.desc == kIdleItem) return;
    const ProjectDesc& sd = descs[it.desc];
    project_z_body(sd.o[0], sd.o[1], sd.o[2], sd.div, sd.mag, sd.v[0], sd.v[1], sd.v[2], sd.p, sd.d, sd.h, sd.pbc, sd.dbc, z_tile_of(sd.d, static_cast<int>(it.block)));
}
__global__ __launch_bounds__(kZThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_divergence_z_batched(const DivergenceDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    if (it.desc == kIdleItem) return;
    const DivergenceDesc& sd = descs[it.desc];
    divergence_z_body(sd.div, sd.v[0], sd.v[1], sd.v[2], sd.d, sd.h, sd.dbc, z_tile_of(sd.d, static_cast<int>(it.block)));
}

uint32_t z_blocks(const Dims d) { return static_cast<uint32_t>(((d.nx + kZX - 1) / kZX) * ((d.ny + kZY - 1) / kZY) * ((d.nz + z_chunk(d) - 1) / z_chunk(d))); }

// Batched elementwise stages: local block b covers elements [4 * 256 * b, + 1024).
__global__ __launch_bounds__(kElemBlock) void k_add_force_batched(const AddForceDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    const AddForceDesc& sd = descs[it.desc];
    const uint64_t i = (static_cast<uint64_t>(it.block) * kElemBlock + threadIdx.x) * 4;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (i + 3 < sd.n) {
            float4 a        = *reinterpret_cast<float4*>(sd.v[c] + i);
            const float4 fv = *reinterpret_cast<const float4*>(sd.f[c] + i);
            a.x += sd.dt * fv.x;
            a.y += sd.dt * fv.y;
            a.z += sd.dt * fv.z;
            a.w += sd.dt * fv.w;
            *reinterpret_cast<float4*>(sd.v[c] + i) = a;
        } else {
            for (uint64_t k = i; k < sd.n; ++k) sd.v[c][k] += sd.dt * sd.f[c][k];
        }
    }
}

__global__ __launch_bounds__(kElemBlock) void k_add_scaled_batched(const AddScaledDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    const AddScaledDesc& sd = descs[it.desc];
    const uint64_t i = (static_cast<uint64_t>(it.block) * kElemBlock + threadIdx.x) * 4;
    if (i + 3 < sd.n) {
        float4 a       = *reinterpret_cast<float4*>(sd.dst + i);
        const float4 f = *reinterpret_cast<const float4*>(sd.src + i);
        a.x += sd.dt * f.x;
        a.y += sd.dt * f.y;
        a.z += sd.dt * f.z;
        a.w += sd.dt * f.w;
        *reinterpret_cast<float4*>(sd.dst + i) = a;
    } else {
        for (uint64_t k = i; k < sd.n; ++k) sd.dst[k] += sd.dt * sd.src[k];
    }
}

inline unsigned elem_blocks(const uint64_t n) { return static_cast<unsigned>((n + 4 * kElemBlock - 1) / (4 * kElemBlock)); }
inline dim3 tile_grid(const Dims d) { return dim3(static_cast<unsigned>((d.nx + kTileX - 1) / kTileX), static_cast<unsigned>((d.ny + kTileY - 1) / kTileY), static_cast<unsigned>(d.nz)); }
inline dim3 tile_block() { return dim3(kTileX, kTileY, 1); }

} // namespace

hipError_t launch_fill(hipStream_t s, float* p, const float value, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_fill<<<elem_blocks(n), kElemBlock, 0, s>>>(p, value, n);
    return hipGetLastError();
}

hipError_t launch_copy_scaled(hipStream_t s, float* dst, const float* src, const float factor, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_copy_scaled<<<elem_blocks(n), kElemBlock, 0, s>>>(dst, src, factor, n);
    return hipGetLastError();
}

hipError_t launch_add_scaled(hipStream_t s, float* dst, const float* src, const float dt, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_add_scaled<<<elem_blocks(n), kElemBlock, 0, s>>>(dst, src, dt, n);
    return hipGetLastError();
}

hipError_t launch_add_force(hipStream_t s, float* vx, float* vy, float* vz, const float* fx, const float* fy, const float* fz, const float dt, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_add_force<<<elem_blocks(n), kElemBlock, 0, s>>>(vx, vy, vz, fx, fy, fz, dt, n);
    return hipGetLastError();
}

hipError_t launch_animate_force(hipStream_t s, float* fx, float* fz, const float* drift, const float* swirl_x, const float* swirl_z, const ForceAnimation& a, const uint32_t* counter, const uint64_t n) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>((n + kElemBlock - 1) / kElemBlock, 4096));
    k_animate_force<<<blocks, kElemBlock, 0, s>>>(fx, fz, drift, swirl_x, swirl_z, a, counter, n);
    return hipGetLastError();
}

hipError_t launch_count_step(hipStream_t s, uint32_t* counter) {
    k_count_step<<<1, 1, 0, s>>>(counter);
    return hipGetLastError();
}

hipError_t launch_advect_velocity(hipStream_t s, float* tx, float* ty, float* tz, const float* vx, const float* vy, const float* vz, const float dt, const Dims d, const float h, const StableFluidsFlowBoundaryConfig& bc) {
    if (advect_buffers_fit(d)) return launch_advect_velocity_buffers(s, tx, ty, tz, vx, vy, vz, dt, d, h, bc);
    k_advect_velocity<<<tile_grid(d), tile_block(), 0, s>>>(tx, ty, tz, vx, vy, vz, dt, d, h, bc);
    return hipGetLastError();
}

hipError_t launch_advect_scalar(hipStream_t s, float* out, const float* src, const float* vx, const float* vy, const float* vz, const float dt, const Dims d, const float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc) {
    if (advect_buffers_fit(d)) return launch_advect_scalar_buffers(s, out, src, vx, vy, vz, dt, d, h, sbc, fbc);
    k_advect_scalar<<<tile_grid(d), tile_block(), 0, s>>>(out, src, vx, vy, vz, dt, d, h, sbc, fbc);
    return hipGetLastError();
}

hipError_t launch_divergence(hipStream_t s, float* div, const float* vx, const float* vy, const float* vz, const Dims d, const float h, const StencilBC& div_bc) {
    if (advect_buffers_fit(d)) k_divergence_z<<<z_blocks(d), kZThreads, 0, s>>>(div, vx, vy, vz, d, h, div_bc);
    else k_divergence<<<tile_grid(d), tile_block(), 0, s>>>(div, vx, vy, vz, d, h, div_bc);
    return hipGetLastError();
}

hipError_t launch_project(hipStream_t s, float* ox, float* oy, float* oz, float* div, float* mag, const float* vx, const float* vy, const float* vz, const float* p, const Dims d, const float h, const StencilBC& p_bc, const StencilBC& div_bc) {
    if (advect_buffers_fit(d)) k_project_z<<<z_blocks(d), kZThreads, 0, s>>>(ox, oy, oz, div, mag, vx, vy, vz, p, d, h, p_bc, div_bc);
    else k_project<<<tile_grid(d), tile_block(), 0, s>>>(ox, oy, oz, div, mag, vx, vy, vz, p, d, h, p_bc, div_bc);
    return hipGetLastError();
}

uint32_t elem_block_count(const uint64_t n) { return static_cast<uint32_t>((n + 4 * kElemBlock - 1) / (4 * kElemBlock)); }

hipError_t launch_add_force_batched(hipStream_t s, const AddForceDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_add_force_batched<<<n_items, kElemBlock, 0, s>>>(descs, items);
    return hipGetLastError();
}
hipError_t launch_add_scaled_batched(hipStream_t s, const AddScaledDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_add_scaled_batched<<<n_items, kElemBlock, 0, s>>>(descs, items);
    return hipGetLastError();
}
hipError_t launch_divergence_z_batched(hipStream_t s, const DivergenceDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_divergence_z_batched<<<n_items, kZThreads, 0, s>>>(descs, items);
    return hipGetLastError();
}
hipError_t launch_project_z_batched(hipStream_t s, const ProjectDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    k_project_z_batched<<<n_items, kZThreads, 0, s>>>(descs, items);
    return hipGetLastError();
}
uint32_t z_march_blocks(const Dims d) { return z_blocks(d); }

} // namespace sf
