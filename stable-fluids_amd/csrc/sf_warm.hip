// sf_warm.hip -- learned pressure warm start and the FP64 CG solve of its training pairs.
//
// No counterpart in the reference (its projection restarts pressure from zero,
// backend-cuda.cu:644). Stage kernels of the warm start, in graph order:
//   restrict_input: coarse divergence (mean of f^3 cells), wall mask, sums of squares;
//   (sigma and the input packing, network forward: sf_unet_kernels.hip, sf_unet.hip)
//   prolong:        trilinear guess on the full grid, scaled by (2h)^2 sigma;
//   guess_test:     ||b - A p0|| against ||b|| (FP64), fallback flag, skip flag.
// All reductions are deterministic two-pass sums (per-block partials, then one block).
#include <algorithm>

#include "sf_warm.hpp"

namespace sf::warm {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of v over the block (256 threads) into *out (thread 0).
__device__ __forceinline__ void block_sum_to(double v, double* out) {
    __shared__ double s[kThreads / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = s[0] + s[1] + s[2] + s[3];
    __syncthreads();
}

__device__ __forceinline__ void decode(const uint64_t i, const Dims d, int& x, int& y, int& z) {
    if (i <= 0xffffffffull) { // 32-bit division: a 64-bit one is a long instruction sequence
        const uint32_t j = static_cast<uint32_t>(i), nx = static_cast<uint32_t>(d.nx), ny = static_cast<uint32_t>(d.ny);
        const uint32_t r = j / nx;
        x                = static_cast<int>(j - r * nx);
        y                = static_cast<int>(r % ny);
        z                = static_cast<int>(r / ny);
        return;
    }
    x = static_cast<int>(i % static_cast<uint64_t>(d.nx));
    const uint64_t r = i / static_cast<uint64_t>(d.nx);
    y = static_cast<int>(r % static_cast<uint64_t>(d.ny));
    z = static_cast<int>(r / static_cast<uint64_t>(d.ny));
}

// 1 when the cell touches a non-periodic face whose pressure rule is a wall (copy).
__device__ __forceinline__ float wall_mask(const int x, const int y, const int z, const Dims d, const StencilBC& bc) {
    const int c[3] = {x, y, z}, n[3] = {d.nx, d.ny, d.nz};
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1;
        if ((bc.periodic >> a) & 1) continue;
        const bool at = (f & 1) ? c[a] == n[a] - 1 : c[a] == 0;
        if (at && bc.face[f].mode == kGhostCopy) return 1.0f;
    }
    return 0.0f;
}

// sum of the 6 ghosted neighbours of field f (T = float or double) at (x, y, z);
// `homogeneous` drops the outflow pressure (ghost value 0)
template <class T>
__device__ __forceinline__ T neighbour_sum(const T* __restrict__ f, const int x, const int y, const int z, const Dims d, const StencilBC& bc, const T centre, const bool homogeneous) {
    const int n[3] = {d.nx, d.ny, d.nz};
    T sum          = 0;
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const int a = fc >> 1;
        int q[3]    = {x, y, z};
        if (step_axis(q[a], (fc & 1) ? 1 : -1, n[a], (bc.periodic >> a) & 1)) {
            sum += f[lin(q[0], q[1], q[2], d)];
        } else {
            const FaceRule r = bc.face[fc];
            sum += r.mode == kGhostCopy ? centre : (homogeneous ? T(0) : T(r.value));
        }
    }
    return sum;
}

// constant part of the ghosts: sum of the outflow pressures around the cell
__device__ __forceinline__ double ghost_constant(const int x, const int y, const int z, const Dims d, const StencilBC& bc) {
    const int c[3] = {x, y, z}, n[3] = {d.nx, d.ny, d.nz};
    double s = 0.0;
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1;
        if ((bc.periodic >> a) & 1) continue;
        const bool at = (f & 1) ? c[a] == n[a] - 1 : c[a] == 0;
        if (at && bc.face[f].mode != kGhostCopy) s += bc.face[f].value;
    }
    return s;
}

__host__ __device__ inline uint64_t cells_of(const Dims d) { return static_cast<uint64_t>(d.nx) * d.ny * d.nz; }

#define GRID_STRIDE(i, n) for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < (n); i += static_cast<uint64_t>(gridDim.x) * kThreads)

// mean of the f^3 fine cells of coarse cell (x, y, z)
template <class T>
__device__ __forceinline__ double block_mean(const T* __restrict__ v, const int x, const int y, const int z, const int f, const Dims fd) {
    double s = 0.0;
    for (int k = 0; k < f; ++k)
        for (int j = 0; j < f; ++j)
            for (int i = 0; i < f; ++i) s += v[lin(f * x + i, f * y + j, f * z + k, fd)];
    return s / static_cast<double>(f * f * f);
}

// Eight lanes per coarse cell: lane s sums the z-slices k = s, s + 8, ... of its f^3
// block (rows of f consecutive cells, so the eight cells of a wave read contiguous
// rows), and the group adds its partial sums by shuffles.
constexpr int kGroup = 8;

__global__ __launch_bounds__(kThreads) void k_restrict_input(const float* __restrict__ div, float* __restrict__ in2, double* __restrict__ partials, const Dims cd, const int f, const StencilBC bc) {
    const Dims fd{f * cd.nx, f * cd.ny, f * cd.nz};
    const int lane_s   = static_cast<int>(threadIdx.x) & (kGroup - 1);
    const uint64_t n   = cells_of(cd);
    const uint64_t row = static_cast<uint64_t>(fd.nx), plane = row * static_cast<uint64_t>(fd.ny);
    double acc         = 0.0;
    for (uint64_t g = (static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x) / kGroup; g < n; g += static_cast<uint64_t>(gridDim.x) * (kThreads / kGroup)) {
        int x, y, z;
        decode(g, cd, x, y, z);
        double s = 0.0;
        for (int k = lane_s; k < f; k += kGroup)
            for (int j = 0; j < f; ++j) {
                const float* r = div + static_cast<uint64_t>(f * z + k) * plane + static_cast<uint64_t>(f * y + j) * row + static_cast<uint64_t>(f * x);
                // the row of f cells in 16-byte loads where f allows (same summation order)
                if (f == 8) {
                    const float4 a = reinterpret_cast<const float4*>(r)[0], b = reinterpret_cast<const float4*>(r)[1];
                    s += a.x; s += a.y; s += a.z; s += a.w;
                    s += b.x; s += b.y; s += b.z; s += b.w;
                } else if (f == 4) {
                    const float4 a = reinterpret_cast<const float4*>(r)[0];
                    s += a.x; s += a.y; s += a.z; s += a.w;
                } else {
                    for (int i = 0; i < f; ++i) s += r[i];
                }
            }
#pragma unroll
        for (int o = 1; o < kGroup; o <<= 1) s += __shfl_xor(s, o, kGroup);
        if (lane_s == 0) {
            const float m  = static_cast<float>(s / static_cast<double>(f * f * f));
            in2[2 * g]     = m;
            in2[2 * g + 1] = wall_mask(x, y, z, cd, bc);
            acc += static_cast<double>(m) * m;
        }
    }
    block_sum_to(acc, partials + blockIdx.x);
}

// Rows of the fine grid handled by this block: a contiguous chunk.
struct Rows {
    uint64_t r, end;
    int y, z;
};

__device__ __forceinline__ Rows block_rows(const int ny, const int nz) {
    const uint64_t rows = static_cast<uint64_t>(ny) * static_cast<uint64_t>(nz);
    const uint64_t per  = (rows + gridDim.x - 1) / gridDim.x;
    Rows w;
    w.r   = min(rows, static_cast<uint64_t>(blockIdx.x) * per);
    w.end = min(rows, w.r + per);
    w.y   = static_cast<int>(w.r % static_cast<uint64_t>(ny));
    w.z   = static_cast<int>(w.r / static_cast<uint64_t>(ny));
    return w;
}

// Cell-centred trilinear prolongation, separable: a block takes its fine rows in batches;
// the coarse line of every row of the batch (the coarse field interpolated at the row's y
// and z) goes to LDS first, one line per wave at a time and four lines' loads in flight
// (one line at a time per barrier left the block waiting out the load latency once per
// row), then every fine cell of the batch lerps its line in x. Coarse lines up to kLine.
constexpr int kLine = 2048;

__device__ __forceinline__ void axis_weights(const int c, const int n, const float inv_f, int& i0, int& i1, float& t) {
    const float cc = fminf(fmaxf((static_cast<float>(c) + 0.5f) * inv_f - 0.5f, 0.0f), static_cast<float>(n - 1));
    i0             = static_cast<int>(cc);
    i1             = min(i0 + 1, n - 1);
    t              = cc - static_cast<float>(i0);
}

__global__ __launch_bounds__(kThreads) void k_prolong(const float* __restrict__ q, const float* __restrict__ scal, const float scale, float* __restrict__ p0, const Dims cd, const int f) {
    constexpr int kWaves = kThreads / 64;
    __shared__ float lines[kLine];
    const Dims fd{f * cd.nx, f * cd.ny, f * cd.nz};
    const float amp    = scale * scal[0];
    const float inv_f  = 1.0f / static_cast<float>(f);
    const uint64_t cpl = static_cast<uint64_t>(cd.nx) * static_cast<uint64_t>(cd.ny);
    const int tid      = static_cast<int>(threadIdx.x);
    const int wave     = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int batch    = kLine / cd.nx;
    const Rows w0      = block_rows(fd.ny, fd.nz);
    for (uint64_t r0 = w0.r; r0 < w0.end; r0 += static_cast<uint64_t>(batch)) {
        const int nb = static_cast<int>(min(static_cast<uint64_t>(batch), w0.end - r0));
        const int yb = static_cast<int>(r0 % static_cast<uint64_t>(fd.ny)), zb = static_cast<int>(r0 / static_cast<uint64_t>(fd.ny));
        __syncthreads(); // the previous batch's lines are no longer read
        for (int cx0 = 0; cx0 < cd.nx; cx0 += 64) {
            const int cx  = cx0 + lane;
            const int cxl = min(cx, cd.nx - 1);
            for (int j0 = wave; j0 < nb; j0 += 4 * kWaves) {
                float a[4][4], wy[4], wz[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int sr = min(j0 + u * kWaves, nb - 1);
                    const int yy = yb + sr;
                    int y0, y1, z0, z1;
                    axis_weights(yy % fd.ny, cd.ny, inv_f, y0, y1, wy[u]);
                    axis_weights(zb + yy / fd.ny, cd.nz, inv_f, z0, z1, wz[u]);
                    const float* b = q + static_cast<uint64_t>(cxl);
                    a[u][0]        = b[static_cast<uint64_t>(z0) * cpl + static_cast<uint64_t>(y0) * cd.nx];
                    a[u][1]        = b[static_cast<uint64_t>(z0) * cpl + static_cast<uint64_t>(y1) * cd.nx];
                    a[u][2]        = b[static_cast<uint64_t>(z1) * cpl + static_cast<uint64_t>(y0) * cd.nx];
                    a[u][3]        = b[static_cast<uint64_t>(z1) * cpl + static_cast<uint64_t>(y1) * cd.nx];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int sr   = j0 + u * kWaves;
                    const float ty = wy[u], tz = wz[u];
                    if (sr < nb && cx < cd.nx) lines[sr * cd.nx + cx] = ((1.0f - ty) * a[u][0] + ty * a[u][1]) * (1.0f - tz) + ((1.0f - ty) * a[u][2] + ty * a[u][3]) * tz;
                }
            }
        }
        __syncthreads();
        for (int x = tid; x < fd.nx; x += kThreads) {
            int x0, x1;
            float tx;
            axis_weights(x, cd.nx, inv_f, x0, x1, tx);
            float* out = p0 + r0 * static_cast<uint64_t>(fd.nx) + static_cast<uint64_t>(x);
            for (int sr = 0; sr < nb; ++sr) {
                const float* l = lines + sr * cd.nx;
                out[static_cast<uint64_t>(sr) * static_cast<uint64_t>(fd.nx)] = amp * ((1.0f - tx) * l[x0] + tx * l[x1]);
            }
        }
    }
}

// Residual sums of the guess and of the zero guess. A block takes one row y and a chunk of
// zc planes and marches up z with the lane's column in registers, so each plane of p0 is
// fetched once from HBM; the y -+ 1 rows come from the L2 the neighbouring blocks share
// (blocks renumbered so that consecutive rows run on one XCD) and x -+ 1 from the lane's
// own cache lines. Every cell, on a face or not, is summed in the march: a neighbour
// across a periodic boundary is loaded from its wrapped position, a ghost takes the
// homogeneous rule (copy: the centre value, else 0) and the outflow pressures around the
// cell enter b. This is neighbour_sum and ghost_constant cell by cell, in their face order.
constexpr int kAhead = 2; // planes of loads in flight (deeper measured slower)
static_assert(kAhead >= 2, "the sum reads planes z - 1 .. z + 1 from the column ring");
constexpr int kSlots = kAhead + 2; // planes z - 1 .. z + kAhead of the column are live

__global__ __launch_bounds__(kThreads) void k_guess_partials(const float* __restrict__ p0, const float* __restrict__ div, const float h2, double* __restrict__ partials, const Dims d, const StencilBC bc, const int zc) {
    double acc_g = 0.0, acc_z = 0.0;
    const uint64_t sy = static_cast<uint64_t>(d.nx);
    const int n = static_cast<int>(gridDim.x), bid = static_cast<int>(blockIdx.x);
    const int wid = (n & 7) == 0 ? (bid & 7) * (n >> 3) + (bid >> 3) : bid;
    const int y   = wid % d.ny;
    const int z0  = (wid / d.ny) * zc;
    const int z1  = min(z0 + zc, d.nz);
    const bool per_x = bc.periodic & 1, per_y = (bc.periodic >> 1) & 1, per_z = (bc.periodic >> 2) & 1;
    // homogeneous ghost of face f: the centre value for copy rules, else 0
    bool copy[6];
    // outflow pressure of face f, counted where the cell touches that (non-periodic) face
    double g_val[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        copy[f]  = bc.face[f].mode == kGhostCopy;
        g_val[f] = ((bc.periodic >> (f >> 1)) & 1) || copy[f] ? 0.0 : static_cast<double>(bc.face[f].value);
    }
    // neighbour rows: wrapped on a periodic axis, -1 for a ghost
    int ym = y, yp = y;
    const bool has_ym = step_axis(ym, -1, d.ny, per_y), has_yp = step_axis(yp, 1, d.ny, per_y);
    const uint64_t o_y = static_cast<uint64_t>(y) * sy, o_ym = static_cast<uint64_t>(has_ym ? ym : y) * sy, o_yp = static_cast<uint64_t>(has_yp ? yp : y) * sy;
    const uint64_t plane = static_cast<uint64_t>(d.ny) * sy;
    // plane p of the march (p may run from z0 - 1 to z1 + kAhead - 1): wrapped on a periodic
    // z axis, else clamped (a clamped plane is never summed)
    auto plane_of = [&](const int p) { return per_z ? wrap_index(p, d.nz) : min(max(p, 0), d.nz - 1); };
    for (int x = static_cast<int>(threadIdx.x); x < d.nx; x += kThreads) {
        int xl = x, xr = x;
        const bool has_xl = step_axis(xl, -1, d.nx, per_x), has_xr = step_axis(xr, 1, d.nx, per_x);
        if (!has_xl) xl = x;
        if (!has_xr) xr = x;
        // ghost_constant's order: x faces, y faces, z faces (each term exact in double)
        const double g_xy = (((0.0 + (x == 0 ? g_val[0] : 0.0)) + (x == d.nx - 1 ? g_val[1] : 0.0)) + (y == 0 ? g_val[2] : 0.0)) + (y == d.ny - 1 ? g_val[3] : 0.0);
        // Register rings of kSlots slots, unrolled so that every slot is a compile-time
        // register set: the centre column of plane p sits in slot (p - z0 + 1) % kSlots,
        // the other taps of plane p in slot (p - z0) % kSlots. The loads of plane
        // z + kAhead are issued while plane z is summed. (Shifting registers from slot
        // to slot instead makes the compiler wait for every load in flight: the moves
        // read the registers being loaded.)
        struct Taps {
            float ym, yp, xm, xp, dv;
        };
        auto col_of = [&](const int p) { return p0[static_cast<uint64_t>(plane_of(p)) * plane + o_y + x]; };
        auto taps   = [&](const int p) {
            const uint64_t r = static_cast<uint64_t>(plane_of(p)) * plane;
            return Taps{p0[r + o_ym + x], p0[r + o_yp + x], p0[r + o_y + xl], p0[r + o_y + xr], div[r + o_y + x]};
        };
        float col[kSlots];
        Taps t[kSlots];
#pragma unroll
        for (int j = 0; j <= kAhead; ++j) col[j] = col_of(z0 + j - 1);
#pragma unroll
        for (int i = 0; i < kAhead; ++i) t[i] = taps(z0 + i);
        // plane z, whose ring slots are those of j = (z - z0) % kSlots
        auto step = [&](const int z, const int j) __attribute__((always_inline)) {
            col[(j + kAhead + 1) % kSlots] = col_of(z + kAhead);
            t[(j + kAhead) % kSlots]       = taps(z + kAhead);
            const float c  = col[(j + 1) % kSlots];
            const Taps& tz = t[j];
            const bool has_zm = z > 0 || per_z, has_zp = z < d.nz - 1 || per_z;
            float sum = has_xl ? tz.xm : (copy[0] ? c : 0.0f);
            sum += has_xr ? tz.xp : (copy[1] ? c : 0.0f);
            sum += has_ym ? tz.ym : (copy[2] ? c : 0.0f);
            sum += has_yp ? tz.yp : (copy[3] ? c : 0.0f);
            sum += has_zm ? col[j] : (copy[4] ? c : 0.0f);
            sum += has_zp ? col[(j + 2) % kSlots] : (copy[5] ? c : 0.0f);
            const double g  = (g_xy + (z == 0 ? g_val[4] : 0.0)) + (z == d.nz - 1 ? g_val[5] : 0.0);
            const double b  = -static_cast<double>(h2) * tz.dv + g;
            const double ap = 6.0 * c - static_cast<double>(sum);
            acc_g += (b - ap) * (b - ap);
            acc_z += b * b;
        };
        for (int zb = z0; zb < z1; zb += kSlots) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                if (zb + j >= z1) break;
                step(zb + j, j);
            }
        }
    }
    block_sum_to(acc_g, partials + 2 * blockIdx.x);
    block_sum_to(acc_z, partials + 2 * blockIdx.x + 1);
}

__global__ __launch_bounds__(kThreads) void k_guess_decide(const double* __restrict__ partials, const int n_partials, const double ratio, double* __restrict__ stats, int* __restrict__ skip) {
    // four independent sums per thread, so that four pairs of loads are in flight
    double g[4] = {0.0, 0.0, 0.0, 0.0}, zz[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < n_partials; b += 4 * kThreads) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int bu = b + u * kThreads;
            if (bu < n_partials) {
                g[u] += partials[2 * bu];
                zz[u] += partials[2 * bu + 1];
            }
        }
    }
    __shared__ double tg, tz;
    block_sum_to((g[0] + g[1]) + (g[2] + g[3]), &tg);
    block_sum_to((zz[0] + zz[1]) + (zz[2] + zz[3]), &tz);
    if (threadIdx.x == 0) {
        const double rg = sqrt(tg), rz = sqrt(tz);
        const bool fallback = !(rg <= ratio * rz); // NaN guesses fall back too
        stats[0]            = rg;
        stats[1]            = rz;
        stats[2]            = fallback ? 1.0 : 0.0;
        skip[0]             = fallback ? 0 : 1;
        skip[2]             = 0; // failure flag of this step's one-launch fallback solve
    }
}

__global__ __launch_bounds__(kThreads) void k_zero_on_fallback(float* __restrict__ p0, const double* __restrict__ stats, const uint64_t n) {
    if (stats[2] == 0.0) return;
    GRID_STRIDE(i, n) p0[i] = 0.0f;
}

// ---- conjugate gradients (FP64) ----

// r = b (= -h2 div + outflow constants); partials: sum b, sum b^2
__global__ __launch_bounds__(kThreads) void k_cg_rhs(const float* __restrict__ div, const float h2, double* __restrict__ r, double* __restrict__ partials, const Dims d, const StencilBC bc) {
    double s = 0.0, s2 = 0.0;
    GRID_STRIDE(i, cells_of(d)) {
        int x, y, z;
        decode(i, d, x, y, z);
        const double b = -static_cast<double>(h2) * div[i] + ghost_constant(x, y, z, d, bc);
        r[i]           = b;
        s += b;
        s2 += b * b;
    }
    block_sum_to(s, partials + 2 * blockIdx.x);
    block_sum_to(s2, partials + 2 * blockIdx.x + 1);
}

// scal[slot] = sum of partials[stride * b + offset]
__global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ partials, const int n_partials, const int stride, const int offset, double* __restrict__ scal, const int slot) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < n_partials; b += kThreads) acc += partials[stride * b + offset];
    __shared__ double t;
    block_sum_to(acc, &t);
    if (threadIdx.x == 0) scal[slot] = t;
}

// v -= scal[slot] / n
__global__ __launch_bounds__(kThreads) void k_sub_mean(double* __restrict__ v, const double* __restrict__ scal, const int slot, const uint64_t n) {
    const double m = scal[slot] / static_cast<double>(n);
    GRID_STRIDE(i, n) v[i] -= m;
}

// r -= A x (homogeneous A); then p = r; partial r.r
__global__ __launch_bounds__(kThreads) void k_cg_init(const double* __restrict__ x, double* __restrict__ r, double* __restrict__ p, double* __restrict__ partials, const Dims d, const StencilBC bc) {
    double acc = 0.0;
    GRID_STRIDE(i, cells_of(d)) {
        int xx, y, z;
        decode(i, d, xx, y, z);
        const double ax = 6.0 * x[i] - neighbour_sum<double>(x, xx, y, z, d, bc, x[i], true);
        const double ri = r[i] - ax;
        r[i]            = ri;
        p[i]            = ri;
        acc += ri * ri;
    }
    block_sum_to(acc, partials + blockIdx.x);
}

// ap = A p; partial p.ap
__global__ __launch_bounds__(kThreads) void k_cg_apply(const double* __restrict__ p, double* __restrict__ ap, double* __restrict__ partials, const Dims d, const StencilBC bc) {
    double acc = 0.0;
    GRID_STRIDE(i, cells_of(d)) {
        int x, y, z;
        decode(i, d, x, y, z);
        const double v = 6.0 * p[i] - neighbour_sum<double>(p, x, y, z, d, bc, p[i], true);
        ap[i]          = v;
        acc += p[i] * v;
    }
    block_sum_to(acc, partials + blockIdx.x);
}

// alpha = rr / pAp; x += alpha p; r -= alpha ap; partial r.r
__global__ __launch_bounds__(kThreads) void k_cg_update(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p, const double* __restrict__ ap, const double* __restrict__ scal, double* __restrict__ partials, const uint64_t n) {
    const double pap   = scal[1];
    const double alpha = pap > 0.0 ? scal[0] / pap : 0.0;
    double acc         = 0.0;
    GRID_STRIDE(i, n) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * ap[i];
        r[i]            = ri;
        acc += ri * ri;
    }
    block_sum_to(acc, partials + blockIdx.x);
}

// beta = rr_new / rr; p = r + beta p; rr = rr_new (one block does the bookkeeping)
__global__ __launch_bounds__(kThreads) void k_cg_direction(double* __restrict__ p, const double* __restrict__ r, double* __restrict__ scal, const uint64_t n) {
    const double rr = scal[0], rr_new = scal[2];
    const double beta = rr > 0.0 ? rr_new / rr : 0.0;
    GRID_STRIDE(i, n) p[i] = r[i] + beta * p[i];
}
__global__ void k_cg_shift(double* __restrict__ scal) { scal[0] = scal[2]; }

// partials[2b] = sum v, partials[2b + 1] = sum v^2
__global__ __launch_bounds__(kThreads) void k_moments(const double* __restrict__ v, double* __restrict__ partials, const uint64_t n) {
    double s = 0.0, s2 = 0.0;
    GRID_STRIDE(i, n) {
        s += v[i];
        s2 += v[i] * v[i];
    }
    block_sum_to(s, partials + 2 * blockIdx.x);
    block_sum_to(s2, partials + 2 * blockIdx.x + 1);
}

// partial sums of the squared coarse divergence (mean of f^3 cells)
__global__ __launch_bounds__(kThreads) void k_coarse_sum_sq(const float* __restrict__ div, double* __restrict__ partials, const Dims cd, const int f) {
    const Dims fd{f * cd.nx, f * cd.ny, f * cd.nz};
    double acc = 0.0;
    GRID_STRIDE(v, cells_of(cd)) {
        int x, y, z;
        decode(v, cd, x, y, z);
        const double m = block_mean(div, x, y, z, f, fd);
        acc += m * m;
    }
    block_sum_to(acc, partials + blockIdx.x);
}

// coarse pair: in2 = (div_c / sigma_c, wall mask), target = x_c / (h_c^2 sigma_c)
__global__ __launch_bounds__(kThreads) void k_pair(const float* __restrict__ div, const double* __restrict__ x, const double hc2, float* __restrict__ in2, float* __restrict__ target, const double* __restrict__ scal, const Dims cd, const int f, const StencilBC bc) {
    const Dims fd{f * cd.nx, f * cd.ny, f * cd.nz};
    const double sigma = sqrt(scal[5] / static_cast<double>(cells_of(cd)));
    const double inv   = sigma > 0.0 ? 1.0 / sigma : 0.0;
    GRID_STRIDE(v, cells_of(cd)) {
        int xx, y, z;
        decode(v, cd, xx, y, z);
        in2[2 * v]     = static_cast<float>(block_mean(div, xx, y, z, f, fd) * inv);
        in2[2 * v + 1] = wall_mask(xx, y, z, cd, bc);
        target[v]      = static_cast<float>(block_mean(x, xx, y, z, f, fd) * inv / hc2);
    }
}

#undef GRID_STRIDE

unsigned blocks_for(const uint64_t n) {
    const uint64_t b = (n + kThreads - 1) / kThreads;
    return static_cast<unsigned>(b < kPartialBlocks ? (b ? b : 1) : kPartialBlocks);
}

bool singular(const Dims, const StencilBC& bc) {
    for (int f = 0; f < 6; ++f)
        if (!((bc.periodic >> (f >> 1)) & 1) && bc.face[f].mode != kGhostCopy) return false;
    return true;
}

} // namespace

unsigned restrict_partials(const Dims cd) { return blocks_for(cells_of(cd) * kGroup); }

hipError_t launch_restrict_input(hipStream_t s, const float* div, float* in2, double* partials, const Dims cd, const int f, const StencilBC& p_bc) {
    k_restrict_input<<<restrict_partials(cd), kThreads, 0, s>>>(div, in2, partials, cd, f, p_bc);
    return hipGetLastError();
}

hipError_t launch_prolong(hipStream_t s, const float* q, const float* scal, const float scale, float* p0, const Dims cd, const int f) {
    const Dims fd{f * cd.nx, f * cd.ny, f * cd.nz};
    if (cd.nx > kLine) return hipErrorInvalidValue;
    // about 32 fine rows per block: one round of blocks on the GPU, each streaming its
    // stores after one batch of line loads
    const uint64_t rows = static_cast<uint64_t>(fd.ny) * static_cast<uint64_t>(fd.nz);
    const unsigned b    = static_cast<unsigned>(std::min<uint64_t>(std::max<uint64_t>((rows + 31) / 32, 1), kPartialBlocks));
    k_prolong<<<b, kThreads, 0, s>>>(q, scal, scale, p0, cd, f);
    return hipGetLastError();
}

hipError_t launch_guess_test(hipStream_t s, float* p0, const float* div, const float h2, const double ratio, double* partials, double* stats, int* skip, const Dims d, const StencilBC& p_bc, const bool zero_on_fallback) {
    // 16-plane chunks, longer where the rows times chunks would exceed the partial slots
    int zc = 16;
    while (zc < d.nz && static_cast<uint64_t>(d.ny) * static_cast<uint64_t>((d.nz + zc - 1) / zc) > static_cast<uint64_t>(kPartialBlocks)) zc *= 2;
    const uint64_t nb = static_cast<uint64_t>(d.ny) * static_cast<uint64_t>((d.nz + zc - 1) / zc);
    if (nb > static_cast<uint64_t>(kPartialBlocks)) return hipErrorInvalidValue;
    const unsigned b = static_cast<unsigned>(nb);
    k_guess_partials<<<b, kThreads, 0, s>>>(p0, div, h2, partials, d, p_bc, zc);
    k_guess_decide<<<1, kThreads, 0, s>>>(partials, static_cast<int>(b), ratio, stats, skip);
    if (zero_on_fallback) k_zero_on_fallback<<<b, kThreads, 0, s>>>(p0, stats, cells_of(d));
    return hipGetLastError();
}

hipError_t cg_solve(hipStream_t s, const float* div, const float h2, double* x, double* r, double* p, double* ap, double* partials, double* scal, const Dims d, const StencilBC& p_bc, const double tol, const int max_iter, CgResult* out) {
    constexpr int kCheck = 10;
    const uint64_t n     = cells_of(d);
    const unsigned b     = blocks_for(n);
    const int nb         = static_cast<int>(b);
    const bool sing      = singular(d, p_bc);
    // r = b, projected to zero mean when the system is singular; ||b||^2 -> scal[3]
    k_cg_rhs<<<b, kThreads, 0, s>>>(div, h2, r, partials, d, p_bc);
    if (sing) {
        k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 2, 0, scal, 4);
        k_sub_mean<<<b, kThreads, 0, s>>>(r, scal, 4, n);
        k_moments<<<b, kThreads, 0, s>>>(r, partials, n);
    }
    k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 2, 1, scal, 3);
    // r = b - A x, p = r, rr -> scal[0]
    k_cg_init<<<b, kThreads, 0, s>>>(x, r, p, partials, d, p_bc);
    k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 1, 0, scal, 0);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;

    double host[4] = {};
    auto fetch = [&]() -> bool {
        if (hipMemcpyAsync(host, scal, sizeof(host), hipMemcpyDeviceToHost, s) != hipSuccess) return false;
        return hipStreamSynchronize(s) == hipSuccess;
    };
    int it = 0;
    for (;; ++it) {
        if (it % kCheck == 0 || it == max_iter) {
            if (!fetch()) return hipErrorUnknown;
            if (!(host[0] > tol * tol * host[3]) || it >= max_iter) break;
        }
        k_cg_apply<<<b, kThreads, 0, s>>>(p, ap, partials, d, p_bc);
        k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 1, 0, scal, 1);
        k_cg_update<<<b, kThreads, 0, s>>>(x, r, p, ap, scal, partials, n);
        k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 1, 0, scal, 2);
        k_cg_direction<<<b, kThreads, 0, s>>>(p, r, scal, n);
        k_cg_shift<<<1, 1, 0, s>>>(scal);
        if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    }
    if (sing) {
        k_moments<<<b, kThreads, 0, s>>>(x, partials, n);
        k_reduce<<<1, kThreads, 0, s>>>(partials, nb, 2, 0, scal, 4);
        k_sub_mean<<<b, kThreads, 0, s>>>(x, scal, 4, n);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return hipErrorUnknown;
    out->iterations   = it;
    out->rel_residual = host[3] > 0.0 ? sqrt(host[0] / host[3]) : 0.0;
    return hipSuccess;
}

hipError_t launch_pair(hipStream_t s, const float* div, const double* x, const float h2, float* in2, float* target, double* partials, double* scal, const Dims d, const int f, const StencilBC& p_bc) {
    if (f < 1 || d.nx % f || d.ny % f || d.nz % f) return hipErrorInvalidValue;
    const Dims cd{d.nx / f, d.ny / f, d.nz / f};
    const unsigned b = blocks_for(cells_of(cd));
    k_coarse_sum_sq<<<b, kThreads, 0, s>>>(div, partials, cd, f);
    k_reduce<<<1, kThreads, 0, s>>>(partials, static_cast<int>(b), 1, 0, scal, 5);
    k_pair<<<b, kThreads, 0, s>>>(div, x, static_cast<double>(h2) * f * f, in2, target, scal, cd, f, p_bc);
    return hipGetLastError();
}

} // namespace sf::warm
