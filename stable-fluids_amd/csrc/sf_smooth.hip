// sf_smooth.hip -- z-marching LDS-tiled smoother passes for gfx950.
//
// One pass = one red-black Gauss-Seidel iteration (or one Jacobi sweep) of
//   diffusion: x = (b + c * sum6) / (1 + 6c)      (reference backend-cuda.cu:354-375)
//   Poisson:   x = (sum6 - c * b) / 6             (reference backend-cuda.cu:398-407)
// read from x_old (or zeros) and b, written to x_new: 12 B/cell of compulsory traffic.
//
// Work decomposition: a 256-thread workgroup owns a kTX x kTY column of cells in x-y and
// a chunk of zc planes, and marches upward in z. For RBGS, plane k of the half-updated
// iterate H (red cells relaxed from the old iterate, black cells old) is computed on the
// tile plus a 1-cell halo from an LDS ring of old planes carrying a 2-cell halo; the
// final plane k-1 (black cells relaxed from H) is then computed on the tile. Red values
// on the halo are recomputed redundantly, so tiles never exchange data within a pass.
// Old and b planes for the next step are fetched into registers before the current
// plane is computed, hiding HBM latency behind the LDS work.
//
// Boundary handling: positions outside the grid on a periodic axis are loaded from the
// wrapped cell (so LDS neighbours are always the right values); on other axes the
// neighbour is replaced by the face's ghost rule applied to the centre value, decided
// per position from its global coordinates. Red/black parity uses wrapped coordinates.
#include "sf_kernels.hpp"

namespace sf {

namespace {

constexpr int kTX      = 64;
constexpr int kTY      = 16;
constexpr int kThreads = 256;
constexpr int kOX      = kTX + 4; // old iterate region (halo 2)
constexpr int kOY      = kTY + 4;
constexpr int kHX      = kTX + 2; // half-updated iterate and b region (halo 1)
constexpr int kHY      = kTY + 2;
constexpr int kOldN    = kOX * kOY;
constexpr int kHN      = kHX * kHY;
constexpr int kOldPer  = (kOldN + kThreads - 1) / kThreads;
constexpr int kHPer    = (kHN + kThreads - 1) / kThreads;
constexpr int kOutPer  = kTX * kTY / kThreads;
static_assert(kTX == 64 && kTX * kTY % kThreads == 0, "one wavefront per output row");

struct TileGeom {
    int tiles_x, tiles_y, chunks, zc;
};

__device__ __forceinline__ int fix_coord(const int c, const int n, const bool periodic) {
    if (periodic) return wrap_index(c, n);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

template <int kSolve>
__device__ __forceinline__ float relax_value(const float bv, const float sum, const float coef) {
    if constexpr (kSolve == kSolvePoisson)
        return (sum - coef * bv) / 6.0f;
    else
        return (bv + coef * sum) / (1.0f + 6.0f * coef);
}

// Workgroup -> (tile x, tile y, z chunk). Workgroups are dispatched round-robin over the
// 8 XCDs; renumbering gives each XCD (and its L2) a contiguous run of chunks and tiles,
// so z-halo planes and x-y halo columns are mostly shared within one L2.
__device__ __forceinline__ void tile_of(const TileGeom g, int& tx, int& ty, int& cz) {
    const int n = static_cast<int>(gridDim.x);
    int w       = static_cast<int>(blockIdx.x);
    if ((n & 7) == 0) w = (w & 7) * (n >> 3) + (w >> 3);
    cz          = w % g.chunks;
    const int t = w / g.chunks;
    tx          = t % g.tiles_x;
    ty          = t / g.tiles_x;
}

__device__ __forceinline__ int mod3(const int k) { return ((k % 3) + 3) % 3; }

template <int kSolve, bool kZero>
__global__ __launch_bounds__(kThreads) void k_rbgs_pass(const float* __restrict__ x_old, const float* __restrict__ b, float* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC bc, const TileGeom g) {
    __shared__ float s_old[3][kOldN];
    __shared__ float s_h[3][kHN];
    __shared__ float s_b[3][kHN];

    int tx, ty, cz;
    tile_of(g, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int x0   = tx * kTX, y0 = ty * kTY;
    const int z0   = cz * g.zc;
    const int z1   = min(z0 + g.zc, d.nz);
    const bool px  = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint64_t plane = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny);

    // in-plane global offsets of this thread's load positions
    int old_off[kOldPer];
#pragma unroll
    for (int i = 0; i < kOldPer; ++i) {
        const int e = tid + i * kThreads;
        old_off[i]  = -1;
        if (e < kOldN) {
            const int gx = fix_coord(x0 - 2 + e % kOX, d.nx, px);
            const int gy = fix_coord(y0 - 2 + e / kOX, d.ny, py);
            old_off[i]   = gy * d.nx + gx;
        }
    }
    // H-stage positions: b offset, validity, parity of (x+y), ghost faces in x/y
    int h_off[kHPer];
    int h_info[kHPer]; // bit 0..3: ghost x-,x+,y-,y+; bit 4: parity(x+y); bit 5: valid
#pragma unroll
    for (int i = 0; i < kHPer; ++i) {
        const int e = tid + i * kThreads;
        h_off[i]    = -1;
        h_info[i]   = 0;
        if (e < kHN) {
            const int gx     = x0 - 1 + e % kHX;
            const int gy     = y0 - 1 + e / kHX;
            const bool valid = (px || (gx >= 0 && gx < d.nx)) && (py || (gy >= 0 && gy < d.ny));
            const int wx     = fix_coord(gx, d.nx, px);
            const int wy     = fix_coord(gy, d.ny, py);
            h_off[i]         = wy * d.nx + wx;
            int info         = 0;
            if (!px && gx - 1 < 0) info |= 1;
            if (!px && gx + 1 >= d.nx) info |= 2;
            if (!py && gy - 1 < 0) info |= 4;
            if (!py && gy + 1 >= d.ny) info |= 8;
            info |= ((wx + wy) & 1) << 4;
            if (valid) info |= 32;
            h_info[i] = info;
        }
    }
    // output positions: one row of 64 per wavefront, rows tid/64 + 4*i
    const int fx = tid & (kTX - 1);
    int f_info[kOutPer]; // ghost bits, parity(x+y) in bit 4, in-grid in bit 5
#pragma unroll
    for (int i = 0; i < kOutPer; ++i) {
        const int fy = (tid >> 6) + 4 * i;
        const int gx = x0 + fx, gy = y0 + fy;
        int info     = 0;
        if (!px && gx - 1 < 0) info |= 1;
        if (!px && gx + 1 >= d.nx) info |= 2;
        if (!py && gy - 1 < 0) info |= 4;
        if (!py && gy + 1 >= d.ny) info |= 8;
        info |= ((gx + gy) & 1) << 4;
        if (gx < d.nx && gy < d.ny) info |= 32;
        f_info[i] = info;
    }

    auto plane_valid = [&](const int k) { return pz || (k >= 0 && k < d.nz); };
    auto plane_base  = [&](const int k) { return static_cast<uint64_t>(fix_coord(k, d.nz, pz)) * plane; };

    float r_old[kOldPer], r_b[kHPer];
    auto fetch = [&](const int k_old, const int k_b) {
        if (!kZero && plane_valid(k_old)) {
            const uint64_t base = plane_base(k_old);
#pragma unroll
            for (int i = 0; i < kOldPer; ++i) r_old[i] = old_off[i] >= 0 ? x_old[base + old_off[i]] : 0.0f;
        } else {
#pragma unroll
            for (int i = 0; i < kOldPer; ++i) r_old[i] = 0.0f;
        }
        if (plane_valid(k_b)) {
            const uint64_t base = plane_base(k_b);
#pragma unroll
            for (int i = 0; i < kHPer; ++i) r_b[i] = h_off[i] >= 0 ? b[base + h_off[i]] : 0.0f;
        }
    };
    auto stash_old = [&](const int k) {
        float* s = s_old[mod3(k)];
#pragma unroll
        for (int i = 0; i < kOldPer; ++i)
            if (old_off[i] >= 0) s[tid + i * kThreads] = r_old[i];
    };
    auto stash_b = [&](const int k) {
        float* s = s_b[mod3(k)];
#pragma unroll
        for (int i = 0; i < kHPer; ++i)
            if (h_off[i] >= 0) s[tid + i * kThreads] = r_b[i];
    };

    // prologue: old planes z0-2 and z0-1 in LDS, old z0 and b z0-1 in registers
    fetch(z0 - 2, z0 - 2);
    stash_old(z0 - 2);
    fetch(z0 - 1, z0 - 1);
    stash_old(z0 - 1);
    fetch(z0, z0 - 1);

    for (int k = z0 - 1; k <= z1; ++k) {
        stash_old(k + 1); // registers hold old(k+1) and b(k)
        stash_b(k);
        if (k + 1 <= z1) fetch(k + 2, k + 1);
        __syncthreads();

        // H(k) on tile + 1 halo
        if (plane_valid(k)) {
            const float* so = s_old[mod3(k)];
            const float* sm = s_old[mod3(k - 1)];
            const float* sp = s_old[mod3(k + 1)];
            const float* sb = s_b[mod3(k)];
            float* sh       = s_h[mod3(k)];
            const int wz    = fix_coord(k, d.nz, pz);
            const bool zlo  = !pz && k - 1 < 0;
            const bool zhi  = !pz && k + 1 >= d.nz;
#pragma unroll
            for (int i = 0; i < kHPer; ++i) {
                const int e    = tid + i * kThreads;
                const int info = h_info[i];
                if (e >= kHN || !(info & 32)) continue;
                const int c        = (e / kHX + 1) * kOX + (e % kHX + 1);
                const float centre = so[c];
                float out          = centre;
                if ((((info >> 4) + wz) & 1) == 0) {
                    float sum = (info & 1) ? apply_ghost(bc.face[0], centre) : so[c - 1];
                    sum       = sum + ((info & 2) ? apply_ghost(bc.face[1], centre) : so[c + 1]);
                    sum       = sum + ((info & 4) ? apply_ghost(bc.face[2], centre) : so[c - kOX]);
                    sum       = sum + ((info & 8) ? apply_ghost(bc.face[3], centre) : so[c + kOX]);
                    sum       = sum + (zlo ? apply_ghost(bc.face[4], centre) : sm[c]);
                    sum       = sum + (zhi ? apply_ghost(bc.face[5], centre) : sp[c]);
                    out       = relax_value<kSolve>(sb[e], sum, coef);
                }
                sh[e] = out;
            }
        }
        __syncthreads();

        // final plane k-1 on the tile
        const int kf = k - 1;
        if (kf >= z0 && kf < z1) {
            const float* hm = s_h[mod3(kf - 1)];
            const float* h0 = s_h[mod3(kf)];
            const float* hp = s_h[mod3(kf + 1)];
            const float* sb = s_b[mod3(kf)];
            const bool zlo  = !pz && kf - 1 < 0;
            const bool zhi  = !pz && kf + 1 >= d.nz;
            const uint64_t base = static_cast<uint64_t>(kf) * plane;
#pragma unroll
            for (int i = 0; i < kOutPer; ++i) {
                const int info = f_info[i];
                if (!(info & 32)) continue;
                const int fy       = (tid >> 6) + 4 * i;
                const int c        = (fy + 1) * kHX + (fx + 1);
                const float centre = h0[c];
                float out          = centre;
                if ((((info >> 4) + kf) & 1) == 1) {
                    float sum = (info & 1) ? apply_ghost(bc.face[0], centre) : h0[c - 1];
                    sum       = sum + ((info & 2) ? apply_ghost(bc.face[1], centre) : h0[c + 1]);
                    sum       = sum + ((info & 4) ? apply_ghost(bc.face[2], centre) : h0[c - kHX]);
                    sum       = sum + ((info & 8) ? apply_ghost(bc.face[3], centre) : h0[c + kHX]);
                    sum       = sum + (zlo ? apply_ghost(bc.face[4], centre) : hm[c]);
                    sum       = sum + (zhi ? apply_ghost(bc.face[5], centre) : hp[c]);
                    out       = relax_value<kSolve>(sb[c], sum, coef);
                }
                x_new[base + static_cast<uint64_t>(y0 + fy) * d.nx + (x0 + fx)] = out * out_scale;
            }
        }
        __syncthreads();
    }
}

// Jacobi: every cell from the old iterate; old planes carry a 1-cell halo, b is read
// at the centre only.
constexpr int kJOldN   = kHX * kHY;
constexpr int kJOldPer = (kJOldN + kThreads - 1) / kThreads;

template <int kSolve, bool kZero>
__global__ __launch_bounds__(kThreads) void k_jacobi_pass(const float* __restrict__ x_old, const float* __restrict__ b, float* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC bc, const TileGeom g) {
    __shared__ float s_old[3][kJOldN];

    int tx, ty, cz;
    tile_of(g, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int x0   = tx * kTX, y0 = ty * kTY;
    const int z0   = cz * g.zc;
    const int z1   = min(z0 + g.zc, d.nz);
    const bool px  = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint64_t plane = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny);

    int old_off[kJOldPer];
#pragma unroll
    for (int i = 0; i < kJOldPer; ++i) {
        const int e = tid + i * kThreads;
        old_off[i]  = -1;
        if (e < kJOldN) {
            const int gx = fix_coord(x0 - 1 + e % kHX, d.nx, px);
            const int gy = fix_coord(y0 - 1 + e / kHX, d.ny, py);
            old_off[i]   = gy * d.nx + gx;
        }
    }
    const int fx = tid & (kTX - 1);
    int f_info[kOutPer];
    int f_off[kOutPer];
#pragma unroll
    for (int i = 0; i < kOutPer; ++i) {
        const int fy = (tid >> 6) + 4 * i;
        const int gx = x0 + fx, gy = y0 + fy;
        int info     = 0;
        if (!px && gx - 1 < 0) info |= 1;
        if (!px && gx + 1 >= d.nx) info |= 2;
        if (!py && gy - 1 < 0) info |= 4;
        if (!py && gy + 1 >= d.ny) info |= 8;
        if (gx < d.nx && gy < d.ny) info |= 32;
        f_info[i] = info;
        f_off[i]  = (info & 32) ? gy * d.nx + gx : 0;
    }
    auto plane_valid = [&](const int k) { return pz || (k >= 0 && k < d.nz); };
    auto plane_base  = [&](const int k) { return static_cast<uint64_t>(fix_coord(k, d.nz, pz)) * plane; };

    float r_old[kJOldPer], r_b[kOutPer];
    auto fetch_old = [&](const int k) {
        if (!kZero && plane_valid(k)) {
            const uint64_t base = plane_base(k);
#pragma unroll
            for (int i = 0; i < kJOldPer; ++i) r_old[i] = old_off[i] >= 0 ? x_old[base + old_off[i]] : 0.0f;
        } else {
#pragma unroll
            for (int i = 0; i < kJOldPer; ++i) r_old[i] = 0.0f;
        }
    };
    auto fetch_b = [&](const int k) {
        if (k >= 0 && k < d.nz) {
            const uint64_t base = static_cast<uint64_t>(k) * plane;
#pragma unroll
            for (int i = 0; i < kOutPer; ++i) r_b[i] = b[base + f_off[i]];
        }
    };
    auto stash_old = [&](const int k) {
        float* s = s_old[mod3(k)];
#pragma unroll
        for (int i = 0; i < kJOldPer; ++i)
            if (old_off[i] >= 0) s[tid + i * kThreads] = r_old[i];
    };

    fetch_old(z0 - 1);
    stash_old(z0 - 1);
    fetch_old(z0);
    stash_old(z0);
    fetch_old(z0 + 1);
    fetch_b(z0);
    for (int k = z0; k < z1; ++k) {
        stash_old(k + 1);
        float bv[kOutPer];
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) bv[i] = r_b[i];
        if (k + 1 < z1) {
            fetch_old(k + 2);
            fetch_b(k + 1);
        }
        __syncthreads();
        const float* so = s_old[mod3(k)];
        const float* sm = s_old[mod3(k - 1)];
        const float* sp = s_old[mod3(k + 1)];
        const bool zlo  = !pz && k - 1 < 0;
        const bool zhi  = !pz && k + 1 >= d.nz;
        const uint64_t base = static_cast<uint64_t>(k) * plane;
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) {
            const int info = f_info[i];
            if (!(info & 32)) continue;
            const int fy       = (tid >> 6) + 4 * i;
            const int c        = (fy + 1) * kHX + (fx + 1);
            const float centre = so[c];
            float sum          = (info & 1) ? apply_ghost(bc.face[0], centre) : so[c - 1];
            sum                = sum + ((info & 2) ? apply_ghost(bc.face[1], centre) : so[c + 1]);
            sum                = sum + ((info & 4) ? apply_ghost(bc.face[2], centre) : so[c - kHX]);
            sum                = sum + ((info & 8) ? apply_ghost(bc.face[3], centre) : so[c + kHX]);
            sum                = sum + (zlo ? apply_ghost(bc.face[4], centre) : sm[c]);
            sum                = sum + (zhi ? apply_ghost(bc.face[5], centre) : sp[c]);
            x_new[base + f_off[i]] = relax_value<kSolve>(bv[i], sum, coef) * out_scale;
        }
        __syncthreads();
    }
}

TileGeom plan_tiles(const Dims d) {
    TileGeom g{};
    g.tiles_x   = (d.nx + kTX - 1) / kTX;
    g.tiles_y   = (d.ny + kTY - 1) / kTY;
    const long tiles = static_cast<long>(g.tiles_x) * g.tiles_y;
    // about 1024 workgroups (4 per CU), at least 4 planes per chunk
    long zc = (static_cast<long>(d.nz) * tiles + 1023) / 1024;
    if (zc < 4) zc = 4;
    if (zc > 64) zc = 64;
    if (zc > d.nz) zc = d.nz;
    g.zc     = static_cast<int>(zc);
    g.chunks = (d.nz + g.zc - 1) / g.zc;
    return g;
}

} // namespace

hipError_t launch_smooth(hipStream_t s, const SmootherKind smoother, const SolveKind solve, const SmoothPass& pass, const Dims d, const StencilBC& bc) {
    const TileGeom g      = plan_tiles(d);
    const unsigned blocks = static_cast<unsigned>(g.tiles_x * g.tiles_y * g.chunks);
    const bool zero       = pass.x_old == nullptr;
#define SF_PASS(KERNEL, SO, Z) KERNEL<SO, Z><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g)
    if (smoother == kSmoothRBGS) {
        if (solve == kSolvePoisson) {
            if (zero) SF_PASS(k_rbgs_pass, kSolvePoisson, true);
            else SF_PASS(k_rbgs_pass, kSolvePoisson, false);
        } else {
            if (zero) SF_PASS(k_rbgs_pass, kSolveDiffusion, true);
            else SF_PASS(k_rbgs_pass, kSolveDiffusion, false);
        }
    } else {
        if (solve == kSolvePoisson) {
            if (zero) SF_PASS(k_jacobi_pass, kSolvePoisson, true);
            else SF_PASS(k_jacobi_pass, kSolvePoisson, false);
        } else {
            if (zero) SF_PASS(k_jacobi_pass, kSolveDiffusion, true);
            else SF_PASS(k_jacobi_pass, kSolveDiffusion, false);
        }
    }
#undef SF_PASS
    return hipGetLastError();
}

} // namespace sf
