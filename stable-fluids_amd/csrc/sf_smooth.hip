// sf_smooth.hip -- z-marching LDS-tiled smoother passes for gfx950.
//
// One pass = one red-black Gauss-Seidel iteration (or one Jacobi sweep) of
//   diffusion: x = (b + c * sum6) / (1 + 6c)      (reference backend-cuda.cu:354-375)
//   Poisson:   x = (sum6 - c * b) / 6             (reference backend-cuda.cu:398-407)
// read from x_old (or zeros) and b, written to x_new: 12 B/cell of compulsory traffic.
//
// Work decomposition. A workgroup owns a 128 x 16 column of cells in x-y and a chunk of
// zc planes and marches upward in z. Each lane works on x-pairs (cells 2i, 2i+1 of a
// row): in a checkerboard a pair holds one red and one black cell, so with 64-bit LDS
// accesses every lane relaxes exactly one cell per stage instead of masking half a
// wavefront.
//  RBGS: plane k of the half-updated iterate H (red relaxed from the old iterate, black
//  kept) is computed on the tile plus a one-pair halo from an LDS ring of three old
//  planes with a two-pair halo; then the black cells of plane k-1 are relaxed from the
//  LDS ring of H planes. Halo red values are recomputed, so tiles exchange nothing.
//  Jacobi: every cell of plane k from the old ring (one-pair halo).
// The next old plane is fetched into registers before the current plane is computed.
//
// Boundaries. A neighbour outside the grid on a non-periodic axis is always adjacent to
// the cell being relaxed, and the reference's ghost value for it is a function of that
// cell's current value (copy, 2*v_face - value, or a constant; backend-cuda.cu:62-189).
// The halo position is therefore filled, when its plane is loaded, with the rule applied
// to the clamped (adjacent) value; the stencil itself reads LDS unconditionally. Out-of-
// grid planes below/above a non-periodic z range hold the z-face rule of the edge plane.
// Periodic axes load the wrapped cell. Red/black parity uses wrapped coordinates, and an
// odd periodic axis, where a pair can hold two cells of one colour, takes a second relax.
#include <atomic>
#include <algorithm>
#include <type_traits>

#include "sf_buffer.hpp"
#include "sf_kernels.hpp"
#include "sf_relax.hpp"
#include "sf_storage.hpp"

namespace sf {

namespace {

constexpr int kThreads = 512;
constexpr int kPX      = 64;       // output pairs per row (128 cells)
constexpr int kTY      = 16;       // output rows
constexpr int kTX      = 2 * kPX;  // output cells per row
constexpr int kOPX     = kPX + 4;  // RBGS old region: pairs x0-4 .. x0+131, rows y0-2 .. y0+17
constexpr int kOY      = kTY + 4;
constexpr int kHPX     = kPX + 2;  // H region (and Jacobi old region): pairs x0-2 .. x0+129, rows y0-1 .. y0+16
constexpr int kHY      = kTY + 2;
constexpr int kOldF    = 2 * kOPX * kOY; // floats
constexpr int kHP      = kHPX * kHY;     // pairs
constexpr int kJOldF   = 2 * kHPX * kHY; // floats
constexpr int kOldPer  = (kOldF + kThreads - 1) / kThreads;
constexpr int kHPer    = (kHP + kThreads - 1) / kThreads;
constexpr int kJOldPer = (kJOldF + kThreads - 1) / kThreads;
constexpr int kOutPer  = kPX * kTY / kThreads;
static_assert(kPX * kTY % kThreads == 0, "whole output rows per lane group");

// Relax element j of the pair C, whose row neighbours are the pairs L and R, from the
// six neighbours in the reference's order x-, x+, y-, y+, z-, z+.
template <int kSolve>
__device__ __forceinline__ float relax_elem(const int j, const float2 L, const float2 C, const float2 R, const float2 YM, const float2 YP, const float2 ZM, const float2 ZP, const float bv, const float coef) {
    const float xm = j ? C.x : L.y;
    const float xp = j ? R.x : C.y;
    float sum      = xm;
    sum            = sum + xp;
    sum            = sum + (j ? YM.y : YM.x);
    sum            = sum + (j ? YP.y : YP.x);
    sum            = sum + (j ? ZM.y : ZM.x);
    sum            = sum + (j ? ZP.y : ZP.x);
    return relax_fast<kSolve>(bv, sum, coef);
}

__device__ __forceinline__ int mod3(const int k) { return ((k % 3) + 3) % 3; }


// Reciprocal of the power-of-two scale of a narrow-storage correction solve (the slot
// holds max|r| of its right-hand side); 1 for FP32 storage.
template <class St>
__device__ __forceinline__ float rhs_scale(const float* __restrict__ slot) {
    if constexpr (St::kNarrow) return slot ? 1.0f / correction_scale<St>(slot) : 1.0f;
    else return 1.0f;
}

// Stores the aligned pair (a, b) at dst[0], dst[1] with one vector store.
template <class St>
__device__ __forceinline__ void store_vec2(typename St::T* dst, const float a, const float b) {
    if constexpr (St::kNarrow) {
        struct alignas(2 * sizeof(typename St::T)) P {
            typename St::T x, y;
        };
        *reinterpret_cast<P*>(dst) = P{St::store(a), St::store(b)};
    } else {
        *reinterpret_cast<float2*>(dst) = make_float2(a, b);
    }
}

// Ghost code of a halo position: 0 inside (or periodic), else 1 + face index.
__device__ __forceinline__ int ghost_code(const int gx, const int gy, const Dims d, const bool px, const bool py) {
    if (!px && gx < 0) return 1;
    if (!px && gx >= d.nx) return 2;
    if (!py && gy < 0) return 3;
    if (!py && gy >= d.ny) return 4;
    return 0;
}

// Loads plane k of a row-major region (`per` floats per lane) into registers, applying
// the ghost rules: in-grid planes take the x/y face rule of ghost positions, out-of-grid
// planes (non-periodic z) take the z face rule of the adjacent edge plane.
template <int kPer, bool kZero, class St = StoreF32>
__device__ __forceinline__ void fetch_region(float (&r)[kPer], const typename St::T* __restrict__ src, const int (&off)[kPer], const int (&gc)[kPer], const int k, const Dims d, const bool pz, const uint64_t plane, const StencilBC& bc) {
    const bool below = !pz && k < 0, above = !pz && k >= d.nz;
    const uint64_t base = static_cast<uint64_t>(fix_coord(k, d.nz, pz)) * plane;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        float v = 0.0f;
        if (!kZero && off[i] >= 0) v = St::load(src[base + static_cast<uint64_t>(off[i])]);
        if (below)
            v = apply_ghost(bc.face[4], v);
        else if (above)
            v = apply_ghost(bc.face[5], v);
        else if (gc[i])
            v = apply_ghost(face_rule(bc, gc[i] - 1), v);
        r[i] = v;
    }
}

// Pair version of fetch_region for even row lengths: slot i loads the aligned pair at
// pair_off[i] (-1: nothing) and keeps element pair_sel[i] (0 both, 1 both .x, 2 both .y:
// a left or right x-ghost pair reads the adjacent edge cell), then applies the ghost
// rule pair_gc[i] like fetch_region does per element.
template <int kPer, bool kZero, class St = StoreF32>
__device__ __forceinline__ void fetch_pairs(float2 (&r)[kPer], const typename St::T* __restrict__ src, const int (&pair_off)[kPer], const int (&pair_sel)[kPer], const int (&pair_gc)[kPer], const int k, const Dims d, const bool pz, const uint64_t plane, const StencilBC& bc) {
    const bool below = !pz && k < 0, above = !pz && k >= d.nz;
    const uint64_t base = static_cast<uint64_t>(fix_coord(k, d.nz, pz)) * plane;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        float2 v = make_float2(0.0f, 0.0f);
        if (!kZero && pair_off[i] >= 0) v = St::load2(src + base + static_cast<uint64_t>(pair_off[i]));
        if (pair_sel[i] == 1) v.y = v.x;
        if (pair_sel[i] == 2) v.x = v.y;
        if (below || above) {
            const FaceRule f = below ? bc.face[4] : bc.face[5];
            v                = make_float2(apply_ghost(f, v.x), apply_ghost(f, v.y));
        } else if (pair_gc[i]) {
            const FaceRule f = face_rule(bc, pair_gc[i] - 1);
            v                = make_float2(apply_ghost(f, v.x), apply_ghost(f, v.y));
        }
        r[i] = v;
    }
}

template <int kSolve, bool kZero>
__global__ __launch_bounds__(kThreads) void k_rbgs_pass(const float* __restrict__ x_old, const float* __restrict__ b, float* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC bc, const TileGeom g, const int* __restrict__ skip) {
    __shared__ float2 s_old[3][kOPX * kOY];
    __shared__ float2 s_h[3][kHPX * kHY];
    if (skip != nullptr && *skip != 0) return;
    const int w = xcd_block();

    int tx, ty, cz;
    tile_of(g, w, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int x0   = tx * kTX, y0 = ty * kTY;
    const int z0   = cz * g.zc;
    const int z1   = min(z0 + g.zc, d.nz);
    const bool px  = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint64_t plane = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny);

    // old region load slots (floats, row-major 2*kOPX wide)
    int o_off[kOldPer], o_gc[kOldPer];
#pragma unroll
    for (int i = 0; i < kOldPer; ++i) {
        const int e = tid + i * kThreads;
        o_off[i]    = -1;
        o_gc[i]     = 0;
        if (e < kOldF) {
            const int gx = x0 - 4 + e % (2 * kOPX), gy = y0 - 2 + e / (2 * kOPX);
            o_off[i]     = fix_coord(gy, d.ny, py) * d.nx + fix_coord(gx, d.nx, px);
            o_gc[i]      = ghost_code(gx, gy, d, px, py);
        }
    }
    // H pair slots: b offsets of both elements, parity and reality bits
    int h_off[kHPer], h_offb[kHPer], h_info[kHPer]; // info: bit0 real a, bit1 real b, bit2 parity a, bit3 parity b, bit4 slot used
#pragma unroll
    for (int i = 0; i < kHPer; ++i) {
        const int e = tid + i * kThreads;
        h_off[i]    = 0;
        h_offb[i]   = 0;
        h_info[i]   = 0;
        if (e < kHP) {
            const int xa = x0 - 2 + 2 * (e % kHPX), y = y0 - 1 + e / kHPX;
            const bool ry  = py || (y >= 0 && y < d.ny);
            const bool ra  = ry && (px || (xa >= 0 && xa < d.nx));
            const bool rb  = ry && (px || (xa + 1 >= 0 && xa + 1 < d.nx));
            const int wy   = fix_coord(y, d.ny, py);
            const int wa   = fix_coord(xa, d.nx, px), wb = fix_coord(xa + 1, d.nx, px);
            h_off[i]       = wy * d.nx + wa;
            h_offb[i]      = wy * d.nx + wb;
            h_info[i]      = (ra ? 1 : 0) | (rb ? 2 : 0) | (((wa + wy) & 1) << 2) | (((wb + wy) & 1) << 3) | 16;
        }
    }
    // output pair slots: one row of 64 pairs per wavefront
    const int fp = tid & (kPX - 1);
    int f_info[kOutPer]; // bit0 in-grid a, bit1 in-grid b, bit2 parity(xa+y)
#pragma unroll
    for (int i = 0; i < kOutPer; ++i) {
        const int fy = (tid >> 6) + (kThreads / 64) * i;
        const int xa = x0 + 2 * fp, y = y0 + fy;
        f_info[i]    = (xa < d.nx && y < d.ny ? 1 : 0) | (xa + 1 < d.nx && y < d.ny ? 2 : 0) | (((xa + y) & 1) << 2);
    }

    auto real_plane = [&](const int k) { return pz || (k >= 0 && k < d.nz); };

    // b of the red element of each H slot at plane k, and of the black element of each
    // output slot at plane k (loaded one iteration ahead)
    float r_old[kOldPer], r_bh[kHPer], r_bf[kOutPer];
    auto fetch_bh = [&](const int k) {
        if (!real_plane(k)) return;
        const int wz        = fix_coord(k, d.nz, pz);
        const uint64_t base = static_cast<uint64_t>(wz) * plane;
#pragma unroll
        for (int i = 0; i < kHPer; ++i) {
            const int info   = h_info[i];
            const bool red_a = (((info >> 2) + wz) & 1) == 0;
            r_bh[i]          = (info & 16) ? b[base + static_cast<uint64_t>(red_a ? h_off[i] : h_offb[i])] : 0.0f;
        }
    };
    auto fetch_bf = [&](const int k) {
        if (k < 0 || k >= d.nz) return;
        const uint64_t base = static_cast<uint64_t>(k) * plane;
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) {
            const int fy   = (tid >> 6) + (kThreads / 64) * i;
            const int info = f_info[i];
            const int j    = (((info >> 2) + k) & 1) ? 0 : 1; // black element
            if (info & (1 << j)) r_bf[i] = b[base + static_cast<uint64_t>((y0 + fy) * d.nx + x0 + 2 * fp + j)];
        }
    };
    auto stash_old = [&](const int k) {
        float* s = reinterpret_cast<float*>(s_old[mod3(k)]);
#pragma unroll
        for (int i = 0; i < kOldPer; ++i)
            if (tid + i * kThreads < kOldF) s[tid + i * kThreads] = r_old[i];
    };

    // prologue: old planes z0-2, z0-1 in LDS; old z0, b for H(z0-1) in registers
    fetch_region<kOldPer, kZero>(r_old, x_old, o_off, o_gc, z0 - 2, d, pz, plane, bc);
    stash_old(z0 - 2);
    fetch_region<kOldPer, kZero>(r_old, x_old, o_off, o_gc, z0 - 1, d, pz, plane, bc);
    stash_old(z0 - 1);
    fetch_region<kOldPer, kZero>(r_old, x_old, o_off, o_gc, z0, d, pz, plane, bc);
    fetch_bh(z0 - 1);

    for (int k = z0 - 1; k <= z1; ++k) {
        stash_old(k + 1);
        float bh[kHPer], bf[kOutPer];
#pragma unroll
        for (int i = 0; i < kHPer; ++i) bh[i] = r_bh[i];
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) bf[i] = r_bf[i];
        if (k + 1 <= z1) {
            fetch_region<kOldPer, kZero>(r_old, x_old, o_off, o_gc, k + 2, d, pz, plane, bc);
            fetch_bh(k + 1);
        }
        if (k >= z0 - 1 && k < z1) fetch_bf(k); // black b of plane k, used by final(k) next iteration
        __syncthreads();

        // H(k) on the tile plus a one-pair halo
        {
            const float2* so = s_old[mod3(k)];
            const float2* sm = s_old[mod3(k - 1)];
            const float2* sp = s_old[mod3(k + 1)];
            float2* sh       = s_h[mod3(k)];
            const bool real  = real_plane(k);
            const int wz     = fix_coord(k, d.nz, pz);
#pragma unroll
            for (int i = 0; i < kHPer; ++i) {
                const int e = tid + i * kThreads;
                if (e >= kHP) continue;
                const int c    = (e / kHPX + 1) * kOPX + (e % kHPX + 1);
                const float2 C = so[c];
                float2 out     = C;
                const int info = h_info[i];
                if (real) {
                    const bool red_a = (info & 1) && ((((info >> 2) + wz) & 1) == 0);
                    const bool red_b = (info & 2) && ((((info >> 3) + wz) & 1) == 0);
                    if (red_a || red_b) {
                        const float2 L = so[c - 1], R = so[c + 1], YM = so[c - kOPX], YP = so[c + kOPX];
                        const float2 ZM = sm[c], ZP = sp[c];
                        const int j     = red_a ? 0 : 1;
                        const float v   = relax_elem<kSolve>(j, L, C, R, YM, YP, ZM, ZP, bh[i], coef);
                        if (j == 0) out.x = v;
                        else out.y = v;
                        if (red_a && red_b) { // odd periodic wrap: both cells red
                            const float bb = b[static_cast<uint64_t>(wz) * plane + static_cast<uint64_t>(h_offb[i])];
                            out.y          = relax_elem<kSolve>(1, L, C, R, YM, YP, ZM, ZP, bb, coef);
                        }
                    }
                }
                sh[e] = out;
            }
        }
        __syncthreads();

        // final plane k-1 on the tile: black cells from H
        const int kf = k - 1;
        if (kf >= z0 && kf < z1) {
            const float2* hm = s_h[mod3(kf - 1)];
            const float2* h0 = s_h[mod3(kf)];
            const float2* hp = s_h[mod3(kf + 1)];
            const uint64_t base = static_cast<uint64_t>(kf) * plane;
#pragma unroll
            for (int i = 0; i < kOutPer; ++i) {
                const int info = f_info[i];
                if (!(info & 1)) continue;
                const int fy   = (tid >> 6) + (kThreads / 64) * i;
                const int c    = (fy + 1) * kHPX + (fp + 1);
                const float2 C = h0[c];
                float2 out     = C;
                const int j    = (((info >> 2) + kf) & 1) ? 0 : 1;
                if (info & (1 << j)) {
                    const float v = relax_elem<kSolve>(j, h0[c - 1], C, h0[c + 1], h0[c - kHPX], h0[c + kHPX], hm[c], hp[c], bf[i], coef);
                    if (j == 0) out.x = v;
                    else out.y = v;
                }
                float* dst = x_new + base + static_cast<uint64_t>((y0 + fy) * d.nx + x0 + 2 * fp);
                if ((d.nx & 1) == 0) {
                    *reinterpret_cast<float2*>(dst) = make_float2(out.x * out_scale, out.y * out_scale);
                } else {
                    dst[0] = out.x * out_scale;
                    if (info & 2) dst[1] = out.y * out_scale;
                }
            }
        }
        __syncthreads();
    }
}

// Jacobi pass for even nx. Each wave fetches whole rows of the old region (lane l holds
// the pair x0 + 2l, x0 + 2l + 1), so ghost rules are wave-uniform per row; lanes 0..3
// of every wave fetch the two halo columns (x0 - 1 and x0 + 128) of two rows. Only the
// cells the 7-point stencil reads are filled: rows y0 - 1 .. y0 + 16 at the 64 interior
// pairs, halo columns on rows y0 .. y0 + 15; corner cells stay unset.
//
// Planes are fetched two ahead into alternating register sets and staged in a ring of
// four LDS slots, so one barrier per plane orders the ring (a fast wave writes slot
// k + 2 while a slow one still reads k - 1 .. k + 1). All loads and stores go through
// plane buffer resources with out-of-range offsets for idle lanes and 0-byte resources
// for idle planes: no memory instruction sits under a branch, the compiler's wait counts
// stay exact and a plane's loads are only waited for when it is staged two planes later.
// Ghost rules and the right-hand-side scale are applied at staging / relaxing time.
template <int kSolve, bool kZero, class St>
__device__ __forceinline__ void jacobi_rows_body(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC& bc, const TileGeom& g, const float* __restrict__ scale_slot, const int w) {
    using T                = typename St::T;
    constexpr int kWaves   = kThreads / 64;
    constexpr int kRowsPer = (kHY + kWaves - 1) / kWaves; // 3
    __shared__ float2 s_old[4][kHY * kHPX];
    const float inv_s = rhs_scale<St>(scale_slot);

    int tx, ty, cz;
    tile_of(g, w, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = tx * kTX, y0 = ty * kTY;
    const int z0 = cz * g.zc;
    const int z1 = min(z0 + g.zc, d.nz);
    const bool px = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint32_t plane = static_cast<uint32_t>(d.nx) * static_cast<uint32_t>(d.ny);
    const bool full_x    = x0 + kTX <= d.nx;

    // this lane's pair of a row; in a tile that crosses x = nx, the pair at nx carries the
    // x+ neighbour of the last cell in .x (ghost or wrapped cell 0), later pairs are unused
    const int xa    = x0 + 2 * lane;
    int lane_src    = xa;
    bool lane_ghost = false;
    if (!full_x && xa >= d.nx) {
        lane_src   = (xa == d.nx && px) ? 0 : d.nx - 2;
        lane_ghost = xa == d.nx && !px;
    }
    int row_off[kRowsPer]; // byte offsets in a plane
    bool row_ghost[kRowsPer];
    FaceRule row_rule[kRowsPer];
#pragma unroll
    for (int j = 0; j < kRowsPer; ++j) {
        const int r  = wave + kWaves * j;
        const int gy = y0 - 1 + r;
        row_ghost[j] = !py && (gy < 0 || gy >= d.ny);
        row_rule[j]  = gy < 0 ? bc.face[2] : bc.face[3];
        // a wave without a third row re-reads its last one (never staged)
        const int gr = r < kHY ? gy : gy - kWaves;
        row_off[j]   = static_cast<int>(sizeof(T)) * (fix_coord(gr, d.ny, py) * d.nx + lane_src);
    }
    // halo columns: lane bit 0 = side (0: x0 - 1, 1: x0 + 128), bit 1 = row of the pair
    // of tile rows 2 wave, 2 wave + 1; columns beyond a partial tile are unused. Lanes 4 ..
    // 63 repeat lanes 0 .. 3 (same addresses, never staged).
    const int h_side = lane & 1;
    const int h_row  = 2 * wave + ((lane >> 1) & 1);
    int h_off        = kOutOfRange;
    bool h_ghost     = false;
    {
        const int gy = min(y0 + h_row, d.ny - 1);
        int gx       = h_side ? x0 + kTX : x0 - 1;
        if (gx < 0) {
            h_ghost = !px;
            gx      = px ? d.nx - 1 : 0;
        } else if (gx >= d.nx) {
            h_ghost = gx == d.nx && !px;
            gx      = (gx == d.nx && px) ? 0 : d.nx - 1;
        }
        h_off = static_cast<int>(sizeof(T)) * (gy * d.nx + gx);
    }
    const FaceRule h_rule = h_side ? bc.face[1] : bc.face[0];
    const int h_lds       = 2 * ((1 + h_row) * kHPX + (h_side ? kHPX - 1 : 0)) + (h_side ? 0 : 1);

    int b_off[kOutPer], x_off[kOutPer];
#pragma unroll
    for (int i = 0; i < kOutPer; ++i) {
        const int y    = y0 + wave + kWaves * i;
        const bool ok  = xa < d.nx && y < d.ny;
        b_off[i]       = ok ? static_cast<int>(sizeof(float)) * (y * d.nx + xa) : kOutOfRange;
        x_off[i]       = ok ? static_cast<int>(sizeof(T)) * (y * d.nx + xa) : kOutOfRange;
    }
    const uint32_t x_bytes = plane * static_cast<uint32_t>(sizeof(T));
    const uint32_t b_bytes = plane * static_cast<uint32_t>(sizeof(float));

    struct Regs {
        float2 row[kRowsPer];
        float h;
    };
    Regs ra, rb;
    float2 ba[kOutPer], bb[kOutPer];
    // raw values of plane k (nothing when k is past the last plane the chunk reads)
    auto fetch_old = [&](const int k, Regs& r) {
        if constexpr (kZero) return;
        const Rsrc src = plane_rsrc(x_old + static_cast<uint64_t>(fix_coord(k, d.nz, pz)) * plane, k <= z1 ? x_bytes : 0u);
#pragma unroll
        for (int j = 0; j < kRowsPer; ++j) r.row[j] = buf_load2<St>(src, row_off[j]);
        r.h = buf_load1<St>(src, h_off);
    };
    auto stash_old = [&](const int k, const Regs& r) {
        const bool below = !pz && k < 0, above = !pz && k >= d.nz;
        const FaceRule zr = below ? bc.face[4] : bc.face[5];
        float2* s         = s_old[k & 3];
#pragma unroll
        for (int j = 0; j < kRowsPer; ++j) {
            if (wave + kWaves * j >= kHY) continue;
            float2 v = kZero ? make_float2(0.0f, 0.0f) : r.row[j];
            if (below || above) {
                v = make_float2(apply_ghost(zr, v.x), apply_ghost(zr, v.y));
            } else if (row_ghost[j]) {
                v = make_float2(apply_ghost(row_rule[j], v.x), apply_ghost(row_rule[j], v.y));
            } else if (!full_x && lane_ghost) {
                v.x = apply_ghost(bc.face[1], v.y);
            }
            s[(wave + kWaves * j) * kHPX + 1 + lane] = v;
        }
        if (lane < 4) {
            float v = kZero ? 0.0f : r.h;
            if (below || above) v = apply_ghost(zr, v);
            else if (h_ghost) v = apply_ghost(h_rule, v);
            reinterpret_cast<float*>(s)[h_lds] = v;
        }
    };
    // raw b of plane k (nothing past the chunk)
    auto fetch_b = [&](const int k, float2 (&r_b)[kOutPer]) {
        const Rsrc src = plane_rsrc(b + static_cast<uint64_t>(k) * plane, k < z1 ? b_bytes : 0u);
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) r_b[i] = buf_load2<StoreF32>(src, b_off[i]);
    };
    // relax plane k (stores dropped past the chunk)
    auto compute = [&](const int k, const float2 (&bv)[kOutPer]) {
        const float2* so = s_old[k & 3];
        const float2* sm = s_old[(k - 1) & 3];
        const float2* sp = s_old[(k + 1) & 3];
        const Rsrc dst   = plane_rsrc(x_new + static_cast<uint64_t>(k) * plane, k < z1 ? x_bytes : 0u);
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) {
            const int c    = (wave + kWaves * i + 1) * kHPX + lane + 1;
            const float2 L = so[c - 1], C = so[c], R = so[c + 1], YM = so[c - kHPX], YP = so[c + kHPX], ZM = sm[c], ZP = sp[c];
            const float v0 = relax_elem<kSolve>(0, L, C, R, YM, YP, ZM, ZP, bv[i].x * inv_s, coef) * out_scale;
            const float v1 = relax_elem<kSolve>(1, L, C, R, YM, YP, ZM, ZP, bv[i].y * inv_s, coef) * out_scale;
            buf_store2<St, kDeviceScopeStore>(dst, x_off[i], v0, v1);
        }
    };
    // plane k: stage k + 1 (fetched two steps ago into r) and refill r with k + 3; after the
    // barrier relax k with b of k (fetched two steps ago into bk) and refill bk with k + 2
    auto plane_step = [&](const int k, Regs& r, float2 (&bk)[kOutPer]) {
        stash_old(k + 1, r);
        fetch_old(k + 3, r);
        __syncthreads();
        compute(k, bk);
        fetch_b(k + 2, bk);
    };

    fetch_old(z0 - 1, ra);
    stash_old(z0 - 1, ra);
    fetch_old(z0, ra);
    stash_old(z0, ra);
    fetch_old(z0 + 1, rb); // plane z0 + 1 + 2j in rb, z0 + 2j in ra (j >= 1)
    fetch_old(z0 + 2, ra);
    fetch_b(z0, ba);       // b of z0 + 2j in ba, of z0 + 1 + 2j in bb
    fetch_b(z0 + 1, bb);
    // an odd chunk ends with one idle step (its loads and stores are switched off)
    for (int k = z0; k < z1; k += 2) {
        plane_step(k, rb, ba);
        plane_step(k + 1, ra, bb);
    }
}

// Jacobi pass for odd nx: the old region is fetched element by element (kJOldPer per
// lane) with per-element ghost codes; same ring and prefetch as jacobi_rows_body.
template <int kSolve, bool kZero, class St>
__device__ __forceinline__ void jacobi_body(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC& bc, const TileGeom& g, const float* __restrict__ scale_slot, const int w) {
    __shared__ float2 s_old[4][kHPX * kHY];
    const float inv_s = rhs_scale<St>(scale_slot);

    int tx, ty, cz;
    tile_of(g, w, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int x0   = tx * kTX, y0 = ty * kTY;
    const int z0   = cz * g.zc;
    const int z1   = min(z0 + g.zc, d.nz);
    const bool px  = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint64_t plane = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny);

    int o_off[kJOldPer], o_gc[kJOldPer];
#pragma unroll
    for (int i = 0; i < kJOldPer; ++i) {
        const int e = tid + i * kThreads;
        o_off[i]    = -1;
        o_gc[i]     = 0;
        if (e < kJOldF) {
            const int gx = x0 - 2 + e % (2 * kHPX), gy = y0 - 1 + e / (2 * kHPX);
            o_off[i]     = fix_coord(gy, d.ny, py) * d.nx + fix_coord(gx, d.nx, px);
            o_gc[i]      = ghost_code(gx, gy, d, px, py);
        }
    }
    const int fp = tid & (kPX - 1);
    int f_info[kOutPer];
#pragma unroll
    for (int i = 0; i < kOutPer; ++i) {
        const int fy = (tid >> 6) + (kThreads / 64) * i;
        const int xa = x0 + 2 * fp, y = y0 + fy;
        f_info[i]    = (xa < d.nx && y < d.ny ? 1 : 0) | (xa + 1 < d.nx && y < d.ny ? 2 : 0);
    }
    struct OldRegs {
        float f[kJOldPer];
    };
    OldRegs ra, rb;
    float2 ba[kOutPer], bb[kOutPer];
    auto fetch_b = [&](const int k, float2 (&r_b)[kOutPer]) {
        const uint64_t base = static_cast<uint64_t>(k) * plane;
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) {
            const int fy  = (tid >> 6) + (kThreads / 64) * i;
            const float* p = b + base + static_cast<uint64_t>((y0 + fy) * d.nx + x0 + 2 * fp);
            if ((d.nx & 1) == 0 && (f_info[i] & 2)) {
                const float2 q = *reinterpret_cast<const float2*>(p);
                r_b[i]         = make_float2(q.x * inv_s, q.y * inv_s);
            } else {
                if (f_info[i] & 1) r_b[i].x = p[0] * inv_s;
                if (f_info[i] & 2) r_b[i].y = p[1] * inv_s;
            }
        }
    };
    auto stash_old = [&](const int k, const OldRegs& r) {
        float* s = reinterpret_cast<float*>(s_old[k & 3]);
#pragma unroll
        for (int i = 0; i < kJOldPer; ++i)
            if (tid + i * kThreads < kJOldF) s[tid + i * kThreads] = r.f[i];
    };
    auto fetch_old = [&](const int k, OldRegs& r) { fetch_region<kJOldPer, kZero, St>(r.f, x_old, o_off, o_gc, k, d, pz, plane, bc); };
    auto compute = [&](const int k, const float2 (&bv)[kOutPer]) {
        const float2* so    = s_old[k & 3];
        const float2* sm    = s_old[(k - 1) & 3];
        const float2* sp    = s_old[(k + 1) & 3];
        const uint64_t base = static_cast<uint64_t>(k) * plane;
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) {
            const int info = f_info[i];
            if (!(info & 1)) continue;
            const int fy   = (tid >> 6) + (kThreads / 64) * i;
            const int c    = (fy + 1) * kHPX + (fp + 1);
            const float2 L = so[c - 1], C = so[c], R = so[c + 1], YM = so[c - kHPX], YP = so[c + kHPX], ZM = sm[c], ZP = sp[c];
            auto* dst      = x_new + base + static_cast<uint64_t>((y0 + fy) * d.nx + x0 + 2 * fp);
            const float v0 = relax_elem<kSolve>(0, L, C, R, YM, YP, ZM, ZP, bv[i].x, coef) * out_scale;
            if ((d.nx & 1) == 0) {
                const float v1 = relax_elem<kSolve>(1, L, C, R, YM, YP, ZM, ZP, bv[i].y, coef) * out_scale;
                store_vec2<St>(dst, v0, v1);
            } else {
                dst[0] = St::store(v0);
                if (info & 2) dst[1] = St::store(relax_elem<kSolve>(1, L, C, R, YM, YP, ZM, ZP, bv[i].y, coef) * out_scale);
            }
        }
    };
    // plane k: stage k+1 (fetched two planes ago into `r`), refill `r` with k+3 and `bn`
    // with b of k+2, then relax plane k with b of k (held in `bn` until now)
    auto plane_step = [&](const int k, OldRegs& r, float2 (&bn)[kOutPer]) {
        stash_old(k + 1, r);
        float2 bv[kOutPer];
#pragma unroll
        for (int i = 0; i < kOutPer; ++i) bv[i] = bn[i];
        if (k + 3 <= z1) fetch_old(k + 3, r);
        if (k + 2 < z1) fetch_b(k + 2, bn);
        __syncthreads();
        compute(k, bv);
    };

    fetch_old(z0 - 1, ra);
    stash_old(z0 - 1, ra);
    fetch_old(z0, ra);
    stash_old(z0, ra);
    fetch_old(z0 + 1, rb); // plane z0 + 1 + 2j in rb, z0 + 2j in ra (j >= 1)
    if (z0 + 2 <= z1) fetch_old(z0 + 2, ra);
    fetch_b(z0, ba);       // b of z0 + 2j in ba, of z0 + 1 + 2j in bb
    if (z0 + 1 < z1) fetch_b(z0 + 1, bb);
    for (int k = z0; k < z1; k += 2) {
        plane_step(k, rb, ba);
        if (k + 1 < z1) plane_step(k + 1, ra, bb);
    }
}

// ---------------------------------------------------------------------------------------
// RBGS fast path (no periodic axis of odd length). LDS planes are split into even-x and
// odd-x arrays, so the colour of an array alternates by row: the black stage of a row is
// wave-uniform (one row per wavefront) and every stencil read is a contiguous 32-bit LDS
// access. H planes hold only the red value of each pair. With even periodic lengths the
// unwrapped parity equals the wrapped one.
// ---------------------------------------------------------------------------------------
constexpr int kW        = kPX + 4;     // pairs per LDS row, x from xs = x0 - 4
constexpr int kR        = kTY + 4;     // rows, y from ys = y0 - 2
constexpr int kArr      = kR * kW;     // floats per even/odd array of a plane
constexpr int kOldPlane = 2 * kArr;
constexpr int kHPlane   = kArr;
constexpr int kHVals    = (kTY + 2) * (kPX + 2);                      // red values per H plane
constexpr int kHSlots3  = (kHVals + kThreads - 1) / kThreads;
constexpr int kOldPairs = kArr;                                        // pairs per old plane
constexpr int kOldPairPer = (kOldPairs + kThreads - 1) / kThreads;

// Periodic plane index for k in [-2, n + 1] without a division (n >= 2); n = 1 maps to 0.
__device__ __forceinline__ int wrap_near(const int k, const int n) {
    if (n < 2) return 0;
    return k < 0 ? k + n : (k >= n ? k - n : k);
}

// RBGS split pass for even nx. Each wave owns two tile rows (y0 + wave, y0 + wave + 8):
// lane l holds the pair x0 + 2l (LDS pair i = l + 2) of those rows in the old region, the
// H stage and the final stage. The rest of the old region (halo rows 0, 1, 18, 19 and the
// four halo pairs of every row) and of the H region (halo rows 1, 18 and the halo pairs
// 1, 66) is spread evenly as one extra slot per lane, with per-lane ghost codes, so every
// wave does the same work between barriers. The red cell of a pair is the even or the odd
// one by the parity of y + z (x of a pair start is even), so on the tile rows every LDS
// index is a row-uniform base plus the lane. Ghost rules run as one FMA (LinGhost) and
// only where a tile touches a ghost; the right-hand-side scale is applied at use. Loads
// and stores go through plane buffer resources (out-of-range offsets for idle lanes,
// 0-byte resources for idle planes) and planes are fetched two ahead into alternating
// register sets. Two barriers per plane: after staging old(k + 1) and after H(k); the
// next plane's first barrier already orders the reuse of the H ring, the second one that
// of the old ring.
template <int kSolve, bool kZero, class St, bool kWhole>
__device__ __forceinline__ void rbgs_rows_body(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC& bc, const TileGeom& g, const float* __restrict__ scale_slot, const int w) {
    using T              = typename St::T;
    constexpr int kT     = static_cast<int>(sizeof(T));
    constexpr int kWaves = kThreads / 64;
    constexpr int kRowsW = kTY / kWaves;                                    // tile rows per wave
    constexpr int kOMisc = (4 * kPX + 4 * kR) / kWaves;                     // extra old pairs per wave
    constexpr int kHAll  = 2 * (kPX + 2) + 2 * kTY;                         // extra H pairs
    constexpr int kHMisc = (kHAll + kWaves - 1) / kWaves;                   // per wave
    static_assert(kRowsW == kOutPer && kOMisc * kWaves == 4 * kPX + 4 * kR && kOMisc <= 64 && kHMisc <= 64, "tile split");
    __shared__ float s_old[3][kOldPlane];
    __shared__ float s_h[3][kHPlane];
    const float inv_s = rhs_scale<St>(scale_slot);

    int tx, ty, cz;
    tile_of(g, w, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = tx * kTX, y0 = ty * kTY;
    const int xs = x0 - 4, ys = y0 - 2;
    const int z0 = cz * g.zc;
    const int z1 = min(z0 + g.zc, d.nz);
    const bool px = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint32_t plane   = static_cast<uint32_t>(d.nx) * static_cast<uint32_t>(d.ny);
    const uint32_t x_bytes = plane * static_cast<uint32_t>(sizeof(T));
    const uint32_t b_bytes = plane * static_cast<uint32_t>(sizeof(float));
    const bool full_x      = x0 + kTX <= d.nx;
    const bool full_y      = y0 + kTY <= d.ny;
    // kWhole (fields below 2 GiB): one resource per field, the plane chosen by the scalar
    // offset, and a 0-byte resource for idle planes; otherwise one resource per plane
    // (a resource is rebuilt from the field base with the size selected, which keeps it in
    // scalar registers; selecting between two resources would not)
    const uint32_t x_field = kWhole ? x_bytes * static_cast<uint32_t>(d.nz) : 0u;
    const uint32_t b_field = kWhole ? b_bytes * static_cast<uint32_t>(d.nz) : 0u;

    // the lane's pair on the tile rows; past a wall (partial tile) both cells take the x+
    // ghost of cell nx - 1, across a periodic x face the pair wraps
    const int xa    = x0 + 2 * lane;
    int lane_src    = xa;
    bool lane_xg    = false;
    if (!full_x && xa >= d.nx) {
        if (px) {
            lane_src = wrap_index(xa, d.nx);
        } else {
            lane_src = d.nx - 2;
            lane_xg  = true;
        }
    }
    const bool lane_xreal = px || xa < d.nx;
    const LinGhost g_xp   = lin_ghost(bc.face[1]);
    const LinGhost g_yp   = lin_ghost(bc.face[3]);

    int t_off[kRowsW], hb_off[kRowsW], bf_off[kRowsW], st_off[kRowsW];
    bool t_yg[kRowsW], t_rl[kRowsW];
#pragma unroll
    for (int j = 0; j < kRowsW; ++j) {
        const int y   = y0 + wave + kWaves * j;
        const int yy  = fix_coord(y, d.ny, py);
        t_yg[j]       = !py && y >= d.ny;
        t_off[j]      = kT * (yy * d.nx + lane_src);
        t_rl[j]       = (py || y < d.ny) && lane_xreal;
        hb_off[j]     = t_rl[j] ? 4 * (yy * d.nx + fix_coord(xa, d.nx, px)) : kOutOfRange;
        const bool ok = xa < d.nx && y < d.ny;
        bf_off[j]     = ok ? 4 * (y * d.nx + xa) : kOutOfRange;
        st_off[j]     = ok ? kT * (y * d.nx + xa) : kOutOfRange;
    }
    // extra old pairs: halo rows 0, 1, 18, 19 at the 64 tile pairs, then the halo pairs
    // 0, 1, 66, 67 of all 20 rows
    // (lanes past a wave's share repeat an earlier lane's item: same load, same LDS value)
    int m_off = kOutOfRange, m_sel = 0, m_lds = 0;
    LinGhost m_g = lin_identity();
    {
        const int m = kOMisc * wave + lane % kOMisc;
        int r, i;
        if (m < 4 * kPX) {
            const int hr = m / kPX;
            r            = hr < 2 ? hr : hr + kR - 4;
            i            = m % kPX + 2;
        } else {
            const int m2 = m - 4 * kPX, q = m2 & 3;
            r            = m2 >> 2;
            i            = q < 2 ? q : kW - 4 + q;
        }
        {
            const int gx = xs + 2 * i, gy = ys + r;
            const int row = fix_coord(gy, d.ny, py) * d.nx;
            int off;
            if (!px && gx < 0) {
                off   = row;
                m_sel = 1;
            } else if (!px && gx >= d.nx) {
                off   = row + d.nx - 2;
                m_sel = 2;
            } else {
                off = row + fix_coord(gx, d.nx, px);
            }
            m_off        = kT * off;
            const int gc = ghost_code(gx, gy, d, px, py);
            if (gc) m_g = lin_ghost(face_rule(bc, gc - 1));
            m_lds = r * kW + i;
        }
    }
    // extra H pairs: halo rows 1 and 18 (pairs 1 .. 66), then the halo pairs 1 and 66 of
    // rows 2 .. 17
    int n_off = kOutOfRange, n_base = 0, n_ypar = 0;
    bool n_rl = false;
    {
        const int n_cnt = min(kHMisc, kHAll - kHMisc * wave);
        const int m     = kHMisc * wave + lane % n_cnt;
        int r, i;
        if (m < 2 * (kPX + 2)) {
            r = m < kPX + 2 ? 1 : kTY + 2;
            i = 1 + m % (kPX + 2);
        } else {
            const int m2 = m - 2 * (kPX + 2);
            r            = 2 + (m2 >> 1);
            i            = (m2 & 1) ? kPX + 2 : 1;
        }
        {
            const int y = ys + r, gx = xs + 2 * i;
            n_rl        = (py || (y >= 0 && y < d.ny)) && (px || (gx >= 0 && gx < d.nx));
            if (n_rl) n_off = 4 * (fix_coord(y, d.ny, py) * d.nx + fix_coord(gx, d.nx, px));
            n_base = r * kW + i;
            n_ypar = y & 1;
        }
    }

    auto plane_of = [&](const int k) { return pz ? wrap_near(k, d.nz) : (k < 0 ? 0 : (k >= d.nz ? d.nz - 1 : k)); };

    struct OldRegs {
        float2 t[kRowsW];
        float2 m;
    };
    struct HRegs {
        float t[kRowsW];
        float m;
    };
    OldRegs oa, ob;
    HRegs ha, hb;
    float fa[kOutPer], fb[kOutPer];

    auto fetch_old = [&](const int k, OldRegs& r) {
        if constexpr (kZero) return;
        Rsrc src;
        int so = 0;
        if constexpr (kWhole) {
            src = plane_rsrc(x_old, k <= z1 + 1 ? x_field : 0u);
            so  = plane_of(k) * static_cast<int>(x_bytes);
        } else {
            src = plane_rsrc(x_old + static_cast<uint64_t>(plane_of(k)) * plane, k <= z1 + 1 ? x_bytes : 0u);
        }
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) r.t[j] = buf_load2<St>(src, t_off[j], so);
        r.m = buf_load2<St>(src, m_off, so);
    };
    auto stash_old = [&](const int k, float* s, const OldRegs& r) {
        const bool zg      = !pz && (k < 0 || k >= d.nz);
        const LinGhost g_z = lin_ghost(k < 0 ? bc.face[4] : bc.face[5]);
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            float2 v = kZero ? make_float2(0.0f, 0.0f) : r.t[j];
            if (zg || t_yg[j] || !full_x) {
                if (lane_xg) v.x = v.y;
                const LinGhost gr = zg ? g_z : (lane_xg ? g_xp : (t_yg[j] ? g_yp : lin_identity()));
                v                 = make_float2(lin_apply(gr, v.x), lin_apply(gr, v.y));
            }
            const int c = (2 + wave + kWaves * j) * kW + lane + 2;
            s[c]        = v.x;
            s[kArr + c] = v.y;
        }
        {
            float2 v = kZero ? make_float2(0.0f, 0.0f) : r.m;
            if (m_sel == 1) v.y = v.x;
            if (m_sel == 2) v.x = v.y;
            const LinGhost gm = zg ? g_z : m_g;
            v                 = make_float2(lin_apply(gm, v.x), lin_apply(gm, v.y));
            s[m_lds]          = v.x;
            s[kArr + m_lds]   = v.y;
        }
    };
    // b of the red cells of H plane k (H planes z0 - 1 .. z1, real planes only)
    auto fetch_bh = [&](const int k, HRegs& r) {
        const bool live = k >= z0 - 1 && k <= z1 && (pz || (k >= 0 && k < d.nz));
        const int a_t   = (y0 + wave + k) & 1; // 1: the red cell of a tile-row pair is the odd one
        Rsrc src;
        int so = 4 * a_t;
        if constexpr (kWhole) {
            src = plane_rsrc(b, live ? b_field : 0u);
            so += plane_of(k) * static_cast<int>(b_bytes);
        } else {
            src = plane_rsrc(b + static_cast<uint64_t>(plane_of(k)) * plane, live ? b_bytes : 0u);
        }
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) r.t[j] = buf_load1<StoreF32>(src, hb_off[j], so);
        r.m = buf_load1<StoreF32>(src, n_off + 4 * ((n_ypar + k) & 1), so - 4 * a_t);
    };
    // b of the black cells of final plane k
    auto fetch_bf = [&](const int k, float (&r)[kOutPer]) {
        const bool live = k >= z0 && k < z1;
        const int e     = ((y0 + wave + k) & 1) == 0 ? 1 : 0;
        Rsrc src;
        int so = 4 * e;
        if constexpr (kWhole) {
            src = plane_rsrc(b, live ? b_field : 0u);
            so += (live ? k : 0) * static_cast<int>(b_bytes);
        } else {
            src = plane_rsrc(b + static_cast<uint64_t>(k) * plane, live ? b_bytes : 0u);
        }
#pragma unroll
        for (int j = 0; j < kOutPer; ++j) r[j] = buf_load1<StoreF32>(src, bf_off[j], so);
    };
    // H(k): red value of every pair on the tile plus a one-pair halo; ghosts keep the old value
    auto compute_h = [&](const int k, const float* om, const float* o0, const float* op, float* h0, const HRegs& bh) {
        const bool preal    = pz || (k >= 0 && k < d.nz);
        const bool all_real = preal && (full_x || px) && (full_y || py); // every tile-row pair real
        StageDiv<kSolve> dvs(coef);
        auto num = [&](const int base, const int a, const float bv) {
            const int ca = base + a * kArr;
            const int cx = base + (1 - a) * (kArr - 1);
            float sum    = o0[cx];
            sum          = sum + o0[cx + 1];
            sum          = sum + o0[ca - kW];
            sum          = sum + o0[ca + kW];
            sum          = sum + om[ca];
            sum          = sum + op[ca];
            return dvs.num(bv * inv_s, sum, coef);
        };
        const int a_t = (y0 + wave + k) & 1;
        const int a_m = (n_ypar + k) & 1;
        float x[kRowsW + 1], q[kRowsW + 1];
        int base[kRowsW + 1], av[kRowsW + 1];
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            base[j] = (2 + wave + kWaves * j) * kW + lane + 2;
            av[j]   = a_t;
            x[j]    = num(base[j], a_t, bh.t[j]);
        }
        base[kRowsW] = n_base;
        av[kRowsW]   = a_m;
        x[kRowsW]    = num(n_base, a_m, bh.m);
#pragma unroll
        for (int j = 0; j <= kRowsW; ++j) q[j] = dvs.quot(x[j]);
        if (dvs.bad()) {
#pragma unroll
            for (int j = 0; j <= kRowsW; ++j) q[j] = dvs.exact(x[j]);
        }
        if (!all_real) { // tile rows past a wall or off a ghost plane
#pragma unroll
            for (int j = 0; j < kRowsW; ++j)
                if (!(preal && t_rl[j])) q[j] = o0[base[j] + a_t * kArr];
        }
        const float n_old = o0[n_base + a_m * kArr]; // halo pairs are ghosts at walls
        q[kRowsW]         = (preal && n_rl) ? q[kRowsW] : n_old;
#pragma unroll
        for (int j = 0; j <= kRowsW; ++j) h0[base[j]] = q[j];
    };
    // final plane k - 1: black cells relaxed from the red values, red cells copied; stores
    // switched off outside the chunk
    auto compute_final = [&](const int k, const float* hm2, const float* hm1, const float* h0, const float (&bf)[kOutPer]) {
        const int kf     = k - 1;
        const bool live  = kf >= z0 && kf < z1;
        Rsrc dst;
        int so = 0;
        if constexpr (kWhole) {
            dst = plane_rsrc(x_new, live ? x_field : 0u);
            so  = (live ? kf : 0) * static_cast<int>(x_bytes);
        } else {
            dst = plane_rsrc(x_new + static_cast<uint64_t>(live ? kf : 0) * plane, live ? x_bytes : 0u);
        }
        const int e      = ((y0 + wave + kf) & 1) == 0 ? 1 : 0; // 1: the red cell is the even one
        StageDiv<kSolve> dvs(coef);
        float x[kOutPer], q[kOutPer], red[kOutPer];
#pragma unroll
        for (int j = 0; j < kOutPer; ++j) {
            const int c = (2 + wave + kWaves * j) * kW + lane + 2;
            float sum   = hm1[c - 1 + e];
            sum         = sum + hm1[c + e];
            sum         = sum + hm1[c - kW];
            sum         = sum + hm1[c + kW];
            sum         = sum + hm2[c];
            sum         = sum + h0[c];
            x[j]        = dvs.num(bf[j] * inv_s, sum, coef);
            red[j]      = hm1[c];
            q[j]        = dvs.quot(x[j]);
        }
        if (dvs.bad()) {
#pragma unroll
            for (int j = 0; j < kOutPer; ++j) q[j] = dvs.exact(x[j]);
        }
#pragma unroll
        for (int j = 0; j < kOutPer; ++j) {
            const float ve = (e ? red[j] : q[j]) * out_scale;
            const float vo = (e ? q[j] : red[j]) * out_scale;
            buf_store2<St>(dst, st_off[j], ve, vo, so);
        }
    };
    // plane k; register sets alternate with k - z0 (set 0 at k = z0 - 1): old(m), bh(m)
    // and bf(m) live in set (m - z0 + 1) & 1. Ring slots rotate: old planes k - 1, k,
    // k + 1 in (o_a, o_b, o_c), H planes k - 2, k - 1, k in (h_a, h_b, h_c); old(k + 1)
    // and H(k) take the slots that held old(k - 2) and H(k - 3).
    float *o_a = s_old[0], *o_b = s_old[1], *o_c = s_old[2];
    float *h_a = s_h[0], *h_b = s_h[1], *h_c = s_h[2];
    auto body = [&](const int k, OldRegs& o_next, HRegs& h_cur, float (&f_prev)[kOutPer]) {
        stash_old(k + 1, o_c, o_next);
        fetch_old(k + 3, o_next);
        __syncthreads();
        compute_h(k, o_a, o_b, o_c, h_c, h_cur);
        fetch_bh(k + 2, h_cur);
        __syncthreads();
        compute_final(k, h_a, h_b, h_c, f_prev);
        fetch_bf(k + 1, f_prev);
        float* o_t = o_a;
        float* h_t = h_a;
        o_a           = o_b;
        o_b           = o_c;
        o_c           = o_t;
        h_a           = h_b;
        h_b           = h_c;
        h_c           = h_t;
    };

    fetch_old(z0 - 2, oa);
    stash_old(z0 - 2, s_old[0], oa);
    fetch_old(z0 - 1, oa);
    stash_old(z0 - 1, s_old[1], oa);
    fetch_old(z0, ob);     // set 1
    fetch_old(z0 + 1, oa); // set 0
    fetch_bh(z0 - 1, ha);
    fetch_bh(z0, hb);
    // an odd number of planes z0 - 1 .. z1 ends with one idle plane
    for (int k = z0 - 1; k <= z1; k += 2) {
        body(k, ob, ha, fb);
        body(k + 1, oa, hb, fa);
    }
}

template <int kSolve, bool kZero, class St, bool kPairs>
__device__ __forceinline__ void rbgs_split_body(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC& bc, const TileGeom& g, const float* __restrict__ scale_slot, const int w) {
    __shared__ float s_old[3][kOldPlane];
    __shared__ float s_h[3][kHPlane];
    const float inv_s = rhs_scale<St>(scale_slot);

    int tx, ty, cz;
    tile_of(g, w, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int x0   = tx * kTX, y0 = ty * kTY;
    const int xs   = x0 - 4, ys = y0 - 2;
    const int z0   = cz * g.zc;
    const int z1   = min(z0 + g.zc, d.nz);
    const bool px  = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint64_t plane = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny);

    // Old-plane load slots: contiguous floats of the (2*kW)-wide region rows, stored to the
    // even or odd LDS array by column parity. Lanes past the region repeat its last slot
    // (same address, same value), so no lane needs a branch. With even rows (kPairs) a slot
    // is an aligned pair: .x goes to the even array, .y to the odd one; x-ghost pairs read
    // the adjacent edge cell (sel 1: both from .x, sel 2: both from .y).
    constexpr int kSlots = kPairs ? kOldPairPer : kOldPer;
    uint32_t o_off[kSlots];
    int o_gc[kSlots], o_lds[kSlots], o_sel[kSlots];
    bool any_gc = false;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
        if constexpr (kPairs) {
            const int e   = min(tid + j * kThreads, kOldPairs - 1);
            const int r   = e / kW, i = e % kW;
            const int gx  = xs + 2 * i, gy = ys + r;
            const int row = fix_coord(gy, d.ny, py) * d.nx;
            o_sel[j]      = 0;
            if (!px && gx < 0) {
                o_off[j] = static_cast<uint32_t>(row);
                o_sel[j] = 1;
            } else if (!px && gx >= d.nx) {
                o_off[j] = static_cast<uint32_t>(row + d.nx - 2);
                o_sel[j] = 2;
            } else {
                o_off[j] = static_cast<uint32_t>(row + fix_coord(gx, d.nx, px));
            }
            o_gc[j]  = ghost_code(gx, gy, d, px, py);
            o_lds[j] = r * kW + i;
        } else {
            const int e  = min(tid + j * kThreads, kOldPlane - 1);
            const int r  = e / (2 * kW), c = e % (2 * kW);
            const int gx = xs + c, gy = ys + r;
            o_off[j]     = static_cast<uint32_t>(fix_coord(gy, d.ny, py) * d.nx + fix_coord(gx, d.nx, px));
            o_gc[j]      = ghost_code(gx, gy, d, px, py);
            o_lds[j]     = (c & 1) * kArr + r * kW + (c >> 1);
            o_sel[j]     = 0;
        }
        any_gc = any_gc || o_gc[j] != 0;
    }
    const bool wave_gc = __builtin_amdgcn_readfirstlane(__any(any_gc)) != 0;

    // H slots: row r in [1, kTY+3), pair i in [1, kPX+3), clamped like the load slots. At
    // plane parity q the red cell is the even one when (ys + r + q) is even. One array per
    // parity: a runtime select between two elements of one array would become a scratch load.
    int h_ca0[kHSlots3], h_ca1[kHSlots3], h_cx0[kHSlots3], h_cx1[kHSlots3], h_st[kHSlots3];
    uint32_t h_ob0[kHSlots3], h_ob1[kHSlots3];
    bool h_rl0[kHSlots3], h_rl1[kHSlots3]; // the red cell at parity q is a real cell
#pragma unroll
    for (int j = 0; j < kHSlots3; ++j) {
        const int p = min(tid + j * kThreads, kHVals - 1);
        const int r = 1 + p / (kPX + 2), i = 1 + p % (kPX + 2);
        const int y = ys + r;
        h_st[j]     = r * kW + i;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int a       = (y + q) & 1; // 0: red in the even array
            const int ca      = a * kArr + r * kW + i;
            const int cx      = (1 - a) * kArr + r * kW + (a == 0 ? i - 1 : i);
            const int gx      = xs + 2 * i + a;
            const bool rl     = (py || (y >= 0 && y < d.ny)) && (px || (gx >= 0 && gx < d.nx));
            const uint32_t ob = static_cast<uint32_t>(fix_coord(y, d.ny, py) * d.nx + fix_coord(gx, d.nx, px));
            if (q == 0) {
                h_ca0[j] = ca;
                h_cx0[j] = cx;
                h_ob0[j] = ob;
                h_rl0[j] = rl;
            } else {
                h_ca1[j] = ca;
                h_cx1[j] = cx;
                h_ob1[j] = ob;
                h_rl1[j] = rl;
            }
        }
    }
    // output: wavefront w handles rows w and w+8, lane = pair; row validity is wave-uniform
    const int lane     = tid & 63;
    const int wrow     = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool full_x  = x0 + kTX <= d.nx && (d.nx & 1) == 0;
    const bool even_nx = (d.nx & 1) == 0;

    auto plane_of = [&](const int k) { return pz ? wrap_near(k, d.nz) : (k < 0 ? 0 : (k >= d.nz ? d.nz - 1 : k)); };
    auto real_plane = [&](const int k) { return pz || (k >= 0 && k < d.nz); };

    float r_old[kPairs ? 1 : kOldPer], r_bh[kHSlots3], r_bf[kOutPer];
    float2 r_pair[kPairs ? kOldPairPer : 1];
    auto fetch_old = [&](const int k) {
        const typename St::T* src = x_old + static_cast<uint64_t>(plane_of(k)) * plane;
        if constexpr (kPairs) {
#pragma unroll
            for (int j = 0; j < kOldPairPer; ++j) {
                float2 v = kZero ? make_float2(0.0f, 0.0f) : St::load2(src + o_off[j]);
                if (o_sel[j] == 1) v.y = v.x;
                if (o_sel[j] == 2) v.x = v.y;
                r_pair[j] = v;
            }
            if (!pz && (k < 0 || k >= d.nz)) {
                const FaceRule f = k < 0 ? FaceRule{bc.face[4].mode, bc.face[4].value} : FaceRule{bc.face[5].mode, bc.face[5].value};
#pragma unroll
                for (int j = 0; j < kOldPairPer; ++j) r_pair[j] = make_float2(apply_ghost(f, r_pair[j].x), apply_ghost(f, r_pair[j].y));
            } else if (wave_gc) {
#pragma unroll
                for (int j = 0; j < kOldPairPer; ++j) {
                    const FaceRule f = face_rule(bc, o_gc[j] - 1);
                    const float2 gv  = make_float2(apply_ghost(f, r_pair[j].x), apply_ghost(f, r_pair[j].y));
                    r_pair[j]        = o_gc[j] ? gv : r_pair[j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < kOldPer; ++j) r_old[j] = kZero ? 0.0f : St::load(src[o_off[j]]);
            if (!pz && (k < 0 || k >= d.nz)) {
                const FaceRule f = k < 0 ? FaceRule{bc.face[4].mode, bc.face[4].value} : FaceRule{bc.face[5].mode, bc.face[5].value};
#pragma unroll
                for (int j = 0; j < kOldPer; ++j) r_old[j] = apply_ghost(f, r_old[j]);
            } else if (wave_gc) {
#pragma unroll
                for (int j = 0; j < kOldPer; ++j) {
                    const float gv = apply_ghost(face_rule(bc, o_gc[j] - 1), r_old[j]);
                    r_old[j]       = o_gc[j] ? gv : r_old[j];
                }
            }
        }
    };
    auto fetch_bh = [&](const int k, const bool odd) {
        const float* src = b + static_cast<uint64_t>(plane_of(k)) * plane;
#pragma unroll
        for (int j = 0; j < kHSlots3; ++j) {
            r_bh[j] = src[odd ? h_ob1[j] : h_ob0[j]];
            if constexpr (St::kNarrow) r_bh[j] *= inv_s;
        }
    };
    auto fetch_bf = [&](const int k) {
        const float* src = b + static_cast<uint64_t>(k) * plane;
#pragma unroll
        for (int j = 0; j < kOutPer; ++j) {
            const int y = min(y0 + wrow + (kThreads / 64) * j, d.ny - 1);
            const int e = ((y + k) & 1) == 0 ? 1 : 0; // black cell: odd when the even one is red
            const int x = min(x0 + 2 * lane + e, d.nx - 1);
            r_bf[j]     = src[static_cast<uint32_t>(y * d.nx + x)];
            if constexpr (St::kNarrow) r_bf[j] *= inv_s;
        }
    };
    auto stash_old = [&](float* s) {
        if constexpr (kPairs) {
#pragma unroll
            for (int j = 0; j < kOldPairPer; ++j) {
                s[o_lds[j]]        = r_pair[j].x;
                s[kArr + o_lds[j]] = r_pair[j].y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kOldPer; ++j) s[o_lds[j]] = r_old[j];
        }
    };

    // LDS rings, rotated each plane: old planes k-1, k, k+1 and H planes k-2, k-1, k
    float* om = s_old[0];
    float* o0 = s_old[1];
    float* op = s_old[2];
    float* hm2 = s_h[0];
    float* hm1 = s_h[1];
    float* h0  = s_h[2];

    // one plane of the pipeline; Q is the parity of k (compile time)
    auto body = [&](const int k, auto qc) {
        constexpr int Q = decltype(qc)::value;
        stash_old(op); // old(k+1), fetched one plane earlier
        float bh[kHSlots3], bf[kOutPer];
#pragma unroll
        for (int j = 0; j < kHSlots3; ++j) bh[j] = r_bh[j];
#pragma unroll
        for (int j = 0; j < kOutPer; ++j) bf[j] = r_bf[j];
        if (k + 1 <= z1) {
            fetch_old(k + 2);
            if (real_plane(k + 1)) fetch_bh(k + 1, Q == 0);
        }
        if (k >= z0 && k < z1) fetch_bf(k);
        __syncthreads();

        // H(k): red value of every pair on the tile plus a one-cell halo
        if (real_plane(k)) {
#pragma unroll
            for (int j = 0; j < kHSlots3; ++j) {
                const int ca   = Q ? h_ca1[j] : h_ca0[j];
                const int cx   = Q ? h_cx1[j] : h_cx0[j];
                const bool rl  = Q ? h_rl1[j] : h_rl0[j];
                float sum      = o0[cx];
                sum            = sum + o0[cx + 1];
                sum            = sum + o0[ca - kW];
                sum            = sum + o0[ca + kW];
                sum            = sum + om[ca];
                sum            = sum + op[ca];
                const float v  = relax_fast<kSolve>(bh[j], sum, coef);
                h0[h_st[j]]    = rl ? v : o0[ca];
            }
        } else {
#pragma unroll
            for (int j = 0; j < kHSlots3; ++j) h0[h_st[j]] = o0[Q ? h_ca1[j] : h_ca0[j]];
        }
        __syncthreads();

        // final plane k-1: black cells relaxed from the red values, red cells copied
        const int kf = k - 1;
        if (kf >= z0 && kf < z1) {
            auto* dst_plane = x_new + static_cast<uint64_t>(kf) * plane;
#pragma unroll
            for (int j = 0; j < kOutPer; ++j) {
                const int rr = wrow + (kThreads / 64) * j;
                const int y  = y0 + rr;
                if (y >= d.ny) continue; // wave-uniform
                const int c         = (rr + 2) * kW + (lane + 2);
                const bool even_red = ((y + (1 - Q)) & 1) == 0; // parity of kf is 1 - Q
                const float red     = hm1[c];
                float sum;
                if (even_red) { // black is the odd cell: x- = red[i], x+ = red[i+1]
                    sum = hm1[c];
                    sum = sum + hm1[c + 1];
                } else { // black is the even cell: x- = red[i-1], x+ = red[i]
                    sum = hm1[c - 1];
                    sum = sum + hm1[c];
                }
                sum               = sum + hm1[c - kW];
                sum               = sum + hm1[c + kW];
                sum               = sum + hm2[c];
                sum               = sum + h0[c];
                const float black = relax_fast<kSolve>(bf[j], sum, coef);
                const float ve    = (even_red ? red : black) * out_scale;
                const float vo    = (even_red ? black : red) * out_scale;
                const int x       = x0 + 2 * lane;
                auto* dst         = dst_plane + static_cast<uint32_t>(y * d.nx + x);
                if (full_x) {
                    store_vec2<St>(dst, ve, vo);
                } else if (even_nx) {
                    if (x < d.nx) store_vec2<St>(dst, ve, vo);
                } else {
                    if (x < d.nx) dst[0] = St::store(ve);
                    if (x + 1 < d.nx) dst[1] = St::store(vo);
                }
            }
        }
        __syncthreads();
        float* t = om;
        om       = o0;
        o0       = op;
        op       = t;
        t        = hm2;
        hm2      = hm1;
        hm1      = h0;
        h0       = t;
    };

    // prologue: old planes z0-2, z0-1 in LDS, old z0 and b for H(z0-1) in registers
    fetch_old(z0 - 2);
    stash_old(om);
    fetch_old(z0 - 1);
    stash_old(o0);
    fetch_old(z0);
    if (real_plane(z0 - 1)) fetch_bh(z0 - 1, ((z0 - 1) & 1) != 0);

    using Even = std::integral_constant<int, 0>;
    using Odd  = std::integral_constant<int, 1>;
    int k = z0 - 1;
    if (k & 1) {
        body(k, Odd{});
        ++k;
    }
    for (; k + 1 <= z1; k += 2) {
        body(k, Even{});
        body(k + 1, Odd{});
    }
    if (k <= z1) body(k, Even{});
}

template <int kSolve, bool kZero, class St, bool kPairs>
__global__ __launch_bounds__(kThreads) void k_jacobi_pass(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC bc, const TileGeom g, const float* __restrict__ scale_slot, const int* __restrict__ skip) {
    if (skip != nullptr && *skip != 0) return;
    if constexpr (kPairs)
        jacobi_rows_body<kSolve, kZero, St>(x_old, b, x_new, coef, out_scale, d, bc, g, scale_slot, xcd_block());
    else
        jacobi_body<kSolve, kZero, St>(x_old, b, x_new, coef, out_scale, d, bc, g, scale_slot, xcd_block());
}

template <int kSolve, bool kZero, class St, bool kPairs, bool kWhole = false>
__global__ __launch_bounds__(kThreads) void k_rbgs_split_pass(const typename St::T* __restrict__ x_old, const float* __restrict__ b, typename St::T* __restrict__ x_new, const float coef, const float out_scale, const Dims d, const StencilBC bc, const TileGeom g, const float* __restrict__ scale_slot, const int* __restrict__ skip) {
    if (skip != nullptr && *skip != 0) return;
    if constexpr (kPairs)
        rbgs_rows_body<kSolve, kZero, St, kWhole>(x_old, b, x_new, coef, out_scale, d, bc, g, scale_slot, xcd_block());
    else
        rbgs_split_body<kSolve, kZero, St, false>(x_old, b, x_new, coef, out_scale, d, bc, g, scale_slot, xcd_block());
}

// Batched passes: workgroup i runs block items[i].block of pass descs[items[i].desc]. The
// descriptor is read in place (a local copy of it would be indexed into scratch memory);
// all items of one launch share the row parity kPairs.
template <int kSolve, bool kZero, bool kPairs>
__global__ __launch_bounds__(kThreads) void k_rbgs_split_batched(const SmoothDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it   = items[blockIdx.x];
    const SmoothDesc& sd = descs[it.desc];
    if constexpr (kPairs)
        rbgs_rows_body<kSolve, kZero, StoreF32, false>(sd.x_old, sd.b, sd.x_new, sd.coef, sd.out_scale, sd.d, sd.bc, sd.g, nullptr, static_cast<int>(it.block));
    else
        rbgs_split_body<kSolve, kZero, StoreF32, false>(sd.x_old, sd.b, sd.x_new, sd.coef, sd.out_scale, sd.d, sd.bc, sd.g, nullptr, static_cast<int>(it.block));
}

template <int kSolve, bool kZero, bool kPairs>
__global__ __launch_bounds__(kThreads) void k_jacobi_batched(const SmoothDesc* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it   = items[blockIdx.x];
    const SmoothDesc& sd = descs[it.desc];
    if constexpr (kPairs)
        jacobi_rows_body<kSolve, kZero, StoreF32>(sd.x_old, sd.b, sd.x_new, sd.coef, sd.out_scale, sd.d, sd.bc, sd.g, nullptr, static_cast<int>(it.block));
    else
        jacobi_body<kSolve, kZero, StoreF32>(sd.x_old, sd.b, sd.x_new, sd.coef, sd.out_scale, sd.d, sd.bc, sd.g, nullptr, static_cast<int>(it.block));
}

// Spins (s_sleep 4, about 0.1 us each) one waiting lane may spend over the whole launch:
// about 0.4 s, far above a solve. The ticket order below guarantees every wait ends; the
// budget only keeps the launch finite if that guarantee were ever broken.
constexpr unsigned kDepSpins = 1u << 22;

// `passes` RBGS passes from zero in one launch, run only when skip[0] == 0 (the learned
// warm start's fallback): io -> other -> io -> ... (the result lands in other for an odd
// count, in io for an even one); the first pass reads the implicit zero iterate, as the
// single pass from zero does. The work is the items (pass j, block t), handed out in
// pass-major order by an atomic ticket (ctr[0]); block t is the single pass's block t
// (tile_of). Before item (j, t) runs, the blocks whose cells its halo reads (its own and
// the up to 26 around it, wrapped on periodic axes) must have finished pass j - 1
// (done[u] = ctr[4 + u] counts the passes block u has finished): that covers the reads of
// pass j - 1's output and the stores into the buffer pass j - 1 read. Every item waited
// for has a smaller ticket, taken by a workgroup that is already running and waits only
// for smaller tickets itself, so the launch completes whether or not all its workgroups
// are resident at once (other kernels may hold part of the GPU). Stores are published at
// device scope before a block's count moves (the XCDs' L2s are not coherent with each
// other) and stale lines dropped after a wait; counts move by atomic read-modify-write,
// which other XCDs observe (a plain store would sit in the writer's L2). The last
// workgroup to leave resets the counters for the next launch. A waiting lane that used
// up its spin budget (kDepSpins over the launch) waits no more and sets skip[2]; the last
// workgroup then fills the result with NaN so the failure cannot pass unseen.
template <int kSolve>
__global__ __launch_bounds__(kThreads) void k_rbgs_chain(float* io, float* other, const float* __restrict__ b, const float coef, const int passes, const Dims d, const StencilBC bc, const TileGeom g, int* __restrict__ skip, unsigned* __restrict__ ctr) {
    if (skip[0] != 0) return;
    __shared__ unsigned s_ticket;
    __shared__ int s_late;
    unsigned* ticket = ctr;
    unsigned* exits  = ctr + 1;
    unsigned* done   = ctr + 4;
    const int n           = g.tiles_x * g.tiles_y * g.chunks;
    const unsigned total  = static_cast<unsigned>(passes) * static_cast<unsigned>(n);
    const bool px = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    unsigned budget = kDepSpins; // this lane's spins left over the launch
    for (;;) {
        if (threadIdx.x == 0) s_ticket = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const unsigned k = s_ticket;
        if (k >= total) break;
        const int j = static_cast<int>(k / static_cast<unsigned>(n));
        const int t = static_cast<int>(k % static_cast<unsigned>(n));
        if (j > 0) {
            // lanes 0 .. 26 of the first wave each wait for one neighbour block
            if (threadIdx.x < 27) {
                int tx, ty, cz;
                tile_of(g, t, tx, ty, cz);
                const int q = static_cast<int>(threadIdx.x);
                int ux = tx + q % 3 - 1, uy = ty + (q / 3) % 3 - 1, uz = cz + q / 9 - 1;
                bool in = true;
                if (ux < 0 || ux >= g.tiles_x) { ux = px ? wrap_index(ux, g.tiles_x) : 0; in = in && px; }
                if (uy < 0 || uy >= g.tiles_y) { uy = py ? wrap_index(uy, g.tiles_y) : 0; in = in && py; }
                if (uz < 0 || uz >= g.chunks) { uz = pz ? wrap_index(uz, g.chunks) : 0; in = in && pz; }
                if (in) {
                    const int u = (uy * g.tiles_x + ux) * g.chunks + uz;
                    while (budget > 0 && __hip_atomic_load(done + u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < static_cast<unsigned>(j)) {
                        if (--budget == 0) __hip_atomic_fetch_or(skip + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __builtin_amdgcn_s_sleep(4);
                    }
                }
            }
            __syncthreads();
            __threadfence();
        }
        const float* src = (j & 1) ? other : io;
        float* dst       = (j & 1) ? io : other;
        if (j == 0) rbgs_rows_body<kSolve, true, StoreF32, true>(src, b, dst, coef, 1.0f, d, bc, g, nullptr, t);
        else rbgs_rows_body<kSolve, false, StoreF32, true>(src, b, dst, coef, 1.0f, d, bc, g, nullptr, t);
        __threadfence();
        __syncthreads(); // every wave's stores published; the LDS rings free for the next item
        if (threadIdx.x == 0) __hip_atomic_fetch_add(done + t, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); // j -> j + 1
    }
    // leave; the last workgroup out resets the counters (and poisons a failed result)
    __threadfence();
    if (threadIdx.x == 0) s_late = __hip_atomic_fetch_add(exits, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u;
    __syncthreads();
    if (!s_late) return;
    __threadfence();
    for (int i = static_cast<int>(threadIdx.x); i < n; i += kThreads) done[i] = 0u;
    if (threadIdx.x == 0) {
        ticket[0] = 0u;
        exits[0]  = 0u;
    }
    if (__hip_atomic_fetch_or(skip + 2, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        float* result        = (passes & 1) ? other : io;
        const uint64_t cells = static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny) * static_cast<uint64_t>(d.nz);
        for (uint64_t i = threadIdx.x; i < cells; i += kThreads) result[i] = __builtin_nanf("");
    }
}

bool has_odd_periodic(const Dims d, const StencilBC& bc) {
    return ((bc.periodic & 1) && (d.nx & 1)) || ((bc.periodic & 2) && (d.ny & 1)) || ((bc.periodic & 4) && (d.nz & 1));
}

} // namespace

TileGeom plan_tiles(const Dims d, const int zc_override) {
    TileGeom g{};
    g.tiles_x        = (d.nx + kTX - 1) / kTX;
    g.tiles_y        = (d.ny + kTY - 1) / kTY;
    const long tiles = static_cast<long>(g.tiles_x) * g.tiles_y;
    // about 512 workgroups (two 512-thread groups per CU), at least 4 planes per chunk
    long zc = (static_cast<long>(d.nz) * tiles + 511) / 512;
    if (zc < 4) zc = 4;
    if (zc > 64) zc = 64;
    if (zc_override > 0) zc = zc_override;
    if (zc > d.nz) zc = d.nz;
    g.zc     = static_cast<int>(zc);
    g.chunks = (d.nz + g.zc - 1) / g.zc;
    return g;
}

uint32_t smooth_blocks(const TileGeom& g) { return static_cast<uint32_t>(g.tiles_x * g.tiles_y * g.chunks); }

bool smooth_batchable(const SmootherKind smoother, const Dims d, const StencilBC& bc) { return smoother == kSmoothJacobi || !has_odd_periodic(d, bc); }

hipError_t launch_smooth_batched(hipStream_t s, const SmootherKind smoother, const SolveKind solve, const bool zero, const bool pairs, const SmoothDesc* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
#define SF_B(KERNEL, SO, Z)                                                          \
    do {                                                                             \
        if (pairs) KERNEL<SO, Z, true><<<n_items, kThreads, 0, s>>>(descs, items);   \
        else KERNEL<SO, Z, false><<<n_items, kThreads, 0, s>>>(descs, items);        \
    } while (0)
    if (smoother == kSmoothRBGS) {
        if (solve == kSolvePoisson) {
            if (zero) SF_B(k_rbgs_split_batched, kSolvePoisson, true);
            else SF_B(k_rbgs_split_batched, kSolvePoisson, false);
        } else {
            if (zero) SF_B(k_rbgs_split_batched, kSolveDiffusion, true);
            else SF_B(k_rbgs_split_batched, kSolveDiffusion, false);
        }
    } else {
        if (solve == kSolvePoisson) {
            if (zero) SF_B(k_jacobi_batched, kSolvePoisson, true);
            else SF_B(k_jacobi_batched, kSolvePoisson, false);
        } else {
            if (zero) SF_B(k_jacobi_batched, kSolveDiffusion, true);
            else SF_B(k_jacobi_batched, kSolveDiffusion, false);
        }
    }
#undef SF_B
    return hipGetLastError();
}


// Whether a field of `elem` byte elements is below 2 GiB (one buffer resource spans it).
static bool field_below_2g(const Dims d, const int elem) { return static_cast<uint64_t>(elem) * static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny) * static_cast<uint64_t>(d.nz) < (1ull << 31); }

// Compute units of the current device, read once per device.
int device_cus() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

hipError_t launch_smooth(hipStream_t s, const SmootherKind smoother, const SolveKind solve, const SmoothPass& pass, const Dims d, const StencilBC& bc) {
    // Jacobi passes run at most one 512-thread workgroup per CU (long z-chunks, not capped
    // at plan_tiles' 64 planes): 33.2 us at 256^3 against 35.1 us with two per CU and 37.4
    // us with four; 290 us at 512^3 (256-plane chunks) against 327 us with 64-plane ones.
    // The chunk count comes first, so the workgroups never outnumber the CUs by a few (the
    // 128x160x128 grid: 22 chunks of 6 planes, 220 workgroups, not 26 of 5, 260).
    const long tiles_j    = static_cast<long>((d.nx + kTX - 1) / kTX) * ((d.ny + kTY - 1) / kTY);
    const long cus        = device_cus();
    const long chunks_j   = std::max<long>(1, cus / tiles_j);
    const int zc_j        = static_cast<int>(std::max<long>(4, (d.nz + chunks_j - 1) / chunks_j));
    const TileGeom g      = plan_tiles(d, smoother == kSmoothJacobi ? zc_j : 0);
    const unsigned blocks = static_cast<unsigned>(g.tiles_x * g.tiles_y * g.chunks);
    const bool zero       = pass.x_old == nullptr;
    const bool whole      = field_below_2g(d, 4);
#define SF_PASS_GENERAL(SO, Z) k_rbgs_pass<SO, Z><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, pass.skip)
#define SF_JPASS(SO, Z)                                                                                                                                                            \
    do {                                                                                                                                                                            \
        if ((d.nx & 1) == 0) k_jacobi_pass<SO, Z, StoreF32, true><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip); \
        else k_jacobi_pass<SO, Z, StoreF32, false><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip);         \
    } while (0)
#define SF_SPASS(SO, Z)                                                                                                                                                                \
    do {                                                                                                                                                                                \
        if ((d.nx & 1) == 0 && whole) k_rbgs_split_pass<SO, Z, StoreF32, true, true><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip); \
        else if ((d.nx & 1) == 0) k_rbgs_split_pass<SO, Z, StoreF32, true><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip); \
        else k_rbgs_split_pass<SO, Z, StoreF32, false><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip);         \
    } while (0)
#define SF_PASS(KERNEL, SO, Z) KERNEL<SO, Z, StoreF32><<<blocks, kThreads, 0, s>>>(pass.x_old, pass.b, pass.x_new, pass.coef, pass.out_scale, d, bc, g, nullptr, pass.skip)
    if (smoother == kSmoothRBGS && !has_odd_periodic(d, bc)) {
        if (solve == kSolvePoisson) {
            if (zero) SF_SPASS(kSolvePoisson, true);
            else SF_SPASS(kSolvePoisson, false);
        } else {
            if (zero) SF_SPASS(kSolveDiffusion, true);
            else SF_SPASS(kSolveDiffusion, false);
        }
    } else if (smoother == kSmoothRBGS) {
        if (solve == kSolvePoisson) {
            if (zero) SF_PASS_GENERAL(kSolvePoisson, true);
            else SF_PASS_GENERAL(kSolvePoisson, false);
        } else {
            if (zero) SF_PASS_GENERAL(kSolveDiffusion, true);
            else SF_PASS_GENERAL(kSolveDiffusion, false);
        }
    } else {
        if (solve == kSolvePoisson) {
            if (zero) SF_JPASS(kSolvePoisson, true);
            else SF_JPASS(kSolvePoisson, false);
        } else {
            if (zero) SF_JPASS(kSolveDiffusion, true);
            else SF_JPASS(kSolveDiffusion, false);
        }
    }
#undef SF_PASS
#undef SF_SPASS
#undef SF_JPASS
#undef SF_PASS_GENERAL
    return hipGetLastError();
}

bool smooth_chain_supported(const SmootherKind smoother, const Dims d, const StencilBC& bc) {
    return smoother == kSmoothRBGS && (d.nx & 1) == 0 && !has_odd_periodic(d, bc) && field_below_2g(d, 4);
}

hipError_t smooth_chain_grid(const Dims d, uint32_t* out_grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rbgs_chain<kSolvePoisson>, kThreads, 0);
    if (e != hipSuccess) return e;
    int per_cu_d = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_d, k_rbgs_chain<kSolveDiffusion>, kThreads, 0);
    if (e != hipSuccess) return e;
    const uint64_t resident = static_cast<uint64_t>(cus) * static_cast<uint64_t>(std::min(per_cu, per_cu_d));
    if (resident == 0) return hipErrorInvalidConfiguration;
    const uint64_t n = smooth_blocks(plan_tiles(d, 0));
    // no more workgroups than can be resident at once (they need not all be)
    const uint64_t grid = std::min(n, resident);
    *out_grid = static_cast<uint32_t>(grid);
    return hipSuccess;
}

uint32_t smooth_chain_counters(const Dims d) { return 4u + smooth_blocks(plan_tiles(d, 0)); }

hipError_t launch_smooth_chain(hipStream_t s, const SolveKind solve, float* io, float* other, const float* b, const float coef, const int passes, const Dims d, const StencilBC& bc, int* skip, unsigned* counters, const uint32_t grid) {
    if (passes <= 0 || io == other || grid == 0) return hipErrorInvalidValue;
    const TileGeom g = plan_tiles(d, 0);
    if (solve == kSolvePoisson) k_rbgs_chain<kSolvePoisson><<<grid, kThreads, 0, s>>>(io, other, b, coef, passes, d, bc, g, skip, counters);
    else k_rbgs_chain<kSolveDiffusion><<<grid, kThreads, 0, s>>>(io, other, b, coef, passes, d, bc, g, skip, counters);
    return hipGetLastError();
}

namespace {

template <class St>
hipError_t correction_pass(hipStream_t s, const SmootherKind smoother, const SolveKind solve, const CorrectionPass& pass, const Dims d, const StencilBC& bc) {
    const TileGeom g      = plan_tiles(d, 0);
    const unsigned blocks = static_cast<unsigned>(g.tiles_x * g.tiles_y * g.chunks);
    const auto* e_old     = static_cast<const typename St::T*>(pass.e_old);
    auto* e_new           = static_cast<typename St::T*>(pass.e_new);
    const bool whole      = field_below_2g(d, 4); // the right-hand side r is FP32
#define SF_SCORR(SO, Z)                                                                                                                                      \
    do {                                                                                                                                                     \
        if ((d.nx & 1) == 0 && whole) k_rbgs_split_pass<SO, Z, St, true, true><<<blocks, kThreads, 0, s>>>(e_old, pass.r, e_new, pass.coef, 1.0f, d, bc, g, pass.scale_slot, nullptr); \
        else if ((d.nx & 1) == 0) k_rbgs_split_pass<SO, Z, St, true><<<blocks, kThreads, 0, s>>>(e_old, pass.r, e_new, pass.coef, 1.0f, d, bc, g, pass.scale_slot, nullptr); \
        else k_rbgs_split_pass<SO, Z, St, false><<<blocks, kThreads, 0, s>>>(e_old, pass.r, e_new, pass.coef, 1.0f, d, bc, g, pass.scale_slot, nullptr);         \
    } while (0)
#define SF_JCORR(SO, Z)                                                                                                                                      \
    do {                                                                                                                                                     \
        if ((d.nx & 1) == 0) k_jacobi_pass<SO, Z, St, true><<<blocks, kThreads, 0, s>>>(e_old, pass.r, e_new, pass.coef, 1.0f, d, bc, g, pass.scale_slot, nullptr); \
        else k_jacobi_pass<SO, Z, St, false><<<blocks, kThreads, 0, s>>>(e_old, pass.r, e_new, pass.coef, 1.0f, d, bc, g, pass.scale_slot, nullptr);         \
    } while (0)
    const bool zero = e_old == nullptr;
    if (smoother == kSmoothRBGS) {
        if (solve == kSolvePoisson) {
            if (zero) SF_SCORR(kSolvePoisson, true);
            else SF_SCORR(kSolvePoisson, false);
        } else {
            if (zero) SF_SCORR(kSolveDiffusion, true);
            else SF_SCORR(kSolveDiffusion, false);
        }
    } else {
        if (solve == kSolvePoisson) {
            if (zero) SF_JCORR(kSolvePoisson, true);
            else SF_JCORR(kSolvePoisson, false);
        } else {
            if (zero) SF_JCORR(kSolveDiffusion, true);
            else SF_JCORR(kSolveDiffusion, false);
        }
    }
#undef SF_SCORR
#undef SF_JCORR
    return hipGetLastError();
}

} // namespace

hipError_t launch_correction(hipStream_t s, const SmootherKind smoother, const SolveKind solve, const StorageKind storage, const CorrectionPass& pass, const Dims d, const StencilBC& homogeneous_bc) {
    // the split RBGS kernel covers every grid except an odd periodic axis
    if (smoother == kSmoothRBGS && has_odd_periodic(d, homogeneous_bc)) return hipErrorInvalidValue;
    switch (storage) {
    case kStoreF16: return correction_pass<StoreF16>(s, smoother, solve, pass, d, homogeneous_bc);
    case kStoreBF16: return correction_pass<StoreBF16>(s, smoother, solve, pass, d, homogeneous_bc);
    default: return correction_pass<StoreF32>(s, smoother, solve, pass, d, homogeneous_bc);
    }
}

} // namespace sf
