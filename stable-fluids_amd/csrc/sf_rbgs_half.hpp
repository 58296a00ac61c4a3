// sf_rbgs_half.hpp -- red-black Gauss-Seidel passes that carry only the black cells from one
// pass to the next (gfx950).
//
// One pass is one reference RBGS iteration, a red half-sweep then a black one
// (backend-cuda.cu:354-375, :398-407, loops :628-640, :646-649):
//   diffusion: x = (b + c * sum6) / (1 + 6c)       Poisson: x = (sum6 - c * b) / 6
// A red cell is relaxed from its six neighbours, which are all black, and a black cell
// from six red ones. So the old value of a red cell is read only where the cell sits on a
// non-periodic face, whose ghost is a rule of the cell itself (copy, 2 v_face - value or a
// constant, backend-cuda.cu:62-189). Between passes the iterate is therefore kept as
//   * the black cell of every x pair, in a "half" buffer of nx/2 floats per row, and
//   * the ghost values of the red face cells (the rule applied to the red value) in the
//     padding of that buffer: rows -1 and ny, planes -1 and nz, and two short x-face
//     columns per plane.
// A middle pass reads 4 B/cell of b and 2 B/cell of black cells and writes 2 B/cell
// (8 B/cell against 12 for a full-iterate pass). The first pass of a solve reads the
// natural-layout start iterate (or none: zeros) and the last one writes the natural-layout
// result, both colours, scaled by out_scale.
//
// Padded half plane k (k = -1 .. nz; planes -1 and nz are the z-face ghosts), ps floats:
//   entry (pair i, row y), y = -1 .. ny:  (y + 1) * nxh + i
//   x-face ghosts (side s, row y):         (ny + 2) * nxh + s * ny + y
// A pair (2i, 2i + 1) of row y, plane k holds its red cell at x = 2i + p, p = (y + k) & 1,
// and its black cell at 2i + 1 - p. The six black neighbours of red cell i are the
// black cells i - 1 + p and i + p of its row, and i of rows y -+ 1 and planes k -+ 1; the
// red neighbours of black cell i are red cells i - p and i + 1 - p of its row and i of
// the rows and planes above and below. Every sum keeps the reference order x-, x+, y-,
// y+, z-, z+, and divisions are the correctly rounded reciprocal forms (sf_relax.hpp),
// so a solve is bit-identical to the in-place reference iteration.
//
// Work decomposition. A 512-thread workgroup owns 64 pairs x 8*kRowsW rows of a z-chunk
// and marches upward. Per plane step s, with one barrier:
//   stage A(s+1) (black cells of the old iterate, tile plus two pairs and two rows of
//   halo, ghosts filled) into a ring of four LDS planes -> barrier -> red half-update
//   R(s) on the tile plus one pair and one row of halo (the halo is recomputed, so tiles
//   exchange nothing) into a ring of two LDS planes -> black cells of plane s - 1 from
//   R(s-2 .. s) -> stores.
// A lane holds the same pair of its rows in every plane, so the z neighbours of its own
// cells and one of the two x neighbours come from registers; LDS supplies the other x
// neighbour and the y neighbours. Loads run two planes ahead through buffer resources
// (idle lanes get an out-of-range offset, idle planes a 0-byte resource), so no memory
// instruction sits under a branch.
#pragma once

#include <algorithm>
#include <type_traits>
#include <utility>

#include "sf_buffer.hpp"
#include "sf_kernels.hpp"
#include "sf_relax.hpp"

namespace sf {

namespace {

#ifndef SF_HALF_DEPTH
#define SF_HALF_DEPTH 2
#endif
constexpr int kHalfDepth = SF_HALF_DEPTH; // planes of loads in flight
// Cache policy of the black-cell stores: device scope (kDeviceScopeStore): 28.8 -> 28.0 us per
// middle pass at 256^3 against the default policy; non-temporal stores measured slower.
constexpr int kHalfStoreAux = kDeviceScopeStore;
// Minimum waves per SIMD the register allocation must leave room for. Passes that read the
// natural-layout start iterate hold float2 loads in flight and went to 133 VGPRs (three
// waves per SIMD, one 512-thread workgroup per CU); capped at 128 they keep two workgroups
// per CU without spilling (about 9 us less per first pass at 256^3).
template <int kIn>
constexpr int kHalfMinWaves = kIn == kHalfInNatural ? 4 : 2;
// progress-based wave priority in the plane loop of single-grid passes (0 turns it off)
#ifndef SF_HALF_PRIO
#define SF_HALF_PRIO 1
#endif

// Tile shape: kTXP x-pairs (2 kTXP cells) by TY rows for kWaves waves of kRowsW row
// slots; a slot of a wave covers kR = 64 / kTXP rows of kTXP lanes each.
template <int kTXP, int kWaves, int kRowsW>
struct HalfTile {
    static constexpr int kR       = 64 / kTXP;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int TY     = kWaves * kRowsW * kR; // tile rows
    static constexpr int AW     = kTXP + 4;         // A plane: pairs -2 .. kTXP + 1
    static constexpr int AR     = TY + 4;           //          rows -2 .. TY + 1
    static constexpr int AF     = AW * AR;
    static constexpr int RW     = kTXP + 2;         // R plane: pairs -1 .. kTXP
    static constexpr int RR     = TY + 2;           //          rows -1 .. TY
    static constexpr int RF     = RW * RR;
    // ring plane strides: each plane ends in 64 words that take the writes of halo-slot
    // lanes with no position (one word per lane), so the write offset is loop-invariant
    static constexpr int AS     = AF + 64;
    static constexpr int RS     = RF + 64;
    static constexpr int kAHalo = 4 * AW + 4 * TY;  // A halo positions (one extra slot per lane)
    static constexpr int kRHalo = 2 * RW + 2 * TY;  // R halo positions
    static constexpr int kRHaloWaves = (kRHalo + 63) / 64;
    static constexpr int kAPer  = (kAHalo + kThreads - 1) / kThreads; // A halo slots per lane
    static_assert(kRHalo <= kThreads && 2 * kTXP + 2 * TY <= kThreads, "one R halo and one ghost-store slot per lane");
    static_assert(kTXP * kR == 64 && (kWaves % 2 == 0 || kR % 2 == 0), "own rows of a lane share one parity");
};

// Position classes of a (pair, row) of the neighbourhood.
enum HalfKind : int {
    kPosReal = 0, // a grid cell (wrapped on periodic axes)
    kPosXm   = 1, // x- ghost (pair -1), non-periodic x
    kPosXp   = 2, // x+ ghost (pair nxh)
    kPosYm   = 3, // y- ghost (row -1), non-periodic y
    kPosYp   = 4, // y+ ghost (row ny)
    kPosNone = 5, // further out, or outside on both axes: read by no real cell
};

struct HalfPos {
    int kind;
    int off_h; // float offset in a padded half plane (real: the black cell; ghost: its entry)
    int off_n; // float offset of the aligned natural pair holding the cell (ghost: the adjacent pair)
};

__device__ __forceinline__ HalfPos classify(const int i, const int y, const Dims d, const HalfLayout& L, const bool px, const bool py) {
    const int nxh = L.nxh;
    const bool xo = !px && (i < 0 || i >= nxh);
    const bool yo = !py && (y < 0 || y >= d.ny);
    const int iw  = px ? wrap_index(i, nxh) : min(max(i, 0), nxh - 1);
    const int yw  = py ? wrap_index(y, d.ny) : min(max(y, 0), d.ny - 1);
    HalfPos p;
    p.kind  = kPosReal;
    p.off_n = yw * d.nx + 2 * iw;
    p.off_h = (yw + 1) * nxh + iw;
    if (xo && yo) {
        p.kind = kPosNone;
    } else if (xo) {
        p.kind  = i == -1 ? kPosXm : (i == nxh ? kPosXp : kPosNone);
        p.off_h = (d.ny + 2) * nxh + (i < 0 ? 0 : d.ny) + yw;
    } else if (yo) {
        p.kind  = y == -1 ? kPosYm : (y == d.ny ? kPosYp : kPosNone);
        p.off_h = (y < 0 ? 0 : d.ny + 1) * nxh + iw;
    }
    return p;
}

// c ? a : b on values (a conditional on two lvalues selects their addresses, which keeps
// the selected variables in memory)
template <class T>
__device__ __forceinline__ T pick(const bool c, const T a, const T b) { return c ? a : b; }

// LDS of one workgroup: the A and R rings and the ghost-rule table (declared by the kernel,
// so both march directions share one allocation)
template <class TT>
struct HalfLds {
    float a[4 * TT::AS];
    float r[2 * TT::RS];
    // ghost rules (a, c, k) of faces 0..5 and the identity at 6, read with runtime indices
    // at set-up (indexing the kernel argument itself would copy it to scratch memory)
    float rule_a[8], rule_c[8];
    int rule_k[8];
};

template <int kRowsW, int kAPer, int kIn>
struct PfSet {
    using T = std::conditional_t<kIn == kHalfInNatural, float2, float>;
    T own[kRowsW];
    T ex[kAPer];
};

// kDown: the chunk is marched from its top plane down (odd chunks), so two neighbouring
// chunks reach the planes around their shared boundary at the same time and the second
// read of each halo plane hits the XCD's L2. A red-black half-sweep has no order between
// planes, so either direction gives the same bits; sums keep the order z-, z+ of the
// grid, whatever the direction. Ring positions follow the march coordinate u; memory,
// ghosts and parities the grid plane pl(u). kPrio: progress-based wave priority (below).
template <int kSolve, int kIn, int kOut, int kTXP, int kWaves, int kRowsW, int kDepth, bool kDown, bool kPrio>
__device__ __forceinline__ void rbgs_half_body(const HalfPassArgs& args, const int w_block, HalfLds<HalfTile<kTXP, kWaves, kRowsW>>& lds) {
    using TT           = HalfTile<kTXP, kWaves, kRowsW>;
    constexpr int kR   = TT::kR;
    constexpr int kAPer = TT::kAPer;
    constexpr int TY   = TT::TY;
    constexpr int AW   = TT::AW;
    constexpr int RW   = TT::RW;
    constexpr int AF   = TT::AF;
    constexpr int RF   = TT::RF;
    constexpr int AS   = TT::AS;
    constexpr int RS   = TT::RS;
    float* const s_a      = lds.a;
    float* const s_r      = lds.r;
    float* const s_rule_a = lds.rule_a;
    float* const s_rule_c = lds.rule_c;
    int* const s_rule_k   = lds.rule_k;

    const Dims d       = args.d;
    const StencilBC& bc = args.bc;
    // (the rule table is filled after the first plane loads are issued, so their latency
    // overlaps the set-up; rules are looked up after its barrier)
    auto rule = [&](const int f) { return LinGhost{s_rule_a[f], s_rule_c[f], s_rule_k[f] != 0}; };
    auto kind_rule = [&](const int kind) { return rule(kind >= kPosXm && kind <= kPosYp ? kind - 1 : 6); };
    const HalfLayout L = args.L;
    const float coef   = args.coef;
    int tx, ty, cz;
    tile_of(args.g, w_block, tx, ty, cz);
    const int tid  = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l    = lane % kTXP; // the lane's pair in its rows, and its row within a slot
    const int sub  = lane / kTXP;
    // tile row of own slot j
    auto orow = [&](const int j) { return (wave + kWaves * j) * kR + sub; };
    const int i0 = tx * kTXP, y0 = ty * TY;
    const int z0 = cz * args.g.zc;
    const int z1 = min(z0 + args.g.zc, d.nz);
    const int nx = d.nx, ny = d.ny, nz = d.nz, nxh = L.nxh;
    const bool px = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
    const uint32_t nplane_b = 4u * static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny); // natural plane bytes
    const uint32_t nfield_b = nplane_b * static_cast<uint32_t>(nz);
    const uint32_t hplane_b = 4u * static_cast<uint32_t>(L.ps);
    const uint32_t hfield_b = hplane_b * static_cast<uint32_t>(nz + 2);

    // ---- own slots: rows wave + 8 j of the tile, pair i0 + lane ----
    int o_src[kRowsW], o_b[kRowsW], o_st[kRowsW], o_kind[kRowsW];
    bool o_xflip[kRowsW];
    LinGhost o_rule[kRowsW];
    bool tile_ghost = false; // some own slot is not a real cell (partial tile)
#pragma unroll
    for (int j = 0; j < kRowsW; ++j) {
        const int y       = y0 + orow(j);
        const int i       = i0 + l;
        const HalfPos P   = classify(i, y, d, L, px, py);
        o_kind[j]         = P.kind;
        o_xflip[j]        = P.kind == kPosXm || P.kind == kPosXp;
        o_src[j]          = kIn == kHalfInNatural ? 4 * P.off_n : 4 * P.off_h;
        o_b[j]            = P.kind == kPosReal ? 4 * P.off_n : kOutOfRange;
        const bool owned  = i < nxh && y < ny;
        o_st[j]           = !owned ? kOutOfRange : (kOut == kHalfOutHalf ? 4 * ((y + 1) * nxh + i) : 4 * (y * nx + 2 * i));
        tile_ghost        = tile_ghost || P.kind != kPosReal;
    }
    tile_ghost = __builtin_amdgcn_readfirstlane(__any(tile_ghost)) != 0;
    const int ypar = (y0 + orow(0)) & 1; // row parity of every own slot of the lane

    // ---- extra A slot: the halo of the A planes ----
    int ex_src[kAPer], ex_lds[kAPer], ex_y[kAPer];
    bool ex_xflip[kAPer];
    int ex_kind[kAPer];
    LinGhost ex_rule[kAPer];
#pragma unroll
    for (int q = 0; q < kAPer; ++q) {
        const int e = tid + q * TT::kThreads;
        ex_src[q]   = kOutOfRange;
        ex_lds[q]   = AF + lane; // no position: the plane's sink word
        ex_y[q]     = 0;
        ex_xflip[q] = false;
        ex_kind[q]  = kPosReal; // identity rule
        if (e < TT::kAHalo) {
            int r, hl;
            if (e < 4 * AW) {
                const int rr = e / AW;
                r            = rr < 2 ? rr - 2 : TY + rr - 2;
                hl           = e % AW - 2;
            } else {
                const int e2 = e - 4 * AW, c = e2 & 3;
                r            = e2 >> 2;
                hl           = c < 2 ? c - 2 : kTXP + c - 2;
            }
            const HalfPos P = classify(i0 + hl, y0 + r, d, L, px, py);
            ex_src[q]       = kIn == kHalfInNatural ? 4 * P.off_n : 4 * P.off_h;
            ex_lds[q]       = (r + 2) * AW + hl + 2;
            ex_y[q]         = y0 + r;
            ex_xflip[q]     = P.kind == kPosXm || P.kind == kPosXp;
            ex_kind[q]      = P.kind;
        }
    }

    // ---- R halo slot (waves 0 .. kRHaloWaves - 1): the halo of the R planes ----
    // A ghost position takes the face rule of the adjacent cell's old value; the adjacent
    // cell is black where it matters (x-: p = 1, x+: p = 0), so its A entry is at a fixed
    // offset h_goff from the position (x- +1, x+ -1, y- +AW, y+ -AW).
    int h_b = kOutOfRange, h_alds = 0, h_rlds = RF + lane, h_y = 0, h_goff = 0; // no position: the sink word
    bool h_real = false;
    int h_kind  = kPosReal; // identity rule
    LinGhost h_rule;
    if (tid < TT::kRHalo) {
        int r, hl;
        if (tid < 2 * RW) {
            r = tid < RW ? -1 : TY;
            hl = tid % RW - 1;
        } else {
            const int e2 = tid - 2 * RW;
            r            = e2 >> 1;
            hl           = (e2 & 1) ? kTXP : -1;
        }
        const HalfPos P = classify(i0 + hl, y0 + r, d, L, px, py);
        h_real          = P.kind == kPosReal;
        h_b             = h_real ? 4 * P.off_n : kOutOfRange;
        h_alds          = (r + 2) * AW + hl + 2;
        h_rlds          = (r + 1) * RW + hl + 1;
        h_y             = y0 + r;
        h_kind          = P.kind;
        h_goff          = P.kind == kPosXm ? 1 : (P.kind == kPosXp ? -1 : (P.kind == kPosYm ? AW : (P.kind == kPosYp ? -AW : 0)));
    }

    // ---- ghost-store slot (half output): red face cells of the tile -> padded entries ----
    //   threads 0 .. kTXP - 1: row 0 -> row -1; the next kTXP: row ny - 1 -> row ny;
    //   the next TY: pair 0 -> x- column; the next TY: pair nxh - 1 -> x+ column
    int g_off = kOutOfRange, g_rlds = 0;
    int g_face = 6; // identity
    LinGhost g_rule;
    if constexpr (kOut == kHalfOutHalf) {
        if (tid < 2 * kTXP) {
            const int side = tid / kTXP, gl = tid % kTXP;
            const int y    = side ? ny - 1 : 0;
            if (!py && y >= y0 && y < y0 + TY && i0 + gl < nxh) {
                g_off  = 4 * ((side ? ny + 1 : 0) * nxh + i0 + gl);
                g_rlds = (y - y0 + 1) * RW + gl + 1;
                g_face = 2 + side;
            }
        } else if (tid < 2 * kTXP + 2 * TY) {
            const int side = (tid - 2 * kTXP) / TY, r = (tid - 2 * kTXP) % TY;
            const int i    = side ? nxh - 1 : 0;
            if (!px && i >= i0 && i < i0 + kTXP && y0 + r < ny) {
                g_off  = 4 * ((ny + 2) * nxh + side * ny + y0 + r);
                g_rlds = (r + 1) * RW + i - i0 + 1;
                g_face = side;
            }
        }
    }
    const bool g_any = __builtin_amdgcn_readfirstlane(__any(g_off != kOutOfRange)) != 0; // wave-uniform
    LinGhost zr_lo, zr_hi; // z-face rules (set after the rule table's barrier)
    const DiffusionDiv ddv = diffusion_div(coef); // loop-invariant reciprocal of 1 + 6 coef

    auto zghost = [&](const int k) { return !pz && (k < 0 || k >= nz); };
    const int zmir = z0 + z1 - 1;
    auto pl        = [&](const int u) { return kDown ? zmir - u : u; }; // grid plane of march plane u
    // periodic plane of k in [-2, nz + 1] (the planes a chunk reads; nz >= 2)
    auto wrap_near_z = [&](const int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
    // buffer resources: uniform bases once, the size selected per plane (0 bytes: off)
    const void* p_hin  = uniform_ptr(args.h_in);
    const void* p_xnat = uniform_ptr(args.x_nat);
    const void* p_b    = uniform_ptr(args.b);
    const void* p_hout = uniform_ptr(args.h_out);
    const void* p_xout = uniform_ptr(args.x_out);

    // ---- loads (kLive: the plane is known to be needed and real) ----
    using Pf = PfSet<kRowsW, kAPer, kIn>;
    auto load_a = [&]<bool kLive>(const int u, Pf& pf) __attribute__((always_inline)) {
        const int k = pl(u);
        if constexpr (kIn == kHalfInHalf) {
            const int kk    = pz ? wrap_near_z(k) : k;
            const bool live = kLive || (k >= z0 - 2 && k <= z1 + 1 && kk >= -1 && kk <= nz);
            const Rsrc r    = rsrc_at(p_hin, live ? hfield_b : 0u);
            const int so    = live ? (kk + 1) * static_cast<int>(hplane_b) : 0;
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) pf.own[j] = buf_load1<StoreF32>(r, o_src[j], so);
#pragma unroll
            for (int q = 0; q < kAPer; ++q) pf.ex[q] = buf_load1<StoreF32>(r, ex_src[q], so);
        } else if constexpr (kIn == kHalfInNatural) {
            const int kc    = pz ? wrap_near_z(k) : min(max(k, 0), nz - 1);
            const bool live = kLive || (k >= z0 - 2 && k <= z1 + 1 && (pz || (k >= -1 && k <= nz)));
            const Rsrc r    = rsrc_at(p_xnat, live ? nfield_b : 0u);
            const int so    = live ? kc * static_cast<int>(nplane_b) : 0;
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) pf.own[j] = buf_load2<StoreF32>(r, o_src[j], so);
#pragma unroll
            for (int q = 0; q < kAPer; ++q) pf.ex[q] = buf_load2<StoreF32>(r, ex_src[q], so);
        }
    };
    struct PfB {
        float2 own[kRowsW];
        float2 h;
    };
    auto load_b = [&]<bool kLive>(const int u, PfB& pf) __attribute__((always_inline)) {
        const int k     = pl(u);
        const int kc    = pz ? wrap_near_z(k) : (kLive ? k : min(max(k, 0), nz - 1));
        const bool live = kLive || (k >= z0 - 1 && k <= z1 && (pz || (k >= 0 && k < nz)));
        const Rsrc r    = rsrc_at(p_b, live ? nfield_b : 0u);
        const int so    = live ? kc * static_cast<int>(nplane_b) : 0;
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) pf.own[j] = buf_load2<StoreF32>(r, o_b[j], so);
        pf.h = buf_load2<StoreF32>(r, h_b, so);
    };

    // value of an A position at plane k from its raw load (kReal: k is not a z ghost plane)
    auto a_value = [&]<bool kReal>(const auto raw, const int k, const int y, const bool xflip, const LinGhost rl) __attribute__((always_inline)) -> float {
        if constexpr (kIn == kHalfInHalf) {
            return raw; // ghost entries hold the rule already applied
        } else {
            const LinGhost gr = kReal ? rl : pick(zghost(k), pick(k < 0, zr_lo, zr_hi), rl);
            float v           = 0.0f;
            if constexpr (kIn == kHalfInNatural) {
                // black element of the pair by the position's parity; on y / z ghosts this is
                // the red cell of the adjacent pair, on x ghosts the red cell is the other one
                const int e = ((y + k) & 1) ^ 1 ^ (xflip ? 1 : 0);
                v           = e ? raw.y : raw.x;
            }
            return lin_apply(gr, v);
        }
    };

    // ---- register state (march planes) ----
    float a_prev[kRowsW], a_cur[kRowsW], a_next[kRowsW]; // own A(s-1), A(s), A(s+1)
    float r_mm[kRowsW], r_m[kRowsW], r_0[kRowsW];        // own R(s-2), R(s-1), R(s)
    float bb[kRowsW], bb_new[kRowsW];                    // b of the own black cells, planes s-1 and s
#pragma unroll
    for (int j = 0; j < kRowsW; ++j) {
        a_prev[j] = a_cur[j] = a_next[j] = 0.0f;
        r_mm[j] = r_m[j] = r_0[j] = 0.0f;
        bb[j] = bb_new[j] = 0.0f;
    }

    // stage A(u) from `pf` into the ring plane at `A`; own values into `own`
    auto stage_a = [&]<bool kReal>(const int u, float* A, const Pf& pf, float (&own)[kRowsW]) __attribute__((always_inline)) {
        const int k = pl(u);
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            float v;
            if constexpr (kIn == kHalfInZero) v = a_value.template operator()<kReal>(0.0f, k, 0, o_xflip[j], o_rule[j]);
            else v = a_value.template operator()<kReal>(pf.own[j], k, y0 + orow(j), o_xflip[j], o_rule[j]);
            own[j] = v;
            A[(orow(j) + 2) * AW + l + 2] = v;
        }
#pragma unroll
        for (int q = 0; q < kAPer; ++q) {
            float v;
            if constexpr (kIn == kHalfInZero) v = a_value.template operator()<kReal>(0.0f, k, 0, ex_xflip[q], ex_rule[q]);
            else v = a_value.template operator()<kReal>(pf.ex[q], k, ex_y[q], ex_xflip[q], ex_rule[q]);
            A[ex_lds[q]] = v;
        }
    };

    // red half-update of plane s (own slots and the R halo) from the A planes s-1, s, s+1
    // into R0 (kReal: s is not a z ghost plane)
    auto red = [&]<bool kReal>(const int u, const float* Am, const float* A0, const float* Ap, float* R0, const PfB& pb) __attribute__((always_inline)) {
        const int s   = pl(u);
        const bool zg = !kReal && zghost(s);
        const int p   = (ypar + s) & 1;
        // A planes below and above in the grid (Am / Ap are the march neighbours)
        const float* Ab   = kDown ? Ap : Am;
        const float* Aa   = kDown ? Am : Ap;
        auto a_bl = [&](const int j) { if constexpr (kDown) return a_next[j]; else return a_prev[j]; };
        auto a_ab = [&](const int j) { if constexpr (kDown) return a_prev[j]; else return a_next[j]; };
        float q[kRowsW], qh = 0.0f;
        if (zg) {
            // ghost plane: the z-face rule of the adjacent (black) old value
            const LinGhost zr = pick(s < 0, zr_lo, zr_hi);
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) q[j] = lin_apply(zr, pick(s < 0, a_ab(j), a_bl(j)));
            if (wave < TT::kRHaloWaves) qh = lin_apply(zr, pick(s < 0, Aa[h_alds], Ab[h_alds]));
        } else {
            StageDiv<kSolve> dv(ddv);
            float x[kRowsW], xm_[kRowsW], ym_[kRowsW], xh = 0.0f;
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) {
                const int c       = (orow(j) + 2) * AW + l + 2;
                const float other = A0[c - 1 + 2 * p];
                const float xm    = p ? a_cur[j] : other;
                const float xp    = p ? other : a_cur[j];
                const float ym    = A0[c - AW];
                float sum         = xm;
                sum               = sum + xp;
                sum               = sum + ym;
                sum               = sum + A0[c + AW];
                sum               = sum + a_bl(j);
                sum               = sum + a_ab(j);
                const float2 bp   = pb.own[j];
                x[j]              = dv.num(p ? bp.y : bp.x, sum, coef);
                bb_new[j]         = p ? bp.x : bp.y;
                xm_[j]            = xm;
                ym_[j]            = ym;
            }
            if (tile_ghost) {
#pragma unroll
                for (int j = 0; j < kRowsW; ++j)
                    if (o_kind[j] != kPosReal) x[j] = 0.0f;
            }
            float hg = 0.0f;
            if (wave < TT::kRHaloWaves) {
                const int c    = h_alds;
                const int ph   = (h_y + s) & 1;
                float sum      = A0[c - 1 + ph];
                sum            = sum + A0[c + ph];
                sum            = sum + A0[c - AW];
                sum            = sum + A0[c + AW];
                sum            = sum + Ab[c];
                sum            = sum + Aa[c];
                const float xv = dv.num(ph ? pb.h.y : pb.h.x, sum, coef);
                xh             = h_real ? xv : 0.0f;
                hg             = lin_apply(h_rule, A0[c + h_goff]);
            }
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) q[j] = dv.quot(x[j]);
            qh = dv.quot(xh);
            if (__builtin_expect(dv.bad(), 0)) {
#pragma unroll
                for (int j = 0; j < kRowsW; ++j) q[j] = dv.exact(x[j]);
                qh = dv.exact(xh);
            }
            qh = h_real ? qh : hg;
            if (tile_ghost) {
                // past the x+ face the ghost is the rule of the adjacent cell (black when
                // it matters: p = 0), past the y+ face the rule of the cell below
#pragma unroll
                for (int j = 0; j < kRowsW; ++j) {
                    const int kd = o_kind[j];
                    if (kd != kPosReal) q[j] = kd == kPosXp ? lin_apply(o_rule[j], xm_[j]) : (kd == kPosYp ? lin_apply(o_rule[j], ym_[j]) : 0.0f);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            r_0[j]                                       = q[j];
            R0[(orow(j) + 1) * RW + l + 1] = q[j];
        }
        if (wave < TT::kRHaloWaves) R0[h_rlds] = qh;
    };

    // LDS neighbours of the black cells of plane m = s - 1 in R1 = R(m): the other x
    // neighbour and the rows above and below; g: the red face cell of the lane's
    // ghost-store slot (waves that hold one)
    struct BlackIn {
        float other[kRowsW], ym[kRowsW], yp[kRowsW];
        float g;
    };
    auto black_fetch = [&](const int u, const float* R1) __attribute__((always_inline)) {
        const int p = (ypar + pl(u - 1)) & 1;
        BlackIn bi;
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            const int c    = (orow(j) + 1) * RW + l + 1;
            bi.other[j]    = R1[c + 1 - 2 * p];
            bi.ym[j]       = R1[c - RW];
            bi.yp[j]       = R1[c + RW];
        }
        bi.g = 0.0f;
        if constexpr (kOut == kHalfOutHalf)
            if (g_any) bi.g = R1[g_rlds];
        return bi;
    };

    // black half-update of plane m = s - 1 from the LDS neighbours `bi`, the own red
    // values of planes m - 1 .. m + 1, and the stores of plane m (kLive: m is a plane of
    // the chunk and not a z face)
    auto black = [&]<bool kLive>(const int u, const BlackIn& bi) __attribute__((always_inline)) {
        const int m     = pl(u - 1);
        const bool live = kLive || (m >= z0 && m < z1); // planes of the chunk (never a ghost plane)
        const int p     = (ypar + m) & 1;
        // own red values of the grid planes m - 1 and m + 1
        auto r_bl = [&](const int j) { if constexpr (kDown) return r_0[j]; else return r_mm[j]; };
        auto r_ab = [&](const int j) { if constexpr (kDown) return r_mm[j]; else return r_0[j]; };
        StageDiv<kSolve> dv(ddv);
        float x[kRowsW], q[kRowsW];
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            const float xm = p ? bi.other[j] : r_m[j];
            const float xp = p ? r_m[j] : bi.other[j];
            float sum      = xm;
            sum            = sum + xp;
            sum            = sum + bi.ym[j];
            sum            = sum + bi.yp[j];
            sum            = sum + r_bl(j);
            sum            = sum + r_ab(j);
            x[j]           = dv.num(bb[j], sum, coef);
        }
        if (!live || tile_ghost) {
#pragma unroll
            for (int j = 0; j < kRowsW; ++j)
                if (!live || o_kind[j] != kPosReal) x[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) q[j] = dv.quot(x[j]);
        if (__builtin_expect(dv.bad(), 0)) {
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) q[j] = dv.exact(x[j]);
        }
        if constexpr (kOut == kHalfOutHalf) {
            const Rsrc r = rsrc_at(p_hout, live ? hfield_b : 0u);
            const int so = live ? (m + 1) * static_cast<int>(hplane_b) : 0;
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(q[j]), r, o_st[j], so, kHalfStoreAux);
            // red face cells of plane m -> ghost entries: rows -1 / ny and the x columns
            // (one slot per lane, waves that hold face cells), the z planes -1 / nz (own
            // slots). Only stores sit under these branches: a skipped store leaves every
            // later wait count valid.
            if (g_any && live) {
                const float g = lin_apply(g_rule, bi.g);
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(g), r, g_off, so, kHalfStoreAux);
            }
            if (!kLive && live && !pz && (m == 0 || m == nz - 1)) {
                const int zso     = m == 0 ? 0 : (nz + 1) * static_cast<int>(hplane_b);
                const LinGhost zr = pick(m == 0, zr_lo, zr_hi);
#pragma unroll
                for (int j = 0; j < kRowsW; ++j) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(lin_apply(zr, r_m[j])), r, o_st[j], zso, kHalfStoreAux);
            }
        } else {
            const Rsrc r  = rsrc_at(p_xout, live ? nfield_b : 0u);
            const int so  = live ? m * static_cast<int>(nplane_b) : 0;
            const float f = args.out_scale;
#pragma unroll
            for (int j = 0; j < kRowsW; ++j) {
                // p = 0: the red cell is the even one
                const float ve = (p ? q[j] : r_m[j]) * f;
                const float vo = (p ? r_m[j] : q[j]) * f;
                buf_store2<StoreF32, kHalfStoreAux>(r, o_st[j], ve, vo, so);
            }
        }
    };

    // One plane step s, t = s - (z0 - 1): stage A(s+1) (fetched kDepth steps ago into pa)
    // and refetch pa with A(s+1+kDepth); barrier; red of plane s with b(s) (fetched kDepth
    // steps ago into pb), refetch pb with b(s+kDepth); black of plane s - 1. Ring slots:
    // A(s+1) (t+2)&3, A(s) (t+1)&3, A(s-1) t&3, R(s) (t+1)&1, R(s-1) t&1. kSteady: s in
    // [z0 + 2, z1 - 3], where every load reads a plane the chunk needs (at most plane z1
    // of A, z1 - 1 of b, so padded plane nz at most), no plane is a z ghost and the black
    // plane s - 1 is neither z face; the slots are then compile-time (t mod 4 = kT).
    auto step = [&]<bool kSteady, int kT>(const int s, const int t, Pf& pa, PfB& pb) __attribute__((always_inline)) {
        const int tt  = kSteady ? kT : t;
        float* A_next = s_a + ((tt + 2) & 3) * AS;
        float* A_cur  = s_a + ((tt + 1) & 3) * AS;
        float* A_prev = s_a + (tt & 3) * AS;
        float* R_cur  = s_r + ((tt + 1) & 1) * RS;
        float* R_prev = s_r + (tt & 1) * RS;
        stage_a.template operator()<kSteady>(s + 1, A_next, pa, a_next);
        load_a.template operator()<kSteady>(s + 1 + kDepth, pa);
        __syncthreads();
        // the black half-update's LDS reads (R(s-1), written before this barrier) go out
        // first, so their latency overlaps the red half-update
        const BlackIn bi = black_fetch(s, R_prev);
        red.template operator()<kSteady>(s, A_prev, A_cur, A_next, R_cur, pb);
        load_b.template operator()<kSteady>(s + kDepth, pb);
        black.template operator()<kSteady>(s, bi);
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) {
            a_prev[j] = a_cur[j];
            a_cur[j]  = a_next[j];
            r_mm[j]   = r_m[j];
            r_m[j]    = r_0[j];
            bb[j]     = bb_new[j];
        }
    };
    static_assert(kDepth == 2, "register sets alternate with the step parity");

    Pf pa[2];
    PfB pb[2];
    // prologue: every load of A(z0-2) .. A(z0+1), b(z0-1), b(z0) goes out first; then the
    // rule table (one barrier) and the rules of every slot; A(z0-2), A(z0-1) are staged
    // into slots 0, 1 while A(z0), A(z0+1) and b(z0-1), b(z0) stay in flight
    {
        Pf p0, p1;
        load_a.template operator()<false>(z0 - 2, p0);
        load_a.template operator()<false>(z0 - 1, p1);
        load_a.template operator()<false>(z0, pa[0]);
        load_a.template operator()<false>(z0 + 1, pa[1]);
        load_b.template operator()<false>(z0 - 1, pb[0]);
        load_b.template operator()<false>(z0, pb[1]);
        if (threadIdx.x < 7) {
            const int f       = static_cast<int>(threadIdx.x);
            const LinGhost lg = f == 6 ? lin_identity() : lin_ghost(face_rule(bc, f));
            s_rule_a[f]       = lg.a;
            s_rule_c[f]       = lg.c;
            s_rule_k[f]       = lg.k ? 1 : 0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kRowsW; ++j) o_rule[j] = kind_rule(o_kind[j]);
#pragma unroll
        for (int q = 0; q < kAPer; ++q) ex_rule[q] = kind_rule(ex_kind[q]);
        h_rule = kind_rule(h_kind);
        g_rule = rule(g_face);
        zr_lo  = rule(4);
        zr_hi  = rule(5);
        stage_a.template operator()<false>(z0 - 2, s_a, p0, a_cur);
        stage_a.template operator()<false>(z0 - 1, s_a + AS, p1, a_next);
    }
#pragma unroll
    for (int j = 0; j < kRowsW; ++j) {
        a_prev[j] = a_cur[j];
        a_cur[j]  = a_next[j];
    }
    // steps t = 0 .. n - 1 (s = z0 - 1 .. z1), register set t & 1: general steps t = 0, 1, 2;
    // steady steps in groups of four, t = 3 .. 2 + 4 g, s = z0 + 2 .. z0 + 1 + 4 g <= z1 - 3;
    // general steps for the rest (an odd count ends with one idle step, loads and stores off)
    const int n      = z1 - z0 + 2;
    const int groups = n >= 10 ? (n - 6) / 4 : 0;
    // Progress-based wave priority: a CU runs two workgroups, and oldest-first issue lets
    // the one dispatched first run ahead and finish while the other still has a quarter of
    // its planes left, which it then runs alone. Priority falls as a wave advances, so the
    // workgroup that is ahead yields issue slots to the one behind (27.4 -> 27.15 us per
    // middle pass at 256^3). Batched passes run without it: there it cost 3% of the
    // 64-plume batch's throughput.
    auto set_prio = [](const int q) __attribute__((always_inline)) {
        if constexpr (kPrio) {
            if (q <= 0) __builtin_amdgcn_s_setprio(3);
            else if (q == 1) __builtin_amdgcn_s_setprio(2);
            else if (q == 2) __builtin_amdgcn_s_setprio(1);
            else __builtin_amdgcn_s_setprio(0);
        }
    };
    set_prio(0);
    const int pg1 = groups >> 2, pg2 = (2 * groups) >> 2, pg3 = (3 * groups) >> 2;
    step.template operator()<false, 0>(z0 - 1, 0, pa[0], pb[0]);
    step.template operator()<false, 0>(z0, 1, pa[1], pb[1]);
    if (n > 2) step.template operator()<false, 0>(z0 + 1, 2, pa[0], pb[0]);
    for (int g = 0; g < groups; ++g) {
        // level floor((4 g + 3) / groups) rises at g = floor(k groups / 4), k = 1, 2, 3
        if constexpr (kPrio)
            if (g == pg1 || g == pg2 || g == pg3) set_prio((g >= pg1) + (g >= pg2) + (g >= pg3));
        const int s = z0 + 2 + 4 * g; // t = 3 + 4 g
        step.template operator()<true, 3>(s, 0, pa[1], pb[1]);
        step.template operator()<true, 0>(s + 1, 0, pa[0], pb[0]);
        step.template operator()<true, 1>(s + 2, 0, pa[1], pb[1]);
        step.template operator()<true, 2>(s + 3, 0, pa[0], pb[0]);
    }
    set_prio(3);
    for (int t = 3 + 4 * groups; t < n; t += 2) {
        step.template operator()<false, 0>(z0 - 1 + t, t, pa[1], pb[1]);
        step.template operator()<false, 0>(z0 + t, t + 1, pa[0], pb[0]);
    }
}

// one pass over the grid args.g
template <int kSolve, int kIn, int kOut, int kTXP, int kWaves, int kRowsW>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(kHalfMinWaves<kIn>, 4))) void k_rbgs_half(const HalfPassArgs args) {
    __shared__ HalfLds<HalfTile<kTXP, kWaves, kRowsW>> lds;
    const int w = xcd_block();
    constexpr bool kPrio = SF_HALF_PRIO && kWaves == 8; // measured on the 512-thread tiles
    if ((w % args.g.chunks) & 1) rbgs_half_body<kSolve, kIn, kOut, kTXP, kWaves, kRowsW, kHalfDepth, true, kPrio>(args, w, lds);
    else rbgs_half_body<kSolve, kIn, kOut, kTXP, kWaves, kRowsW, kHalfDepth, false, kPrio>(args, w, lds);
}

// batched: workgroup i runs block items[i].block of the pass descs[items[i].desc] (the
// descriptor is read in place); an idle item does nothing
template <int kSolve, int kIn, int kOut, int kTXP, int kWaves, int kRowsW>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(kHalfMinWaves<kIn>, 4))) void k_rbgs_half_batched(const HalfPassArgs* __restrict__ descs, const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    if (it.desc == kIdleItem) return; // the whole workgroup: no barrier is left waiting
    __shared__ HalfLds<HalfTile<kTXP, kWaves, kRowsW>> lds;
    const HalfPassArgs& a = descs[it.desc];
    const int w           = static_cast<int>(it.block);
    if ((w % a.g.chunks) & 1) rbgs_half_body<kSolve, kIn, kOut, kTXP, kWaves, kRowsW, kHalfDepth, true, false>(a, w, lds);
    else rbgs_half_body<kSolve, kIn, kOut, kTXP, kWaves, kRowsW, kHalfDepth, false, false>(a, w, lds);
}

// Calls f.template operator()<solve, in, out, pairs, waves, row slots>() for the runtime choice.
template <class F>
void half_dispatch(const SolveKind solve, const int in_mode, const int out_mode, const int tile, F&& f) {
    auto shape = [&]<int SO, int IN, int OUT>() {
        switch (tile) {
        case kHalfTile64x32: f.template operator()<SO, IN, OUT, 32, 8, 2>(); break;
        case kHalfTile64x16: f.template operator()<SO, IN, OUT, 32, 4, 2>(); break;
        case kHalfTile32x32: f.template operator()<SO, IN, OUT, 16, 4, 2>(); break;
        case kHalfTile32x16: f.template operator()<SO, IN, OUT, 16, 2, 2>(); break;
        default: f.template operator()<SO, IN, OUT, 64, 8, 2>(); break;
        }
    };
    auto out = [&]<int SO, int IN>() {
        if (out_mode == kHalfOutHalf) shape.template operator()<SO, IN, kHalfOutHalf>();
        else shape.template operator()<SO, IN, kHalfOutNatural>();
    };
    auto in = [&]<int SO>() {
        if (in_mode == kHalfInZero) out.template operator()<SO, kHalfInZero>();
        else if (in_mode == kHalfInNatural) out.template operator()<SO, kHalfInNatural>();
        else out.template operator()<SO, kHalfInHalf>();
    };
    if (solve == kSolvePoisson) in.template operator()<kSolvePoisson>();
    else in.template operator()<kSolveDiffusion>();
}

} // namespace

} // namespace sf
