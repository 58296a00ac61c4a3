// sf_rbgs_half.hip -- host side of the black-cell RBGS passes (sf_rbgs_half.hpp): layout,
// tile plans and the single-context launches. The batched launches live in
// sf_rbgs_half_batch.hip (built in parallel: each instantiates the kernel set once).
#include "sf_rbgs_half.hpp"

namespace sf {

// ---- host side ----

bool rbgs_half_supported(const Dims d, const StencilBC& bc) {
    if ((d.nx & 1) || d.nx < 4 || d.ny < 2 || d.nz < 2) return false;
    if (((bc.periodic & 2) && (d.ny & 1)) || ((bc.periodic & 4) && (d.nz & 1))) return false;
    const HalfLayout L = half_layout(d);
    const uint64_t nat = 4ull * static_cast<uint64_t>(d.nx) * static_cast<uint64_t>(d.ny) * static_cast<uint64_t>(d.nz);
    const uint64_t hb  = 4ull * static_cast<uint64_t>(L.ps) * static_cast<uint64_t>(d.nz + 2);
    return nat < (1ull << 31) && hb < (1ull << 31);
}

HalfLayout half_layout(const Dims d) {
    HalfLayout L{};
    L.nxh        = d.nx / 2;
    const long f = static_cast<long>(d.ny + 2) * L.nxh + 2L * d.ny;
    L.ps         = static_cast<int>((f + 63) / 64 * 64);
    return L;
}

uint64_t half_buffer_floats(const Dims d) {
    const HalfLayout L = half_layout(d);
    return static_cast<uint64_t>(L.ps) * static_cast<uint64_t>(d.nz + 2);
}

namespace {
struct HalfShapeInfo {
    int pairs, waves, rows;
};
// pairs per tile row, waves, tile rows (kTXP, kWaves, kWaves * 2 * 64 / kTXP)
constexpr HalfShapeInfo kHalfShapes[kHalfTileCount] = {{64, 8, 16}, {32, 8, 32}, {32, 4, 16}, {16, 4, 32}, {16, 2, 16}};
HalfShapeInfo shape_of(const int tile) { return kHalfShapes[tile >= 0 && tile < kHalfTileCount ? tile : 0]; }
} // namespace

int half_tile_rows(const int tile) { return shape_of(tile).rows; }
int half_tile_threads(const int tile) { return 64 * shape_of(tile).waves; }
int half_tile_cells(const int tile) { return 2 * shape_of(tile).pairs; }

TileGeom half_plan(const Dims d, const int tile, const int zc_override) {
    const HalfShapeInfo sh = shape_of(tile);
    TileGeom g{};
    const int nxh    = d.nx / 2;
    g.tiles_x        = (nxh + sh.pairs - 1) / sh.pairs;
    g.tiles_y        = (d.ny + sh.rows - 1) / sh.rows;
    const long tiles = static_cast<long>(g.tiles_x) * g.tiles_y;
    // about 4096 waves (16 per CU: two 512-thread workgroups), even chunk lengths of at
    // least 4 planes (the CU count of MI355X; the adjustments below read the device's)
    const long target = 4096 / sh.waves;
    long zc = (static_cast<long>(d.nz) * tiles + target - 1) / target;
    zc      = std::max<long>(4, zc + (zc & 1));
    const long cus = device_cus();
    // Chunks past 64 planes at two 512-thread workgroups per CU: one workgroup per CU with
    // chunks twice as long is faster still (512^3 middle pass: 64-plane chunks, four per
    // CU, 242 us; 128 planes, two per CU, 224 us; 256 planes, one per CU, 214 us). At
    // 256^3 (16 planes) the two are within 1%. Otherwise chunks stop at 64 planes.
    const bool one_per_cu = zc_override <= 0 && zc > 64 && sh.waves == 8 && tiles <= cus;
    if (zc_override > 0) {
        zc = zc_override;
    } else if (one_per_cu) {
        const long chunks = cus / tiles;
        zc                = (d.nz + chunks - 1) / chunks;
        zc                = zc + (zc & 1);
    } else if (sh.waves == 8 && zc <= 64) {
        // More 512-thread workgroups than CUs but at most 1.5 per CU leaves a minority of
        // CUs running two of them, one after the other in effect, and the pass lasts that
        // long; longer chunks that put at most one on each CU finish sooner (the reference
        // benchmark's 128x160x128 grid: 320 workgroups of 4 planes 2.68 ms per step, 220 of
        // 6 planes 2.43 ms).
        auto blocks = [&](const long z) { return tiles * ((d.nz + z - 1) / z); };
        if (blocks(zc) > cus && 2 * blocks(zc) <= 3 * cus)
            while (blocks(zc) > cus && zc < std::min<long>(64, d.nz)) zc += 2;
    }
    if (zc_override <= 0 && !one_per_cu) zc = std::min<long>(zc, 64);
    if (zc > d.nz) zc = d.nz;
    g.zc     = static_cast<int>(zc);
    g.chunks = (d.nz + g.zc - 1) / g.zc;
    return g;
}

int default_half_tile(const Dims d) {
    // idle x positions (pairs) of the last tile column under 64- and 32-pair tiles: 64
    // and 192 cells wide grids leave half or a quarter of a 128-cell tile idle, which the
    // 64x16 tiles avoid (64^3 plume 2.85 -> 2.14 ms per step, 192^3 6.21 -> 4.83 ms)
    const int nxh  = d.nx / 2;
    const int idle0 = (nxh + 63) / 64 * 64 - nxh, idle2 = (nxh + 31) / 32 * 32 - nxh;
    return idle2 < idle0 ? kHalfTile64x16 : kHalfTile128x16;
}

hipError_t launch_rbgs_half(hipStream_t s, const SolveKind solve, const int in_mode, const int out_mode, const HalfPassArgs& a_in, const int tile) {
    HalfPassArgs a        = a_in;
    a.L                   = half_layout(a.d);
    const unsigned blocks = smooth_blocks(a.g);
    half_dispatch(solve, in_mode, out_mode, tile, [&]<int SO, int IN, int OUT, int TXP, int W, int RW>() { k_rbgs_half<SO, IN, OUT, TXP, W, RW><<<blocks, 64 * W, 0, s>>>(a); });
    return hipGetLastError();
}

} // namespace sf
