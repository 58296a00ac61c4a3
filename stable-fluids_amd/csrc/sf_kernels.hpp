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
// When skip is given and *skip != 0 at run time, the pass writes nothing.
struct SmoothPass {
    const float* x_old;
    const float* b;
    float* x_new;
    float coef;
    float out_scale;
    const int* skip = nullptr;
};

// Element storage of a correction iterate (arithmetic is FP32 in every case).
enum StorageKind : int {
    kStoreF32  = 0,
    kStoreF16  = 1,
    kStoreBF16 = 2,
};

// One smoother pass of a correction solve A e = r with homogeneous boundary rules: reads
// e_old (zeros when nullptr) in `storage` and r in FP32, writes e_new in `storage`.
// The iterate is held scaled by 1/s, s the power of two at least *scale_slot (max|r|).
struct CorrectionPass {
    const void* e_old;
    const float* r;
    void* e_new;
    float coef; // diffusion: dt*diffusivity/h^2; Poisson: -1 (6e - sum = r)
    const float* scale_slot;
};

// ---- batched launches: one launch runs the same stage for many simulations ----

// Workgroup i of a batched launch runs local block `block` of descriptor `desc`.
struct BatchItem {
    uint32_t desc;
    uint32_t block;
};
// A work-list entry that does nothing (pads the black-cell lists to a multiple of 8).
constexpr uint32_t kIdleItem = 0xffffffffu;

// Smoother tiling: tiles_x * tiles_y tiles of 128 x 16 cells, z split in chunks of zc planes.
struct TileGeom {
    int tiles_x, tiles_y, chunks, zc;
};
// zc_override > 0 replaces the default chunk length (results do not depend on it).
TileGeom plan_tiles(Dims d, int zc_override);
uint32_t smooth_blocks(const TileGeom& g);

struct SmoothDesc {
    const float* x_old; // nullptr only with a zero-first launch
    const float* b;
    float* x_new;
    float coef;
    float out_scale;
    Dims d;
    StencilBC bc;
    TileGeom g;
};
// RBGS passes need even extents on periodic axes in a batch (no odd-periodic variant).
bool smooth_batchable(SmootherKind smoother, Dims d, const StencilBC& bc);
// Every item's descriptor must have even nx when `pairs`, odd nx otherwise.
hipError_t launch_smooth_batched(hipStream_t s, SmootherKind smoother, SolveKind solve, bool zero, bool pairs, const SmoothDesc* descs, const BatchItem* items, uint32_t n_items);

// ---- RBGS passes that carry only the black cells between passes (sf_rbgs_half.hip) ----
// A solve of L passes: the first reads the natural start iterate (or zeros), the middle
// ones read and write padded half buffers (the black cell of every x pair plus the ghost
// values of the red face cells), the last writes the natural result scaled by out_scale.
enum HalfIn : int { kHalfInZero = 0, kHalfInNatural = 1, kHalfInHalf = 2 };
enum HalfOut : int { kHalfOutHalf = 0, kHalfOutNatural = 1 };
struct HalfLayout {
    int nxh; // pairs per row
    int ps;  // floats per padded half plane
};
struct HalfPassArgs {
    const float* x_nat; // natural start iterate (kHalfInNatural)
    const float* b;
    const float* h_in;  // padded half buffer (kHalfInHalf)
    float* h_out;       // padded half buffer (kHalfOutHalf)
    float* x_out;       // natural result (kHalfOutNatural)
    float coef;
    float out_scale;
    Dims d;
    StencilBC bc;
    TileGeom g;         // half_plan
    HalfLayout L;       // set by the launcher
};
// even nx >= 4, ny >= 2, nz >= 2, even extents on periodic axes, fields below 2 GiB
bool rbgs_half_supported(Dims d, const StencilBC& bc);
HalfLayout half_layout(Dims d);
uint64_t half_buffer_floats(Dims d);
// tile shapes, x cells by rows (threads): 128 x 16 (512), 64 x 32 (512), 64 x 16 (256),
// 32 x 32 (256), 32 x 16 (128)
enum HalfTileShape : int { kHalfTile128x16 = 0, kHalfTile64x32 = 1, kHalfTile64x16 = 2, kHalfTile32x32 = 3, kHalfTile32x16 = 4, kHalfTileCount = 5 };
int half_tile_rows(int tile);
int half_tile_threads(int tile);
int half_tile_cells(int tile);
TileGeom half_plan(Dims d, int tile, int zc_override);
// Tile shape a context starts with: 128x16 cells, or 64x16 where that leaves fewer x
// positions of the last tile column idle.
int default_half_tile(Dims d);
// Compute units of the current device (read once per device; 256 if the query fails).
int device_cus();
hipError_t launch_rbgs_half(hipStream_t s, SolveKind solve, int in_mode, int out_mode, const HalfPassArgs& args, int tile);
// one launch for many passes: workgroup i runs block items[i].block of descs[items[i].desc]
// (every descriptor's L filled in, its g planned for `tile`); kIdleItem entries return
hipError_t launch_rbgs_half_batched(hipStream_t s, SolveKind solve, int in_mode, int out_mode, int tile, const HalfPassArgs* descs, const BatchItem* items, uint32_t n_items);

// Batched advection runs the buffer-resource kernels (advect_buffers_fit must hold) over
// 16 x 16 blocks (advect_blocks); inv_h = 1 / h and pow2 = advect_pow2(h) set by the host.
struct AdvectVelocityDesc {
    float* t[3];
    const float* v[3];
    float dt;
    Dims d;
    float h;
    StableFluidsFlowBoundaryConfig bc;
    float inv_h;
    int pow2;
};
struct AdvectScalarDesc {
    float* out;
    const float* src;
    const float* v[3];
    float dt;
    Dims d;
    float h;
    StableFluidsScalarBoundaryConfig sbc;
    StableFluidsFlowBoundaryConfig fbc;
    float inv_h;
    int pow2;
};
bool advect_pow2(float h);
uint32_t advect_blocks(Dims d);
struct DivergenceDesc {
    float* div;
    const float* v[3];
    Dims d;
    float h;
    StencilBC dbc;
};
struct ProjectDesc {
    float* o[3];
    float* div;
    float* mag;
    const float* v[3];
    const float* p;
    Dims d;
    float h;
    StencilBC pbc;
    StencilBC dbc;
};
struct AddForceDesc {
    float* v[3];
    const float* f[3];
    float dt;
    uint64_t n;
};
struct AddScaledDesc {
    float* dst;
    const float* src;
    float dt;
    uint64_t n;
};
// blocks per simulation of the elementwise kernels (1024 elements)
uint32_t elem_block_count(uint64_t n);
hipError_t launch_add_force_batched(hipStream_t s, const AddForceDesc* descs, const BatchItem* items, uint32_t n_items);
hipError_t launch_add_scaled_batched(hipStream_t s, const AddScaledDesc* descs, const BatchItem* items, uint32_t n_items);
hipError_t launch_advect_velocity_batched(hipStream_t s, const AdvectVelocityDesc* descs, const BatchItem* items, uint32_t n_items);
hipError_t launch_advect_scalar_batched(hipStream_t s, const AdvectScalarDesc* descs, const BatchItem* items, uint32_t n_items);
// z-marching projection and divergence over a batch (fields below 2 GiB): one item per
// (descriptor, block) with block < z_march_blocks(d)
hipError_t launch_project_z_batched(hipStream_t s, const ProjectDesc* descs, const BatchItem* items, uint32_t n_items);
hipError_t launch_divergence_z_batched(hipStream_t s, const DivergenceDesc* descs, const BatchItem* items, uint32_t n_items);
uint32_t z_march_blocks(Dims d);

hipError_t launch_fill(hipStream_t s, float* p, float value, uint64_t n);
hipError_t launch_copy_scaled(hipStream_t s, float* dst, const float* src, float factor, uint64_t n);
hipError_t launch_add_scaled(hipStream_t s, float* dst, const float* src, float dt, uint64_t n);
// Coefficients of an animated force (stable_fluids_ext_set_force_animation_cuda).
struct ForceAnimation {
    float phase_rate, drift_x_amplitude, drift_z_amplitude, drift_z_rate, swirl_amplitude, swirl_rate;
};
// fx, fz of step *counter from the masks; launch_count_step adds one to *counter.
hipError_t launch_animate_force(hipStream_t s, float* fx, float* fz, const float* drift, const float* swirl_x, const float* swirl_z, const ForceAnimation& a, const uint32_t* counter, uint64_t n);
hipError_t launch_count_step(hipStream_t s, uint32_t* counter);
hipError_t launch_add_force(hipStream_t s, float* vx, float* vy, float* vz, const float* fx, const float* fy, const float* fz, float dt, uint64_t n);
hipError_t launch_advect_velocity(hipStream_t s, float* tx, float* ty, float* tz, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsFlowBoundaryConfig& bc);
// sf_advect.hip: the same stages through field buffer resources (fields below 2 GiB)
bool advect_buffers_fit(Dims d);
hipError_t launch_advect_velocity_buffers(hipStream_t s, float* tx, float* ty, float* tz, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsFlowBoundaryConfig& bc);
hipError_t launch_advect_scalar_buffers(hipStream_t s, float* out, const float* src, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc);
hipError_t launch_advect_scalar(hipStream_t s, float* out, const float* src, const float* vx, const float* vy, const float* vz, float dt, Dims d, float h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc);
hipError_t launch_smooth(hipStream_t s, SmootherKind smoother, SolveKind solve, const SmoothPass& pass, Dims d, const StencilBC& bc);
// The first `passes` (> 0) passes of an RBGS solve from zero in one launch that runs only
// when skip[0] == 0 (the learned warm start's fallback), ping-ponging io -> other -> io
// ..., bit-identical to `passes` launch_smooth calls from zero, with the result in io for
// an even count and in other for an odd one. Its `grid` workgroups (smooth_chain_grid,
// computed outside stream capture) take (pass, block) items in order and wait only for
// the neighbouring blocks of the previous pass, so the launch needs no co-residency.
// `counters` (smooth_chain_counters(d) words) must be zero before the first launch; each
// launch leaves them zero. skip[2] is set (and the result filled with NaN) if a wait ever
// exceeded its bound.
bool smooth_chain_supported(SmootherKind smoother, Dims d, const StencilBC& bc);
hipError_t smooth_chain_grid(Dims d, uint32_t* out_grid);
uint32_t smooth_chain_counters(Dims d);
hipError_t launch_smooth_chain(hipStream_t s, SolveKind solve, float* io, float* other, const float* b, float coef, int passes, Dims d, const StencilBC& bc, int* skip, unsigned* counters, uint32_t grid);
hipError_t launch_correction(hipStream_t s, SmootherKind smoother, SolveKind solve, StorageKind storage, const CorrectionPass& pass, Dims d, const StencilBC& homogeneous_bc);
// Defect of a refinement cycle: x' = x_in + s_e*e (x_in nullptr = zeros, e nullptr = no
// correction yet, s_e from e_slot), r = rhs - A x' with the full boundary rules
// (diffusion: rhs = b; Poisson: rhs = -coef*b), max|r| into *r_slot through `partials`
// (defect_partials(d) floats). x' is written to x_out when e is given.
hipError_t launch_defect(hipStream_t s, SolveKind solve, StorageKind storage, const float* x_in, const void* e, const float* e_slot, const float* b, float coef, float* x_out, float* r, float* partials, float* r_slot, Dims d, const StencilBC& bc);
uint64_t defect_partials(Dims d);
// out = (x + s_e*e) * out_scale, element-wise (out may equal x).
hipError_t launch_finalize(hipStream_t s, StorageKind storage, float* out, const float* x, const void* e, const float* e_slot, float out_scale, uint64_t n);
// divergence of (vx,vy,vz); bc.face[2c+side] is the rule of component c on axis c.
hipError_t launch_divergence(hipStream_t s, float* div, const float* vx, const float* vy, const float* vz, Dims d, float h, const StencilBC& div_bc);
// v_out = v_in - grad p; div and |v| of the projected field, all in one pass.
hipError_t launch_project(hipStream_t s, float* ox, float* oy, float* oz, float* div, float* mag, const float* vx, const float* vy, const float* vz, const float* p, Dims d, float h, const StencilBC& p_bc, const StencilBC& div_bc);

} // namespace sf
