// sf_context.hip -- host runtime and C ABI of libstable-fluids.so (gfx950).
//
// Implements the seven entry points of include/stable-fluids-3d.h (reference
// backend-cuda.cu:481-800) and the additive ones of include/stable-fluids-3d-ext.h.
// The whole step is enqueued once on the context stream under hipStreamBeginCapture and
// replayed with one hipGraphLaunch per step (reference :608-675, :721-726). The graph
// holds kernel nodes only: buffer copies of the reference graph are folded into
// out-of-place smoother passes and no memset node is recorded.
//
// Buffers (all packed float[nx*ny*nz], x fastest, so views stay plain device arrays):
//   flow: v{x,y,z} (view), t{x,y,z} (advected velocity = diffusion rhs), a{x,y,z}
//   (diffused velocity, projection input), p (view), p_aux, div (view), mag (view);
//   scalar: data (view), temp (advected = rhs), source (view), aux; vector: x, y, z.
// Smoother passes ping-pong between the output buffer and its aux partner; the pass
// count's parity picks the first target so the last pass lands in the output buffer.
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/stable-fluids-3d-ext.h"
#include "../../include/stable-fluids-3d.h"
#include "../../include/stable-fluids-unet.h"
#include "sf_kernels.hpp"
#include "sf_unet.hpp"
#include "sf_warm.hpp"

namespace sf {

namespace {

struct Range {
    explicit Range(const char* name) { roctxRangePush(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&)            = delete;
    Range& operator=(const Range&) = delete;
};

struct Options {
    SmootherKind smoother = kSmoothRBGS;
    bool refine           = false;     // refinement cycles (sf_mixed.hip) instead of plain passes
    StorageKind storage   = kStoreF32; // correction iterate storage of the cycles
    int interval          = 10;        // smoother passes per refinement cycle
    bool learned          = false;     // learned pressure warm start (sf_warm.hip)
    int warm_iterations   = 0;         // passes after an accepted guess
    int warm_factor       = 2;         // restriction factor of the network input
};

struct ScalarField {
    StableFluidsScalarFieldDesc desc{};
    float* data   = nullptr;
    float* temp   = nullptr;
    float* source = nullptr;
    float* aux    = nullptr;
};

struct VectorField {
    StableFluidsVectorFieldDesc desc{};
    float* x = nullptr;
    float* y = nullptr;
    float* z = nullptr;
    // animated force (stable_fluids_ext_set_force_animation_cuda); masks allocated on first use
    bool animated = false;
    ForceAnimation anim{};
    float* drift   = nullptr;
    float* swirl_x = nullptr;
    float* swirl_z = nullptr;
};

constexpr int kNormBlocks = 1024;
// partial sums of the norm reduction: 4 per workgroup, then the 4 totals
constexpr int kNormScratchDoubles = 4 * kNormBlocks + 4;

// Scratch of stable_fluids_ext_field_norms_noalloc_cuda (32,800 bytes): the same layout
// as a context's norm_partials. Every GPU the library runs on holds its own copy until
// the process ends. In host code this name is only a stand-in for the symbol: the device
// address comes from hipGetSymbolAddress (norm_scratch below), once per device.
static __device__ __align__(256) double scratch[kNormScratchDoubles];

struct Context {
    StableFluidsSimulationConfig config{};
    Options options{};
    hipStream_t stream = nullptr;
    bool owns_stream   = false;
    Dims dims{};
    uint64_t cells = 0;
    size_t bytes   = 0;
    hipGraph_t graph         = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    struct {
        float *vx = nullptr, *vy = nullptr, *vz = nullptr;
        float *tx = nullptr, *ty = nullptr, *tz = nullptr;
        float *ax = nullptr, *ay = nullptr, *az = nullptr;
        float *p = nullptr, *p_aux = nullptr, *div = nullptr, *mag = nullptr;
        float* rhs = nullptr; // pre-projection divergence (pressure right-hand side)
    } flow;
    std::vector<ScalarField> scalars;
    std::vector<VectorField> vectors;
    uint32_t* anim_step   = nullptr; // step counter of the force animations (allocated on first use)
    double* norm_partials = nullptr; // kNormBlocks * 4 partial sums, then 4 results
    struct {
        float* r     = nullptr; // FP32 defect
        void* e[2]   = {nullptr, nullptr}; // correction ping-pong (storage format)
        float* partials = nullptr; // per-workgroup max|r| of a defect pass
        float* slots = nullptr; // max|r| per (solve, cycle)
        int cycles   = 0;       // slots per solve
        int count    = 0;
    } mixed;
    struct {
        unet::Net* net   = nullptr;
        float* in2       = nullptr; // half-resolution (div, wall mask)
        float* scal      = nullptr; // sigma, 1/sigma
        double* partials = nullptr;
        double* stats    = nullptr; // guess residual, zero residual, fallback
        int* skip        = nullptr; // [0] 1: skip the fallback passes; [2] 1: the fallback solve of the last step failed
        uint32_t chain_grid = 0;    // workgroups of the one-launch fallback solve (0: not used)
        unsigned* chain_ctr = nullptr; // its ticket and per-block pass counters
    } warm;
    struct {
        double *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *partials = nullptr, *scal = nullptr;
    } cg;
    float* half[2] = {nullptr, nullptr}; // padded black-cell iterates of the RBGS passes (sf_rbgs_half.hip)
    int half_tile  = kHalfTile128x16;       // their tile shape and z-chunk length (0: default plan)
    int half_zc    = 0;
    std::vector<void*> allocations;

    ~Context() {
        if (warm.net) unet::net_destroy(warm.net);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocations) (void)hipFree(p);
        if (owns_stream && stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    bool alloc(T*& p, size_t nbytes) {
        void* raw = nullptr;
        if (hipMalloc(&raw, std::max<size_t>(nbytes, 4)) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        allocations.push_back(raw);
        p = static_cast<T*>(raw);
        return true;
    }
};

// ---- boundary rules for the stencil kernels (reference backend-cuda.cu:62-189) ----

const StableFluidsFlowBoundaryFaceDesc& flow_face(const StableFluidsFlowBoundaryConfig& b, int k) {
    const StableFluidsFlowBoundaryFaceDesc* faces[6] = {&b.x_minus, &b.x_plus, &b.y_minus, &b.y_plus, &b.z_minus, &b.z_plus};
    return *faces[k];
}
const StableFluidsScalarBoundaryFaceDesc& scalar_face(const StableFluidsScalarBoundaryConfig& b, int k) {
    const StableFluidsScalarBoundaryFaceDesc* faces[6] = {&b.x_minus, &b.x_plus, &b.y_minus, &b.y_plus, &b.z_minus, &b.z_plus};
    return *faces[k];
}

StencilBC scalar_rules(const StableFluidsScalarBoundaryConfig& b) {
    StencilBC bc{};
    for (int k = 0; k < 6; ++k) {
        const auto& f = scalar_face(b, k);
        bc.face[k]    = f.type == STABLE_FLUIDS_SCALAR_BOUNDARY_FIXED_VALUE ? FaceRule{kGhostConst, f.value} : FaceRule{kGhostCopy, 0.0f};
    }
    for (int a = 0; a < 3; ++a)
        if (scalar_face(b, 2 * a).type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC && scalar_face(b, 2 * a + 1).type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC) bc.periodic |= 1 << a;
    return bc;
}

int flow_periodic_mask(const StableFluidsFlowBoundaryConfig& b) {
    int m = 0;
    for (int a = 0; a < 3; ++a)
        if (flow_face(b, 2 * a).type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC && flow_face(b, 2 * a + 1).type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC) m |= 1 << a;
    return m;
}

StencilBC pressure_rules(const StableFluidsFlowBoundaryConfig& b) {
    StencilBC bc{};
    for (int k = 0; k < 6; ++k) {
        const auto& f = flow_face(b, k);
        bc.face[k]    = f.type == STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW ? FaceRule{kGhostConst, f.pressure} : FaceRule{kGhostCopy, 0.0f};
    }
    bc.periodic = flow_periodic_mask(b);
    return bc;
}

FaceRule velocity_rule(const StableFluidsFlowBoundaryFaceDesc& f, const int comp, const int axis) {
    if (f.type == STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW) return FaceRule{kGhostCopy, 0.0f};
    if (f.type == STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL && comp != axis) return FaceRule{kGhostCopy, 0.0f};
    const float prescribed = comp == 0 ? f.velocity_x : (comp == 1 ? f.velocity_y : f.velocity_z);
    return FaceRule{kGhostReflect, 2.0f * prescribed};
}

StencilBC velocity_rules(const StableFluidsFlowBoundaryConfig& b, const int comp) {
    StencilBC bc{};
    for (int k = 0; k < 6; ++k) bc.face[k] = velocity_rule(flow_face(b, k), comp, k >> 1);
    bc.periodic = flow_periodic_mask(b);
    return bc;
}

// face 2c+s: rule of velocity component c on its own axis c (divergence stencil)
StencilBC divergence_rules(const StableFluidsFlowBoundaryConfig& b) {
    StencilBC bc{};
    for (int k = 0; k < 6; ++k) bc.face[k] = velocity_rule(flow_face(b, k), k >> 1, k >> 1);
    bc.periodic = flow_periodic_mask(b);
    return bc;
}

bool valid_flow_axis(const StableFluidsFlowBoundaryFaceDesc& lo, const StableFluidsFlowBoundaryFaceDesc& hi) {
    if (lo.type > STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC || hi.type > STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC) return false;
    return (lo.type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC) == (hi.type == STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC);
}
bool valid_scalar_axis(const StableFluidsScalarBoundaryFaceDesc& lo, const StableFluidsScalarBoundaryFaceDesc& hi) {
    if (lo.type > STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC || hi.type > STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC) return false;
    return (lo.type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC) == (hi.type == STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC);
}

// ---- step enqueue ----

#define SF_TRY(expr)                                  \
    do {                                              \
        if ((expr) != hipSuccess) return hipErrorUnknown; \
    } while (0)

StencilBC homogeneous(StencilBC bc) {
    for (auto& f : bc.face) f.value = 0.0f;
    return bc;
}

// K smoother passes as refinement cycles with a 16-bit correction iterate (sf_mixed.hip):
// per cycle one FP32 defect pass, then up to `interval` correction passes, and a final
// pass writing (x + s*e) * last_scale into `out`. x ping-pongs between aux and out.
hipError_t enqueue_solve_mixed(Context& c, SolveKind solve, const float* b, const float* x0, float* out, float* aux, int passes, float coef, float last_scale, const StencilBC& bc, int slot_base) {
    auto& m                = c.mixed;
    const StencilBC bc0    = homogeneous(bc);
    const float corr_coef  = solve == kSolvePoisson ? -1.0f : coef;
    const int cycles       = (passes + c.options.interval - 1) / c.options.interval;
    const float* x         = x0;
    const void* e          = nullptr;
    const float* e_slot    = nullptr;
    for (int cyc = 0; cyc < cycles; ++cyc) {
        float* slot  = m.slots + slot_base + cyc;
        float* x_out = nullptr;
        if (e != nullptr) x_out = (x == aux) ? out : aux;
        SF_TRY(launch_defect(c.stream, solve, c.options.storage, x, e, e_slot, b, coef, x_out, m.r, m.partials, slot, c.dims, bc));
        if (x_out != nullptr) x = x_out;
        const int sweeps = std::min(c.options.interval, passes - cyc * c.options.interval);
        const void* src  = nullptr;
        for (int i = 0; i < sweeps; ++i) {
            void* dst = m.e[i & 1];
            SF_TRY(launch_correction(c.stream, c.options.smoother, solve, c.options.storage, CorrectionPass{src, m.r, dst, corr_coef, slot}, c.dims, bc0));
            src = dst;
        }
        e      = src;
        e_slot = slot;
    }
    return launch_finalize(c.stream, c.options.storage, out, x, e, e_slot, last_scale, c.cells);
}

// Whether this context's FP32 smoother passes run two per launch under these rules.

// Whether a plain FP32 RBGS solve under these rules runs on the black-cell passes.
bool use_half(const Context& c, const StencilBC& bc) {
    return c.half[0] != nullptr && !c.options.refine && c.options.smoother == kSmoothRBGS && rbgs_half_supported(c.dims, bc);
}

// K RBGS passes from x0 (nullptr = zero guess) into `out` that carry only the black cells
// (plus the ghosts of the red face cells) between passes: pass l writes half[l & 1], the
// first reads the natural x0, the last writes `out` scaled by last_scale.
hipError_t enqueue_solve_half(Context& c, SolveKind solve, const float* b, const float* x0, float* out, int passes, float coef, float last_scale, const StencilBC& bc) {
    HalfPassArgs a{};
    a.b    = b;
    a.coef = coef;
    a.d    = c.dims;
    a.bc   = bc;
    a.g    = half_plan(c.dims, c.half_tile, c.half_zc);
    a.x_nat = x0;
    a.x_out = out;
    for (int l = 0; l < passes; ++l) {
        const bool first = l == 0, last = l == passes - 1;
        const int in     = first ? (x0 != nullptr ? kHalfInNatural : kHalfInZero) : kHalfInHalf;
        a.h_in           = c.half[(l + 1) & 1];
        a.h_out          = c.half[l & 1];
        a.out_scale      = last ? last_scale : 1.0f;
        SF_TRY(launch_rbgs_half(c.stream, solve, in, last ? kHalfOutNatural : kHalfOutHalf, a, c.half_tile));
    }
    return hipSuccess;
}

// K smoother passes from x0 (nullptr = zero guess) into `out`, ping-ponging with `aux`:
// pass l of K writes out when K - 1 - l is even, so the last one lands in out.
hipError_t enqueue_solve(Context& c, SolveKind solve, const float* b, const float* x0, float* out, float* aux, int passes, float coef, float last_scale, const StencilBC& bc, int slot_base) {
    if (c.options.refine) return enqueue_solve_mixed(c, solve, b, x0, out, aux, passes, coef, last_scale, bc, slot_base);
    if (use_half(c, bc)) return enqueue_solve_half(c, solve, b, x0, out, passes, coef, last_scale, bc);
    const float* src = x0;
    for (int l = 0; l < passes; ++l) {
        float* dst = ((passes - 1 - l) % 2 == 0) ? out : aux;
        const SmoothPass pass{src, b, dst, coef, l == passes - 1 ? last_scale : 1.0f};
        SF_TRY(launch_smooth(c.stream, c.options.smoother, solve, pass, c.dims, bc));
        src = dst;
    }
    return hipSuccess;
}

// Learned warm start in place of the zero-initialised pressure solve (reference
// backend-cuda.cu:643-650): network guess from the half-resolution divergence, residual
// test, then P passes of which the first C = P - K are skipped when the guess is
// accepted; on fallback all P passes run from zero, exactly the plain solve.
hipError_t enqueue_learned_pressure(Context& c, const StencilBC& p_bc) {
    auto& f        = c.flow;
    auto& w        = c.warm;
    const int P    = c.config.pressure_iterations;
    const int K    = c.options.warm_iterations;
    const int C    = P - K;
    const float h  = c.config.cell_size;
    const int fac  = c.options.warm_factor;
    const Dims cd  = unet::net_dims(*w.net);
    const uint64_t vc = static_cast<uint64_t>(cd.nx) * cd.ny * cd.nz;
    // The P passes run as a list of launches. The C fallback-only passes come first: one
    // launch of the one-launch fallback solve where the grid allows it (the accepted guess
    // then costs one exiting dispatch for all C passes), else single passes, each skipped
    // dispatch costing a few microseconds; the K passes after them run single. Launch l
    // of L writes dst(l), the last one p. The launches covering the first C passes are
    // skipped on an accepted guess, which therefore goes where the next launch reads it,
    // and the first launch starts from zero when it can be skipped (it only runs on
    // fallback, as the plain solve's first launch). Chain and single launches give
    // identical results.
    const bool chain = C > 0 && w.chain_grid > 0;
    std::vector<int> per; // passes per launch
    if (chain) {
        per.push_back(C);
        per.insert(per.end(), K, 1);
    } else {
        per.assign(P, 1);
    }
    const int L  = static_cast<int>(per.size());
    int skipped  = 0; // launches covering the first C passes
    for (int acc = 0; acc < C;) acc += per[skipped++];
    auto dst     = [&](const int l) { return ((L - 1 - l) & 1) == 0 ? f.p : f.p_aux; };
    float* guess = dst(skipped - 1);
    // restriction, then sigma and the packing of the network input in one launch
    SF_TRY(warm::launch_restrict_input(c.stream, f.rhs, w.in2, w.partials, cd, fac, p_bc));
    SF_TRY(unet::launch_pack_input_sigma(c.stream, w.in2, w.partials, static_cast<int>(warm::restrict_partials(cd)), w.scal, unet::net_input(*w.net), vc));
    SF_TRY(unet::net_forward_packed(*w.net, false));
    SF_TRY(warm::launch_prolong(c.stream, unet::net_output(*w.net), w.scal, static_cast<float>(fac * fac) * h * h, guess, cd, fac));
    // (the chain zeroes the guess itself on fallback)
    SF_TRY(warm::launch_guess_test(c.stream, guess, f.rhs, h * h, 10.0, w.partials, w.stats, w.skip, c.dims, p_bc, !chain));
    // the K passes after the skippable ones as black-cell passes from the guess buffer (the
    // accepted guess, or the fallback result the chain leaves there), when nothing after
    // the chain is skippable
    const bool half_k = K > 0 && (chain || C == 0) && use_half(c, p_bc);
    for (int l = 0; l < L; ++l) {
        if (half_k && l >= skipped) {
            SF_TRY(enqueue_solve_half(c, kSolvePoisson, f.rhs, guess, f.p, K, h * h, 1.0f, p_bc));
            break;
        }
        if (chain && l == 0) {
            // the chain's result lands in its first buffer for an even pass count, in its
            // second for an odd one: either way in the guess buffer
            float* other = guess == f.p ? f.p_aux : f.p;
            if (C % 2 == 0) SF_TRY(launch_smooth_chain(c.stream, kSolvePoisson, guess, other, f.rhs, h * h, C, c.dims, p_bc, w.skip, w.chain_ctr, w.chain_grid));
            else SF_TRY(launch_smooth_chain(c.stream, kSolvePoisson, other, guess, f.rhs, h * h, C, c.dims, p_bc, w.skip, w.chain_ctr, w.chain_grid));
            continue;
        }
        const float* src = l == 0 ? (C > 0 ? nullptr : guess) : dst(l - 1);
        SmoothPass pass{src, f.rhs, dst(l), h * h, 1.0f};
        if (l < skipped) pass.skip = w.skip;
        SF_TRY(launch_smooth(c.stream, c.options.smoother, kSolvePoisson, pass, c.dims, p_bc));
    }
    return hipSuccess;
}

// The pressure solve of the step from f.rhs into f.p: learned warm start or P passes from
// zero (reference :643-650: the pressure restarts from zero).
hipError_t enqueue_pressure_solve(Context& c) {
    const auto& cfg      = c.config;
    auto& f              = c.flow;
    const StencilBC p_bc = pressure_rules(cfg.flow_boundary);
    if (cfg.pressure_iterations > 0 && c.options.learned) return enqueue_learned_pressure(c, p_bc);
    if (cfg.pressure_iterations > 0) return enqueue_solve(c, kSolvePoisson, f.rhs, nullptr, f.p, f.p_aux, cfg.pressure_iterations, cfg.cell_size * cfg.cell_size, 1.0f, p_bc, 3 * c.mixed.cycles);
    return launch_fill(c.stream, f.p, 0.0f, c.cells);
}

// Projection of the diffused velocity u (reference :643-650): divergence, pressure solve,
// then the gradient subtraction fused with the post-projection divergence and |v|.
hipError_t enqueue_projection(Context& c, const float* const (&u)[3]) {
    const auto& fb         = c.config.flow_boundary;
    const float h          = c.config.cell_size;
    auto& f                = c.flow;
    const StencilBC div_bc = divergence_rules(fb);
    const StencilBC p_bc   = pressure_rules(fb);
    SF_TRY(launch_divergence(c.stream, f.rhs, u[0], u[1], u[2], c.dims, h, div_bc));
    SF_TRY(enqueue_pressure_solve(c));
    SF_TRY(launch_project(c.stream, f.vx, f.vy, f.vz, f.div, f.mag, u[0], u[1], u[2], f.p, c.dims, h, p_bc, div_bc));
    return hipSuccess;
}

// The velocity the projection reads: the diffused one, or the advected one when there are
// no diffusion passes (positive viscosity); a copy otherwise (enqueue_step).
void diffused_velocity(Context& c, const float* (&u)[3]) {
    auto& f           = c.flow;
    const auto& cfg   = c.config;
    const float alpha = cfg.dt * cfg.viscosity / (cfg.cell_size * cfg.cell_size);
    const bool adv    = alpha > 0.0f && cfg.diffuse_iterations <= 0;
    u[0]              = adv ? f.tx : f.ax;
    u[1]              = adv ? f.ty : f.ay;
    u[2]              = adv ? f.tz : f.az;
}

// The step sequence of the reference graph (backend-cuda.cu:612-671).
hipError_t enqueue_step(Context& c) {
    const auto& cfg = c.config;
    const auto& fb  = cfg.flow_boundary;
    const float dt  = cfg.dt;
    const float h   = cfg.cell_size;
    auto& f         = c.flow;

    bool animated = false;
    for (auto& v : c.vectors) {
        if (v.desc.usage != STABLE_FLUIDS_VECTOR_FIELD_FORCE) continue;
        if (v.animated) SF_TRY(launch_animate_force(c.stream, v.x, v.z, v.drift, v.swirl_x, v.swirl_z, v.anim, c.anim_step, c.cells));
        animated = animated || v.animated;
        SF_TRY(launch_add_force(c.stream, f.vx, f.vy, f.vz, v.x, v.y, v.z, dt, c.cells));
    }
    if (animated) SF_TRY(launch_count_step(c.stream, c.anim_step));
    SF_TRY(launch_advect_velocity(c.stream, f.tx, f.ty, f.tz, f.vx, f.vy, f.vz, dt, c.dims, h, fb));

    // viscosity; alpha <= 0 keeps the forced velocity (reference :621-642)
    const float alpha = cfg.dt * cfg.viscosity / (cfg.cell_size * cfg.cell_size);
    const float* u[3];
    if (alpha > 0.0f) {
        if (cfg.diffuse_iterations <= 0) {
            u[0] = f.tx;
            u[1] = f.ty;
            u[2] = f.tz;
        } else {
            float* t[3] = {f.tx, f.ty, f.tz};
            float* a[3] = {f.ax, f.ay, f.az};
            float* v[3] = {f.vx, f.vy, f.vz};
            for (int comp = 0; comp < 3; ++comp) SF_TRY(enqueue_solve(c, kSolveDiffusion, t[comp], t[comp], a[comp], v[comp], cfg.diffuse_iterations, alpha, 1.0f, velocity_rules(fb, comp), comp * c.mixed.cycles));
            u[0] = f.ax;
            u[1] = f.ay;
            u[2] = f.az;
        }
    } else {
        SF_TRY(launch_copy_scaled(c.stream, f.ax, f.vx, 1.0f, c.cells));
        SF_TRY(launch_copy_scaled(c.stream, f.ay, f.vy, 1.0f, c.cells));
        SF_TRY(launch_copy_scaled(c.stream, f.az, f.vz, 1.0f, c.cells));
        u[0] = f.ax;
        u[1] = f.ay;
        u[2] = f.az;
    }

    SF_TRY(enqueue_projection(c, u));

    // scalars (reference :652-668)
    for (size_t si = 0; si < c.scalars.size(); ++si) {
        auto& s = c.scalars[si];
        SF_TRY(launch_add_scaled(c.stream, s.data, s.source, dt, c.cells));
        SF_TRY(launch_advect_scalar(c.stream, s.temp, s.data, f.vx, f.vy, f.vz, dt, c.dims, h, s.desc.boundary, fb));
        const float alpha_s = cfg.dt * s.desc.diffusion / (cfg.cell_size * cfg.cell_size);
        const bool dissipate = s.desc.dissipation > 0.0f;
        const float factor  = dissipate ? 1.0f / (1.0f + cfg.dt * s.desc.dissipation) : 1.0f;
        if (alpha_s > 0.0f) {
            if (cfg.diffuse_iterations <= 0)
                SF_TRY(launch_copy_scaled(c.stream, s.data, s.temp, factor, c.cells));
            else
                SF_TRY(enqueue_solve(c, kSolveDiffusion, s.temp, s.temp, s.data, s.aux, cfg.diffuse_iterations, alpha_s, factor, scalar_rules(s.desc.boundary), static_cast<int>(4 + si) * c.mixed.cycles));
        } else if (dissipate) {
            SF_TRY(launch_copy_scaled(c.stream, s.data, s.data, factor, c.cells));
        }
    }
    return hipSuccess;
}

#undef SF_TRY

StableFluidsResult build_graph(Context& c) {
    Range r("stable.create_context.graph");
    if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const hipError_t enq = enqueue_step(c);
    hipGraph_t graph     = nullptr;
    const hipError_t end = hipStreamEndCapture(c.stream, &graph);
    if (enq != hipSuccess || end != hipSuccess || graph == nullptr) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    c.graph = graph;
    if (hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return STABLE_FLUIDS_RESULT_OK;
}

// Re-captures the step graph (after a change of what it launches).
StableFluidsResult rebuild_graph(Context& c) {
    if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph      = nullptr;
    return build_graph(c);
}

StableFluidsResult create(const StableFluidsContextCreateDesc* desc, const StableFluidsExtOptions* opt, void** out_context, StableFluidsScalarFieldHandle* out_s, uint32_t cap_s, StableFluidsVectorFieldHandle* out_v, uint32_t cap_v) {
    Range r("stable.create_context");
    if (out_context == nullptr || desc == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_context = nullptr;
    const auto& cfg = desc->config;

    // validation happens before any HIP call (reference :516-528, :536-545)
    Options options{};
    if (opt != nullptr) {
        if (opt->struct_size != sizeof(StableFluidsExtOptions)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (opt->smoother > STABLE_FLUIDS_EXT_SMOOTHER_JACOBI) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (opt->iterate_storage > STABLE_FLUIDS_EXT_STORAGE_F32_REFINED) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (opt->defect_correction_interval > 4096) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (opt->pressure_warm_start > STABLE_FLUIDS_EXT_WARM_START_LEARNED) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (opt->flags != 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE; // no flag is defined
        options.smoother = opt->smoother == STABLE_FLUIDS_EXT_SMOOTHER_JACOBI ? kSmoothJacobi : kSmoothRBGS;
        options.refine   = opt->iterate_storage != STABLE_FLUIDS_EXT_STORAGE_F32;
        options.storage  = opt->iterate_storage == STABLE_FLUIDS_EXT_STORAGE_F16 ? kStoreF16 : (opt->iterate_storage == STABLE_FLUIDS_EXT_STORAGE_BF16 ? kStoreBF16 : kStoreF32);
        if (opt->defect_correction_interval > 0) options.interval = static_cast<int>(opt->defect_correction_interval);
        options.learned = opt->pressure_warm_start == STABLE_FLUIDS_EXT_WARM_START_LEARNED;
        if (options.learned) {
            const int P = cfg.pressure_iterations;
            int k       = static_cast<int>(std::min<uint32_t>(opt->warm_start_iterations, 1u << 20));
            const uint32_t cl = opt->warm_start_coarsening == 0 ? 1u : opt->warm_start_coarsening;
            if (cl > 3) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            const int fac = 1 << cl, m = 4 * fac;
            if (P <= 0 || k > P || options.refine || cfg.nx % m || cfg.ny % m || cfg.nz % m) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            options.warm_iterations = k;
            options.warm_factor     = fac;
        }
    }
    if (cfg.nx <= 0 || cfg.ny <= 0 || cfg.nz <= 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const auto& fb = cfg.flow_boundary;
    if (!valid_flow_axis(fb.x_minus, fb.x_plus) || !valid_flow_axis(fb.y_minus, fb.y_plus) || !valid_flow_axis(fb.z_minus, fb.z_plus)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (desc->scalar_field_count > 0 && desc->scalar_fields == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (desc->vector_field_count > 0 && desc->vector_fields == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    for (uint32_t i = 0; i < desc->scalar_field_count; ++i) {
        const auto& b = desc->scalar_fields[i].boundary;
        if (!valid_scalar_axis(b.x_minus, b.x_plus) || !valid_scalar_axis(b.y_minus, b.y_plus) || !valid_scalar_axis(b.z_minus, b.z_plus)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }

    // the narrow-storage RBGS correction runs on the split smoother, which needs even
    // extents on periodic axes
    if (options.refine && options.smoother == kSmoothRBGS) {
        const Dims d{cfg.nx, cfg.ny, cfg.nz};
        auto odd_periodic = [&](const int mask) { return ((mask & 1) && (d.nx & 1)) || ((mask & 2) && (d.ny & 1)) || ((mask & 4) && (d.nz & 1)); };
        if (odd_periodic(flow_periodic_mask(fb))) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        for (uint32_t i = 0; i < desc->scalar_field_count; ++i)
            if (odd_periodic(scalar_rules(desc->scalar_fields[i].boundary).periodic)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }

    std::unique_ptr<Context> c{new (std::nothrow) Context{}};
    if (!c) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
    c->config  = cfg;
    c->options = options;
    c->dims    = Dims{cfg.nx, cfg.ny, cfg.nz};
    c->cells   = static_cast<uint64_t>(cfg.nx) * static_cast<uint64_t>(cfg.ny) * static_cast<uint64_t>(cfg.nz);
    c->bytes   = c->cells * sizeof(float);
    c->stream  = static_cast<hipStream_t>(desc->stream);
    c->half_tile = default_half_tile(c->dims);
    if (c->stream == nullptr) {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        c->owns_stream = true;
    }
    for (uint32_t i = 0; i < desc->scalar_field_count; ++i) {
        c->scalars.push_back(ScalarField{desc->scalar_fields[i]});
        if (out_s != nullptr && i < cap_s) out_s[i] = i + 1u;
    }
    for (uint32_t i = 0; i < desc->vector_field_count; ++i) {
        c->vectors.push_back(VectorField{desc->vector_fields[i]});
        if (out_v != nullptr && i < cap_v) out_v[i] = i + 1u;
    }

    {
        Range ra("stable.create_context.alloc");
        auto& f = c->flow;
        for (float** p : {&f.vx, &f.vy, &f.vz, &f.tx, &f.ty, &f.tz, &f.ax, &f.ay, &f.az, &f.p, &f.p_aux, &f.div, &f.mag, &f.rhs})
            if (!c->alloc(*p, c->bytes)) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (c->options.learned) {
            auto& w = c->warm;
            const int fac = c->options.warm_factor;
            if (unet::net_create(cfg.nx / fac, cfg.ny / fac, cfg.nz / fac, c->stream, &w.net) != hipSuccess) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
            if (!c->alloc(w.in2, c->bytes / 4) || !c->alloc(w.scal, 16) || !c->alloc(w.partials, sizeof(double) * 2 * warm::kPartialBlocks) || !c->alloc(w.stats, 4 * sizeof(double)) || !c->alloc(w.skip, 16)) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
            const bool chain = cfg.pressure_iterations > c->options.warm_iterations && smooth_chain_supported(c->options.smoother, c->dims, pressure_rules(cfg.flow_boundary));
            if (chain && smooth_chain_grid(c->dims, &w.chain_grid) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (chain && !c->alloc(w.chain_ctr, sizeof(unsigned) * smooth_chain_counters(c->dims))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        }
        for (auto& s : c->scalars)
            for (float** p : {&s.data, &s.temp, &s.source, &s.aux})
                if (!c->alloc(*p, c->bytes)) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        for (auto& v : c->vectors)
            for (float** p : {&v.x, &v.y, &v.z})
                if (!c->alloc(*p, c->bytes)) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (!c->alloc(c->norm_partials, sizeof(double) * (kNormBlocks * 4 + 4))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        // black-cell iterates, when some solve of a plain RBGS context can use them
        bool half = !c->options.refine && c->options.smoother == kSmoothRBGS;
        if (half) {
            bool any = rbgs_half_supported(c->dims, pressure_rules(cfg.flow_boundary)) || rbgs_half_supported(c->dims, velocity_rules(cfg.flow_boundary, 0));
            for (const auto& sc : c->scalars) any = any || rbgs_half_supported(c->dims, scalar_rules(sc.desc.boundary));
            half = any;
        }
        if (half)
            for (float*& h : c->half)
                if (!c->alloc(h, sizeof(float) * half_buffer_floats(c->dims))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (c->options.refine) {
            auto& m          = c->mixed;
            const size_t eb  = c->cells * (c->options.storage == kStoreF32 ? 4 : 2);
            const int passes = std::max(std::max(cfg.diffuse_iterations, cfg.pressure_iterations), 1);
            m.cycles         = (passes + c->options.interval - 1) / c->options.interval;
            m.count          = (4 + static_cast<int>(c->scalars.size())) * m.cycles;
            if (!c->alloc(m.r, c->bytes) || !c->alloc(m.e[0], eb) || !c->alloc(m.e[1], eb) || !c->alloc(m.slots, sizeof(float) * static_cast<size_t>(m.count)) || !c->alloc(m.partials, sizeof(float) * defect_partials(c->dims))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        }
    }

    {
        // initial state (reference :583-606)
        Range ri("stable.create_context.init");
        auto& f = c->flow;
        for (float* p : {f.vx, f.vy, f.vz, f.tx, f.ty, f.tz, f.ax, f.ay, f.az, f.p, f.p_aux, f.div, f.mag, f.rhs})
            if (launch_fill(c->stream, p, 0.0f, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        for (float* h : c->half)
            if (h != nullptr && launch_fill(c->stream, h, 0.0f, half_buffer_floats(c->dims)) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (c->options.learned && launch_fill(c->stream, reinterpret_cast<float*>(c->warm.stats), 0.0f, 8) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (c->options.learned && launch_fill(c->stream, reinterpret_cast<float*>(c->warm.skip), 0.0f, 4) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        if (c->warm.chain_ctr != nullptr && launch_fill(c->stream, reinterpret_cast<float*>(c->warm.chain_ctr), 0.0f, smooth_chain_counters(c->dims)) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        for (auto& s : c->scalars) {
            if (launch_fill(c->stream, s.data, s.desc.initial_value, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (launch_fill(c->stream, s.temp, s.desc.initial_value, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (launch_fill(c->stream, s.aux, s.desc.initial_value, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (launch_fill(c->stream, s.source, 0.0f, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        }
        for (auto& v : c->vectors) {
            if (launch_fill(c->stream, v.x, v.desc.initial_value_x, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (launch_fill(c->stream, v.y, v.desc.initial_value_y, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            if (launch_fill(c->stream, v.z, v.desc.initial_value_z, c->cells) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        }
    }

    const StableFluidsResult g = build_graph(*c);
    if (g != STABLE_FLUIDS_RESULT_OK) return g;
    if (hipGetLastError() != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_context = c.release();
    return STABLE_FLUIDS_RESULT_OK;
}

StableFluidsResult status(const hipError_t e) { return e == hipSuccess ? STABLE_FLUIDS_RESULT_OK : STABLE_FLUIDS_RESULT_BACKEND_FAILURE; }

// Copy `values` (device, NULL = zeros) into dst on the context stream.
hipError_t upload(Context& c, float* dst, const float* values) {
    if (values == nullptr) return launch_fill(c.stream, dst, 0.0f, c.cells);
    return hipMemcpyAsync(dst, values, c.bytes, hipMemcpyDeviceToDevice, c.stream);
}

// ---- norms: two deterministic passes, FP64 accumulation ----

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <int kCount>
__device__ void block_reduce(double (&acc)[kCount], const bool is_max_last, double* out) {
    __shared__ double lds[4][kCount];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kCount; ++k) acc[k] = (is_max_last && k == kCount - 1) ? wave_max(acc[k]) : wave_sum(acc[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < kCount; ++k) lds[wave][k] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kCount; ++k) {
            double r = lds[0][k];
            for (int w = 1; w < 4; ++w) r = (is_max_last && k == kCount - 1) ? fmax(r, lds[w][k]) : r + lds[w][k];
            out[k] = r;
        }
}

// partials[block] = {sum div^2, sum |v|^2, sum p^2, max |div|}
__global__ __launch_bounds__(256) void k_norm_partials(const float* __restrict__ div, const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ p, const uint64_t n, double* __restrict__ partials) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * 256) {
        const double dv = div[i], x = vx[i], y = vy[i], z = vz[i], pv = p[i];
        acc[0] += dv * dv;
        acc[1] += x * x + y * y + z * z;
        acc[2] += pv * pv;
        acc[3] = fmax(acc[3], fabs(dv));
    }
    block_reduce<4>(acc, true, partials + 4 * blockIdx.x);
}

__global__ __launch_bounds__(256) void k_norm_final(double* __restrict__ partials, const int blocks) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < blocks; b += 256) {
        acc[0] += partials[4 * b + 0];
        acc[1] += partials[4 * b + 1];
        acc[2] += partials[4 * b + 2];
        acc[3] = fmax(acc[3], partials[4 * b + 3]);
    }
    block_reduce<4>(acc, true, partials + 4 * kNormBlocks);
}

// totals {sum div^2, sum |v|^2, sum p^2, max |div|} -> StableFluidsExtNorms in device memory
__global__ __launch_bounds__(64) void k_norm_export(const double* __restrict__ totals, StableFluidsExtNorms* __restrict__ out) {
    if (threadIdx.x != 0) return;
    out->divergence_l2  = sqrt(totals[0]);
    out->divergence_max = totals[3];
    out->velocity_l2    = sqrt(totals[1]);
    out->pressure_l2    = sqrt(totals[2]);
}

// Device address of the current GPU's norm scratch, looked up on the first call on that
// device (with it current) and kept per device.
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

hipError_t norm_scratch(const int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_ptr[device] == nullptr) {
        void* p            = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}


// ---- batched simulations: one launch per stage for many contexts ----
//
// The batch replays each context's step sequence (enqueue_step) with every stage
// launched once for all simulations: workgroup i runs block items[i].block of the
// stage descriptor items[i].desc (a simulation, or a simulation and velocity
// component). The arithmetic of each block is the single-simulation kernel's, so
// every simulation ends bit-identical to stepping it alone. RBGS solves of simulations
// with even nx run on the black-cell passes (one launch per tile shape in the batch),
// the rest on the full-iterate passes (one launch per nx parity). The tile shape and
// z-chunk length per simulation are the tuned parameters (half_plan, plan_tiles); they
// change which workgroup computes a cell, never the value.

struct Batch {
    std::vector<Context*> sims;
    hipStream_t stream = nullptr;
    bool owns_stream   = false;
    std::vector<int> zc;   // smoother z-chunk per simulation (0 = default)
    std::vector<int> tile; // black-cell tile shape per simulation
    // black-cell iterates of velocity components 1 and 2 (component 0, the pressure and
    // the scalars use the context's own pair): all three components advance together
    std::vector<std::array<float*, 4>> hv;
    std::vector<void*> owned; // hv, freed with the batch
    hipGraph_t graph         = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<void*> allocations;
    std::vector<std::function<hipError_t()>> plan; // the step's launches, in order
    int launches = 0;

    ~Batch() {
        release();
        for (void* p : owned) (void)hipFree(p);
        if (owns_stream && stream) (void)hipStreamDestroy(stream);
    }
    void release() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph      = nullptr;
        for (void* p : allocations) (void)hipFree(p);
        allocations.clear();
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        void* raw = nullptr;
        if (hipMalloc(&raw, std::max<size_t>(v.size() * sizeof(T), 16)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        allocations.push_back(raw);
        if (!v.empty() && hipMemcpy(raw, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return static_cast<T*>(raw);
    }
};

bool batch_compatible(const Context& a, const Context& b) {
    const auto &x = a.config, &y = b.config;
    if (x.diffuse_iterations != y.diffuse_iterations || x.pressure_iterations != y.pressure_iterations) return false;
    if (a.scalars.size() != b.scalars.size() || a.vectors.size() != b.vectors.size()) return false;
    if (a.options.smoother != b.options.smoother) return false;
    return true;
}

bool batch_supported(const Context& c) {
    const auto& cfg = c.config;
    if (c.options.refine || c.options.learned) return false;
    for (const auto& v : c.vectors)
        if (v.animated) return false;
    if (cfg.diffuse_iterations <= 0 || cfg.pressure_iterations <= 0 || !advect_buffers_fit(c.dims)) return false;
    if (!(cfg.dt * cfg.viscosity / (cfg.cell_size * cfg.cell_size) > 0.0f)) return false;
    for (const auto& s : c.scalars)
        if (!(cfg.dt * s.desc.diffusion / (cfg.cell_size * cfg.cell_size) > 0.0f)) return false;
    const auto& fb = cfg.flow_boundary;
    if (!smooth_batchable(c.options.smoother, c.dims, pressure_rules(fb))) return false;
    for (int comp = 0; comp < 3; ++comp)
        if (!smooth_batchable(c.options.smoother, c.dims, velocity_rules(fb, comp))) return false;
    for (const auto& s : c.scalars)
        if (!smooth_batchable(c.options.smoother, c.dims, scalar_rules(s.desc.boundary))) return false;
    return true;
}

// Work list of a stage: for each descriptor d, its blocks 0..blocks(d)-1.
template <class F>
std::vector<BatchItem> batch_items(const size_t n_desc, F blocks) {
    std::vector<BatchItem> items;
    for (size_t d = 0; d < n_desc; ++d)
        for (uint32_t b = 0, n = blocks(d); b < n; ++b) items.push_back(BatchItem{static_cast<uint32_t>(d), b});
    return items;
}

// A work list laid out for the round-robin dispatch of workgroups over the 8 XCDs:
// position w holds item (w & 7) * n / 8 + w / 8 of the list padded to a multiple n of 8
// with idle items, so each XCD walks a contiguous run of neighbouring blocks through its
// own L2 (the kernels taking it return on kIdleItem).
std::vector<BatchItem> xcd_order(const std::vector<BatchItem>& logical) {
    const size_t n8 = (logical.size() + 7) / 8 * 8, per = n8 / 8;
    std::vector<BatchItem> items(n8, BatchItem{kIdleItem, 0});
    for (size_t w = 0; w < n8; ++w) {
        const size_t j = (w & 7) * per + (w >> 3);
        if (j < logical.size()) items[w] = logical[j];
    }
    return items;
}

// Smoother work lists split by row parity (the pair-load kernels need even nx): [0] the
// items of descriptors with odd nx, [1] those with even nx; one launch per non-empty list.
struct SmoothLists {
    const BatchItem* items[2] = {nullptr, nullptr};
    uint32_t n[2]             = {0, 0};
    int launches() const { return (n[0] > 0) + (n[1] > 0); }
};

template <class Nx, class F>
bool smooth_lists(Batch& B, const size_t n_desc, Nx nx_of, F blocks, SmoothLists* out) {
    for (int parity = 0; parity < 2; ++parity) {
        std::vector<BatchItem> items;
        for (size_t d = 0; d < n_desc; ++d) {
            if (((nx_of(d) & 1) == 0) != (parity == 1)) continue;
            for (uint32_t b = 0, n = blocks(d); b < n; ++b) items.push_back(BatchItem{static_cast<uint32_t>(d), b});
        }
        out->n[parity] = static_cast<uint32_t>(items.size());
        if (items.empty()) continue;
        out->items[parity] = B.upload(items);
        if (!out->items[parity]) return false;
    }
    return true;
}

hipError_t launch_smooth_lists(hipStream_t s, const SmootherKind sm, const SolveKind solve, const bool zero, const SmoothDesc* descs, const SmoothLists& L) {
    for (int parity = 0; parity < 2; ++parity) {
        if (L.n[parity] == 0) continue;
        const hipError_t e = launch_smooth_batched(s, sm, solve, zero, parity == 1, descs, L.items[parity], L.n[parity]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Whether simulation i runs its RBGS solves under rules bc (velocity component comp, or
// -1 for a pressure or scalar solve) on the black-cell passes.
bool batch_half(const Batch& B, const size_t i, const StencilBC& bc, const int comp) {
    const Context& c = *B.sims[i];
    if (c.half[0] == nullptr || c.options.smoother != kSmoothRBGS || !rbgs_half_supported(c.dims, bc)) return false;
    return comp <= 0 || B.hv[i][0] != nullptr;
}

// Half-buffer pair of simulation i for velocity component comp (-1: pressure, scalars).
std::array<float*, 2> batch_half_pair(const Batch& B, const size_t i, const int comp) {
    if (comp <= 0) return {B.sims[i]->half[0], B.sims[i]->half[1]};
    return {B.hv[i][2 * (comp - 1)], B.hv[i][2 * (comp - 1) + 1]};
}

TileGeom batch_half_plan(const Batch& B, const size_t i) { return half_plan(B.sims[i]->dims, B.tile[i], B.zc[i]); }

// Work lists of one solve family: descriptor d is simulation sim_of(d)'s; it runs the
// black-cell passes when half_of(d), one list per tile shape, otherwise the full-iterate
// passes, one list per nx parity. A black-cell list orders its descriptors by z-chunk
// length, longest first (they finish last), in XCD order (xcd_order).
struct PassLists {
    const BatchItem* half[kHalfTileCount] = {};
    uint32_t n_half[kHalfTileCount]       = {};
    SmoothLists full;
    int launches() const {
        int n = full.launches();
        for (const uint32_t k : n_half) n += k > 0;
        return n;
    }
};

// only: when given, the simulations to include (the others get no items).
template <class SimOf, class HalfOf>
bool pass_lists(Batch& B, const size_t n_desc, SimOf sim_of, HalfOf half_of, PassLists* out, const std::vector<char>* only = nullptr) {
    auto& S      = B.sims;
    auto include = [&](size_t d) { return only == nullptr || (*only)[sim_of(d)] != 0; };
    if (!smooth_lists(B, n_desc, [&](size_t d) { return S[sim_of(d)]->dims.nx; }, [&](size_t d) { return (!include(d) || half_of(d)) ? 0u : smooth_blocks(plan_tiles(S[sim_of(d)]->dims, B.zc[sim_of(d)])); }, &out->full)) return false;
    for (int t = 0; t < kHalfTileCount; ++t) {
        std::vector<size_t> members;
        for (size_t d = 0; d < n_desc; ++d)
            if (include(d) && half_of(d) && B.tile[sim_of(d)] == t) members.push_back(d);
        std::stable_sort(members.begin(), members.end(), [&](size_t a, size_t b) { return batch_half_plan(B, sim_of(a)).zc > batch_half_plan(B, sim_of(b)).zc; });
        std::vector<BatchItem> logical;
        for (const size_t d : members)
            for (uint32_t b = 0, n = smooth_blocks(batch_half_plan(B, sim_of(d))); b < n; ++b) logical.push_back(BatchItem{static_cast<uint32_t>(d), b});
        if (logical.empty()) continue;
        const auto items = xcd_order(logical);
        out->half[t]     = B.upload(items);
        out->n_half[t]   = static_cast<uint32_t>(items.size());
        if (!out->half[t]) return false;
    }
    return true;
}

// Pass `pass` (1..passes) of a batched solve: the black-cell lists read descriptors hd,
// the full-iterate lists od; the first pass starts from zero when zero_first.
hipError_t launch_pass_lists(hipStream_t s, const SmootherKind sm, const SolveKind solve, const int pass, const int passes, const bool zero_first, const HalfPassArgs* hd, const SmoothDesc* od, const PassLists& L) {
    const int in  = pass > 1 ? kHalfInHalf : (zero_first ? kHalfInZero : kHalfInNatural);
    const int out = pass == passes ? kHalfOutNatural : kHalfOutHalf;
    for (int t = 0; t < kHalfTileCount; ++t)
        if (L.n_half[t] > 0) {
            const hipError_t e = launch_rbgs_half_batched(s, solve, in, out, t, hd, L.half[t], L.n_half[t]);
            if (e != hipSuccess) return e;
        }
    return launch_smooth_lists(s, sm, solve, zero_first && pass == 1, od, L.full);
}

// Descriptor of pass `pass` (1..passes) of simulation i's black-cell solve.
HalfPassArgs batch_half_desc(const Batch& B, const size_t i, const int comp, const int pass, const int passes, const float* b, const float* x0, float* out, const float coef, const float last_scale, const StencilBC& bc) {
    const Context& c = *B.sims[i];
    const auto h     = batch_half_pair(B, i, comp);
    HalfPassArgs a{};
    a.x_nat     = x0;
    a.b         = b;
    a.h_in      = h[pass & 1]; // pass l = pass - 1 writes h[l & 1] and reads h[(l + 1) & 1]
    a.h_out     = h[(pass - 1) & 1];
    a.x_out     = out;
    a.coef      = coef;
    a.out_scale = pass == passes ? last_scale : 1.0f;
    a.d         = c.dims;
    a.bc        = bc;
    a.g         = batch_half_plan(B, i);
    a.L         = half_layout(c.dims);
    return a;
}

#define SF_BTRY(expr)                                           \
    do {                                                        \
        if ((expr) != hipSuccess) return hipErrorUnknown;       \
    } while (0)

// The step of every simulation (enqueue_step order), one launch per stage: uploads every
// descriptor table and work list, and records the launches in B.plan (captured later:
// no allocation or copy may happen during stream capture).
hipError_t plan_batch_step(Batch& B) {
    auto& S        = B.sims;
    const size_t N = S.size();
    const auto& c0 = *S[0];
    const int D = c0.config.diffuse_iterations, P = c0.config.pressure_iterations;
    const SmootherKind sm = c0.options.smoother;
    B.launches = 0;
    B.plan.clear();
    std::vector<TileGeom> geo(N);
    for (size_t i = 0; i < N; ++i) geo[i] = plan_tiles(S[i]->dims, B.zc[i]);

    const auto elem_list  = batch_items(N, [&](size_t d) { return elem_block_count(S[d]->cells); });
    const auto tile_list  = xcd_order(batch_items(N, [&](size_t d) { return z_march_blocks(S[d]->dims); })); // divergence, projection
    const auto adv_list   = xcd_order(batch_items(N, [&](size_t d) { return advect_blocks(S[d]->dims); }));
    const auto elem_items = B.upload(elem_list);
    const auto tile_items = B.upload(tile_list);
    const auto adv_items  = B.upload(adv_list);
    const auto n_elem     = static_cast<uint32_t>(elem_list.size());
    const auto n_tile     = static_cast<uint32_t>(tile_list.size());
    const auto n_adv      = static_cast<uint32_t>(adv_list.size());
    if (!elem_items || !tile_items || !adv_items) return hipErrorOutOfMemory;
    // solve families: velocity (descriptor 3 i + comp), pressure, and each scalar
    auto vel_bc  = [&](size_t d) { return velocity_rules(S[d / 3]->config.flow_boundary, static_cast<int>(d % 3)); };
    auto vel_half = [&](size_t d) { return batch_half(B, d / 3, vel_bc(d), static_cast<int>(d % 3)); };
    auto p_half   = [&](size_t d) { return batch_half(B, d, pressure_rules(S[d]->config.flow_boundary), -1); };
    PassLists lv, lp;
    if (!pass_lists(B, 3 * N, [](size_t d) { return d / 3; }, vel_half, &lv)) return hipErrorOutOfMemory;
    if (!pass_lists(B, N, [](size_t d) { return d; }, p_half, &lp)) return hipErrorOutOfMemory;
    std::vector<PassLists> ls(c0.scalars.size());
    for (size_t si = 0; si < ls.size(); ++si)
        if (!pass_lists(B, N, [](size_t d) { return d; }, [&](size_t d) { return batch_half(B, d, scalar_rules(S[d]->scalars[si].desc.boundary), -1); }, &ls[si])) return hipErrorOutOfMemory;
    hipStream_t st = B.stream;
    auto add       = [&](std::function<hipError_t()> f, const int launches = 1) {
        B.plan.push_back(std::move(f));
        B.launches += launches;
        return hipSuccess;
    };

    // forces (reference :612-617)
    for (size_t v = 0; v < c0.vectors.size(); ++v) {
        std::vector<AddForceDesc> d(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c = *S[i];
            d[i]    = AddForceDesc{{c.flow.vx, c.flow.vy, c.flow.vz}, {c.vectors[v].x, c.vectors[v].y, c.vectors[v].z}, c.config.dt, c.cells};
        }
        const auto dd = B.upload(d);
        if (!dd) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_add_force_batched(st, dd, elem_items, n_elem); }));
    }
    {
        std::vector<AdvectVelocityDesc> d(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c = *S[i];
            auto& f = c.flow;
            const float h = c.config.cell_size;
            d[i]          = AdvectVelocityDesc{{f.tx, f.ty, f.tz}, {f.vx, f.vy, f.vz}, c.config.dt, c.dims, h, c.config.flow_boundary, 1.0f / h, advect_pow2(h) ? 1 : 0};
        }
        const auto dd = B.upload(d);
        if (!dd) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_advect_velocity_batched(st, dd, adv_items, n_adv); }));
    }
    // velocity diffusion: pass i of all three components of all simulations per launch
    for (int pass = 1; pass <= D; ++pass) {
        std::vector<SmoothDesc> d(3 * N);
        std::vector<HalfPassArgs> h(3 * N);
        for (size_t i = 0; i < N; ++i) {
            auto& c           = *S[i];
            auto& f           = c.flow;
            const float alpha = c.config.dt * c.config.viscosity / (c.config.cell_size * c.config.cell_size);
            float* t[3] = {f.tx, f.ty, f.tz};
            float* a[3] = {f.ax, f.ay, f.az};
            float* v[3] = {f.vx, f.vy, f.vz};
            for (int comp = 0; comp < 3; ++comp) {
                const size_t k      = 3 * i + comp;
                const StencilBC bc  = velocity_rules(c.config.flow_boundary, comp);
                if (vel_half(k)) {
                    h[k] = batch_half_desc(B, i, comp, pass, D, t[comp], t[comp], a[comp], alpha, 1.0f, bc);
                    continue;
                }
                float* dst       = ((D - pass) % 2 == 0) ? a[comp] : v[comp];
                const float* src = pass == 1 ? t[comp] : (((D - pass + 1) % 2 == 0) ? a[comp] : v[comp]);
                d[k]             = SmoothDesc{src, t[comp], dst, alpha, 1.0f, c.dims, bc, geo[i]};
            }
        }
        const auto dd = B.upload(d);
        const auto dh = B.upload(h);
        if (!dd || !dh) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_pass_lists(st, sm, kSolveDiffusion, pass, D, false, dh, dd, lv); }, lv.launches()));
    }
    {
        std::vector<DivergenceDesc> d(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c = *S[i];
            auto& f = c.flow;
            d[i]    = DivergenceDesc{f.rhs, {f.ax, f.ay, f.az}, c.dims, c.config.cell_size, divergence_rules(c.config.flow_boundary)};
        }
        const auto dd = B.upload(d);
        if (!dd) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_divergence_z_batched(st, dd, tile_items, n_tile); }));
    }
    for (int pass = 1; pass <= P; ++pass) {
        std::vector<SmoothDesc> d(N);
        std::vector<HalfPassArgs> h(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c            = *S[i];
            auto& f            = c.flow;
            const float coef   = c.config.cell_size * c.config.cell_size;
            const StencilBC bc = pressure_rules(c.config.flow_boundary);
            if (p_half(i)) {
                h[i] = batch_half_desc(B, i, -1, pass, P, f.rhs, nullptr, f.p, coef, 1.0f, bc);
                continue;
            }
            float* dst       = ((P - pass) % 2 == 0) ? f.p : f.p_aux;
            const float* src = pass == 1 ? nullptr : (((P - pass + 1) % 2 == 0) ? f.p : f.p_aux);
            d[i]             = SmoothDesc{src, f.rhs, dst, coef, 1.0f, c.dims, bc, geo[i]};
        }
        const auto dd = B.upload(d);
        const auto dh = B.upload(h);
        if (!dd || !dh) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_pass_lists(st, sm, kSolvePoisson, pass, P, true, dh, dd, lp); }, lp.launches()));
    }
    {
        std::vector<ProjectDesc> d(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c = *S[i];
            auto& f = c.flow;
            d[i]    = ProjectDesc{{f.vx, f.vy, f.vz}, f.div, f.mag, {f.ax, f.ay, f.az}, f.p, c.dims, c.config.cell_size, pressure_rules(c.config.flow_boundary), divergence_rules(c.config.flow_boundary)};
        }
        const auto dd = B.upload(d);
        if (!dd) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_project_z_batched(st, dd, tile_items, n_tile); }));
    }
    // scalars (reference :652-668)
    for (size_t si = 0; si < c0.scalars.size(); ++si) {
        std::vector<AddScaledDesc> da(N);
        std::vector<AdvectScalarDesc> dv(N);
        for (size_t i = 0; i < N; ++i) {
            auto& c = *S[i];
            auto& s = c.scalars[si];
            auto& f = c.flow;
            da[i]   = AddScaledDesc{s.data, s.source, c.config.dt, c.cells};
            const float h = c.config.cell_size;
            dv[i]         = AdvectScalarDesc{s.temp, s.data, {f.vx, f.vy, f.vz}, c.config.dt, c.dims, h, s.desc.boundary, c.config.flow_boundary, 1.0f / h, advect_pow2(h) ? 1 : 0};
        }
        const auto dda = B.upload(da);
        const auto ddv = B.upload(dv);
        if (!dda || !ddv) return hipErrorOutOfMemory;
        SF_BTRY(add([=] { return launch_add_scaled_batched(st, dda, elem_items, n_elem); }));
        SF_BTRY(add([=] { return launch_advect_scalar_batched(st, ddv, adv_items, n_adv); }));
        const PassLists& L = ls[si];
        for (int pass = 1; pass <= D; ++pass) {
            std::vector<SmoothDesc> d(N);
            std::vector<HalfPassArgs> h(N);
            for (size_t i = 0; i < N; ++i) {
                auto& c              = *S[i];
                auto& s              = c.scalars[si];
                const float alpha_s  = c.config.dt * s.desc.diffusion / (c.config.cell_size * c.config.cell_size);
                const bool dissipate = s.desc.dissipation > 0.0f;
                const float factor   = dissipate ? 1.0f / (1.0f + c.config.dt * s.desc.dissipation) : 1.0f;
                const StencilBC bc   = scalar_rules(s.desc.boundary);
                if (batch_half(B, i, bc, -1)) {
                    h[i] = batch_half_desc(B, i, -1, pass, D, s.temp, s.temp, s.data, alpha_s, factor, bc);
                    continue;
                }
                float* dst       = ((D - pass) % 2 == 0) ? s.data : s.aux;
                const float* src = pass == 1 ? s.temp : (((D - pass + 1) % 2 == 0) ? s.data : s.aux);
                d[i]             = SmoothDesc{src, s.temp, dst, alpha_s, pass == D ? factor : 1.0f, c.dims, bc, geo[i]};
            }
            const auto dd = B.upload(d);
            const auto dh = B.upload(h);
            if (!dd || !dh) return hipErrorOutOfMemory;
            SF_BTRY(add([=] { return launch_pass_lists(st, sm, kSolveDiffusion, pass, D, false, dh, dd, L); }, L.launches()));
        }
    }
    return hipSuccess;
}

#undef SF_BTRY

StableFluidsResult build_batch_graph(Batch& B) {
    B.release();
    const hipError_t planned = plan_batch_step(B);
    if (planned != hipSuccess) return planned == hipErrorOutOfMemory ? STABLE_FLUIDS_RESULT_OUT_OF_MEMORY : STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    hipError_t enq = hipSuccess;
    for (auto& launch : B.plan)
        if (enq == hipSuccess) enq = launch();
    hipGraph_t graph     = nullptr;
    const hipError_t end = hipStreamEndCapture(B.stream, &graph);
    if (enq != hipSuccess || end != hipSuccess || graph == nullptr) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    B.graph = graph;
    if (hipGraphInstantiate(&B.graph_exec, B.graph, nullptr, nullptr, 0) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return STABLE_FLUIDS_RESULT_OK;
}

// ---- tiling autotuner: per grid shape, the (tile shape, z chunk) of the fastest pass ----

// A tuned tiling for nx x ny x nz under a smoother; members = the simulations of that
// shape timed together (0 for a single context).
struct TilingEntry {
    int nx, ny, nz, smoother, members, tile, zc;
};
constexpr const char* kTilingCacheTag = "stable-fluids-tiling 2";

// Entries of a cache file; a missing file or one in another format gives none.
std::vector<TilingEntry> read_tiling_cache(const char* path) {
    std::vector<TilingEntry> out;
    if (path == nullptr) return out;
    FILE* f = std::fopen(path, "r");
    if (f == nullptr) return out;
    char tag[64] = {};
    if (std::fgets(tag, sizeof tag, f) != nullptr && std::strncmp(tag, kTilingCacheTag, std::strlen(kTilingCacheTag)) == 0) {
        TilingEntry e{};
        while (std::fscanf(f, "%d %d %d %d %d %d %d", &e.nx, &e.ny, &e.nz, &e.smoother, &e.members, &e.tile, &e.zc) == 7)
            if (e.tile >= 0 && e.tile < kHalfTileCount && e.zc >= 0 && e.zc <= e.nz) out.push_back(e);
    }
    std::fclose(f);
    return out;
}

void write_tiling_cache(const char* path, const std::vector<TilingEntry>& entries) {
    if (path == nullptr) return;
    if (FILE* f = std::fopen(path, "w")) {
        std::fprintf(f, "%s\n", kTilingCacheTag);
        for (const auto& e : entries) std::fprintf(f, "%d %d %d %d %d %d %d\n", e.nx, e.ny, e.nz, e.smoother, e.members, e.tile, e.zc);
        std::fclose(f);
    }
}

const TilingEntry* find_tiling(const std::vector<TilingEntry>& cache, const Dims d, const int smoother, const int members) {
    for (const auto& e : cache)
        if (e.nx == d.nx && e.ny == d.ny && e.nz == d.nz && e.smoother == smoother && e.members == members) return &e;
    return nullptr;
}

struct Tiling {
    int tile, zc;
};
constexpr int kTuneRounds = 3;

// Runs `timed` (mean ms of `passes` timed passes) until about 20 ms of GPU work have
// passed, so that clocks have ramped before candidates are compared.
template <class F>
void tune_warmup(F timed, const int passes) {
    double spent = 0.0;
    for (int i = 0; i < 200 && spent < 20.0; ++i) {
        const double t = timed();
        if (!(t > 0.0)) return;
        spent += t * passes;
    }
}
// Candidates: every black-cell tile shape with the default and 4..32-plane chunks, or for
// the full-iterate passes (whose tile is fixed) chunks of 2..64 planes.
std::vector<Tiling> tiling_candidates(const Dims d, const bool half, const int tile) {
    std::vector<Tiling> out;
    if (half) {
        for (int t = 0; t < kHalfTileCount; ++t)
            for (const int zc : {0, 4, 6, 8, 12, 16, 32})
                if (zc <= d.nz) out.push_back(Tiling{t, zc});
    } else {
        for (const int zc : {0, 2, 4, 8, 16, 32, 64})
            if (zc <= d.nz) out.push_back(Tiling{tile, zc});
    }
    return out;
}

// Mean time of `repeats` middle RBGS pressure passes of a single context (half[0] ->
// half[1], scratch between steps) under tile shape `tile` and z chunk zc.
double time_half_pass(Context& c, const int tile, const int zc, const int repeats) {
    HalfPassArgs a{};
    a.b      = c.flow.rhs;
    a.h_in   = c.half[0];
    a.h_out  = c.half[1];
    a.x_out  = c.flow.p_aux;
    a.coef   = c.config.cell_size * c.config.cell_size;
    a.out_scale = 1.0f;
    a.d      = c.dims;
    a.bc     = pressure_rules(c.config.flow_boundary);
    a.g      = half_plan(c.dims, tile, zc);
    auto pass = [&] { return launch_rbgs_half(c.stream, kSolvePoisson, kHalfInHalf, kHalfOutHalf, a, tile) == hipSuccess; };
    double ms = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool ok = pass() && pass();
        ok      = ok && hipEventRecord(e0, c.stream) == hipSuccess;
        for (int r = 0; ok && r < repeats; ++r) ok = pass();
        ok      = ok && hipEventRecord(e1, c.stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float t = 0.0f;
        if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = t / repeats;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return ms;
}

// Mean time of `repeats` rounds of batched middle smoother passes over the simulations
// marked in `only` under the batch's current tiling, in the step's mix: one velocity
// diffusion pass (three components per simulation) and three single-field passes (the
// step runs 38 of the first and 116 of the second per 40 diffusion and 80 pressure
// passes). Black-cell passes write half buffers, full-iterate passes p_aux and the
// diffused velocity a (all scratch between steps); views are not touched. The
// descriptor tables join the batch's allocations (freed by the next graph build).
double time_batch_smoother(Batch& B, const std::vector<char>& only, const int repeats) {
    auto& S        = B.sims;
    const size_t N = S.size();
    auto p_half    = [&](size_t d) { return batch_half(B, d, pressure_rules(S[d]->config.flow_boundary), -1); };
    auto vel_bc    = [&](size_t d) { return velocity_rules(S[d / 3]->config.flow_boundary, static_cast<int>(d % 3)); };
    auto vel_half  = [&](size_t d) { return batch_half(B, d / 3, vel_bc(d), static_cast<int>(d % 3)); };
    PassLists lp, lv;
    if (!pass_lists(B, N, [](size_t d) { return d; }, p_half, &lp, &only)) return -1.0;
    if (!pass_lists(B, 3 * N, [](size_t d) { return d / 3; }, vel_half, &lv, &only)) return -1.0;
    std::vector<SmoothDesc> dp(N), dv(3 * N);
    std::vector<HalfPassArgs> hp(N), hv(3 * N);
    for (size_t i = 0; i < N; ++i) {
        auto& c            = *S[i];
        auto& f            = c.flow;
        const float coef   = c.config.cell_size * c.config.cell_size;
        const StencilBC bc = pressure_rules(c.config.flow_boundary);
        if (p_half(i)) hp[i] = batch_half_desc(B, i, -1, 2, 3, f.rhs, nullptr, f.p_aux, coef, 1.0f, bc);
        else dp[i] = SmoothDesc{f.rhs, f.rhs, f.p_aux, coef, 1.0f, c.dims, bc, plan_tiles(c.dims, B.zc[i])};
        const float alpha = c.config.dt * c.config.viscosity / (c.config.cell_size * c.config.cell_size);
        float* t[3]       = {f.tx, f.ty, f.tz};
        float* a[3]       = {f.ax, f.ay, f.az};
        for (int comp = 0; comp < 3; ++comp) {
            const size_t k = 3 * i + comp;
            if (vel_half(k)) hv[k] = batch_half_desc(B, i, comp, 2, 3, t[comp], nullptr, a[comp], alpha, 1.0f, vel_bc(k));
            else dv[k] = SmoothDesc{t[comp], t[comp], a[comp], alpha, 1.0f, c.dims, vel_bc(k), plan_tiles(c.dims, B.zc[i])};
        }
    }
    const auto ddp = B.upload(dp);
    const auto dhp = B.upload(hp);
    const auto ddv = B.upload(dv);
    const auto dhv = B.upload(hv);
    if (!ddp || !dhp || !ddv || !dhv) return -1.0;
    const SmootherKind sm = S[0]->options.smoother;
    auto round            = [&] {
        bool ok = launch_pass_lists(B.stream, sm, kSolveDiffusion, 2, 3, false, dhv, ddv, lv) == hipSuccess;
        for (int r = 0; ok && r < 3; ++r) ok = launch_pass_lists(B.stream, sm, kSolvePoisson, 2, 3, false, dhp, ddp, lp) == hipSuccess;
        return ok;
    };
    double ms = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool ok = round();
        ok      = ok && hipEventRecord(e0, B.stream) == hipSuccess;
        for (int r = 0; ok && r < repeats; ++r) ok = round();
        ok      = ok && hipEventRecord(e1, B.stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float t = 0.0f;
        if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = t / repeats;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return ms;
}

// State a batched simulation carries from one step to the next or shows as a view (the
// velocity, pressure, divergence and |v| views and the scalar fields; batched simulations
// have constant forces and sources, and every other array is rewritten each step before
// it is read), copied aside while the tuner times real steps.
struct BatchState {
    void* buf = nullptr;
    std::vector<std::pair<float*, size_t>> arrays; // (array, floats)
    ~BatchState() {
        if (buf) (void)hipFree(buf);
    }
};

std::vector<std::pair<float*, size_t>> batch_state_arrays(const Batch& B) {
    std::vector<std::pair<float*, size_t>> out;
    for (const Context* c : B.sims) {
        for (float* p : {c->flow.vx, c->flow.vy, c->flow.vz, c->flow.p, c->flow.div, c->flow.mag}) out.emplace_back(p, c->cells);
        for (const auto& s : c->scalars) out.emplace_back(s.data, c->cells);
    }
    return out;
}

bool copy_batch_state(Batch& B, BatchState& st, const bool save) {
    size_t off = 0;
    for (const auto& [p, n] : st.arrays) {
        float* mine = static_cast<float*>(st.buf) + off;
        if (hipMemcpyAsync(save ? mine : p, save ? p : mine, n * sizeof(float), hipMemcpyDeviceToDevice, B.stream) != hipSuccess) return false;
        off += n;
    }
    return hipStreamSynchronize(B.stream) == hipSuccess;
}

bool save_batch_state(Batch& B, BatchState& st) {
    for (const Context* c : B.sims)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
    st.arrays    = batch_state_arrays(B);
    size_t total = 0;
    for (const auto& a : st.arrays) total += a.second;
    if (hipMalloc(&st.buf, total * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        st.buf = nullptr;
        return false;
    }
    return copy_batch_state(B, st, true);
}

bool restore_batch_state(Batch& B, BatchState& st) { return copy_batch_state(B, st, false); }

// Mean GPU milliseconds of one batch step under the batch's current tiling: the graph
// rebuilt, one untimed step, then `steps` timed ones.
double time_batch_steps(Batch& B, const int steps) {
    if (build_batch_graph(B) != STABLE_FLUIDS_RESULT_OK) return -1.0;
    double ms     = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool ok = hipGraphLaunch(B.graph_exec, B.stream) == hipSuccess && hipEventRecord(e0, B.stream) == hipSuccess;
        for (int r = 0; ok && r < steps; ++r) ok = hipGraphLaunch(B.graph_exec, B.stream) == hipSuccess;
        ok      = ok && hipEventRecord(e1, B.stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float t = 0.0f;
        if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = t / steps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return ms;
}
} // namespace

} // namespace sf

using sf::Context;

extern "C" {

STABLE_FLUIDS_API StableFluidsResult stable_fluids_create_context_cuda(const StableFluidsContextCreateDesc* desc, void** out_context, StableFluidsScalarFieldHandle* out_s, uint32_t cap_s, StableFluidsVectorFieldHandle* out_v, uint32_t cap_v) {
    return sf::create(desc, nullptr, out_context, out_s, cap_s, out_v, cap_v);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_create_context_cuda(const StableFluidsContextCreateDesc* desc, const StableFluidsExtOptions* options, void** out_context, StableFluidsScalarFieldHandle* out_s, uint32_t cap_s, StableFluidsVectorFieldHandle* out_v, uint32_t cap_v) {
    return sf::create(desc, options, out_context, out_s, cap_s, out_v, cap_v);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_destroy_context_cuda(void* context) {
    sf::Range r("stable.destroy_context");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto* c = static_cast<Context*>(context);
    if (c->stream != nullptr) (void)hipStreamSynchronize(c->stream);
    delete c;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_scalar_field_cuda(void* context, StableFluidsScalarFieldHandle field, const float* values) {
    sf::Range r("stable.update_scalar_field");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (field == 0 || field > c.scalars.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return sf::status(sf::upload(c, c.scalars[field - 1u].data, values));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_scalar_field_source_cuda(void* context, StableFluidsScalarFieldHandle field, const float* values) {
    sf::Range r("stable.update_scalar_field_source");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (field == 0 || field > c.scalars.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return sf::status(sf::upload(c, c.scalars[field - 1u].source, values));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_vector_field_cuda(void* context, StableFluidsVectorFieldHandle field, const float* x, const float* y, const float* z) {
    sf::Range r("stable.update_vector_field");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (field == 0 || field > c.vectors.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& v = c.vectors[field - 1u];
    if (sf::upload(c, v.x, x) != hipSuccess || sf::upload(c, v.y, y) != hipSuccess || sf::upload(c, v.z, z) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_force_animation_cuda(void* context, StableFluidsVectorFieldHandle field, const StableFluidsExtForceAnimation* animation, const float* drift, const float* swirl_x, const float* swirl_z) {
    sf::Range r("stable.ext.set_force_animation");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (field == 0 || field > c.vectors.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& v = c.vectors[field - 1u];
    if (v.desc.usage != STABLE_FLUIDS_VECTOR_FIELD_FORCE) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (animation != nullptr && (drift == nullptr || swirl_x == nullptr || swirl_z == nullptr)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (animation != nullptr) {
        if (v.drift == nullptr && (!c.alloc(v.drift, c.bytes) || !c.alloc(v.swirl_x, c.bytes) || !c.alloc(v.swirl_z, c.bytes))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (c.anim_step == nullptr && !c.alloc(c.anim_step, sizeof(uint32_t))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (hipMemcpy(v.drift, drift, c.bytes, hipMemcpyDeviceToDevice) != hipSuccess || hipMemcpy(v.swirl_x, swirl_x, c.bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
            hipMemcpy(v.swirl_z, swirl_z, c.bytes, hipMemcpyDeviceToDevice) != hipSuccess)
            return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        const uint32_t start = animation->start_step;
        if (hipMemcpy(c.anim_step, &start, sizeof(start), hipMemcpyHostToDevice) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        v.anim     = sf::ForceAnimation{animation->phase_rate, animation->drift_x_amplitude, animation->drift_z_amplitude, animation->drift_z_rate, animation->swirl_amplitude, animation->swirl_rate};
        v.animated = true;
    } else {
        v.animated = false;
    }
    // the step graph is captured with the animation kernels in it (or without them)
    return sf::rebuild_graph(c);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_force_animation_step_cuda(void* context, uint32_t* out_step) {
    if (context == nullptr || out_step == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (c.anim_step == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return sf::status(hipMemcpy(out_step, c.anim_step, sizeof(uint32_t), hipMemcpyDeviceToHost));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_step_cuda(void* context) {
    sf::Range r("stable.step");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    return sf::status(hipGraphLaunch(c.graph_exec, c.stream));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_get_view_cuda(void* context, const StableFluidsViewRequest* request, StableFluidsView* out_view) {
    sf::Range r("stable.get_view");
    if (context == nullptr || request == nullptr || out_view == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    StableFluidsView v{};
    v.layout             = STABLE_FLUIDS_VIEW_LAYOUT_F32_3D;
    v.nx                 = c.config.nx;
    v.ny                 = c.config.ny;
    v.nz                 = c.config.nz;
    v.row_stride_bytes   = static_cast<uint64_t>(c.config.nx) * sizeof(float);
    v.slice_stride_bytes = static_cast<uint64_t>(c.config.nx) * static_cast<uint64_t>(c.config.ny) * sizeof(float);
    *out_view            = v;
    switch (request->kind) {
    case STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA:
    case STABLE_FLUIDS_VIEW_SCALAR_FIELD_SOURCE: {
        if (request->scalar_field == 0 || request->scalar_field > c.scalars.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        const auto& s   = c.scalars[request->scalar_field - 1u];
        out_view->data0 = request->kind == STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA ? s.data : s.source;
        break;
    }
    case STABLE_FLUIDS_VIEW_VECTOR_FIELD: {
        if (request->vector_field == 0 || request->vector_field > c.vectors.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        const auto& vf   = c.vectors[request->vector_field - 1u];
        out_view->layout = STABLE_FLUIDS_VIEW_LAYOUT_F32_3D_SOA3;
        out_view->data0  = vf.x;
        out_view->data1  = vf.y;
        out_view->data2  = vf.z;
        break;
    }
    case STABLE_FLUIDS_VIEW_FLOW_VELOCITY:
        out_view->layout = STABLE_FLUIDS_VIEW_LAYOUT_F32_3D_SOA3;
        out_view->data0  = c.flow.vx;
        out_view->data1  = c.flow.vy;
        out_view->data2  = c.flow.vz;
        break;
    case STABLE_FLUIDS_VIEW_FLOW_VELOCITY_MAGNITUDE: out_view->data0 = c.flow.mag; break;
    case STABLE_FLUIDS_VIEW_FLOW_PRESSURE: out_view->data0 = c.flow.p; break;
    case STABLE_FLUIDS_VIEW_FLOW_DIVERGENCE: out_view->data0 = c.flow.div; break;
    default: return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    // hand-off to the consumer stream (reference :743-759)
    if (request->consumer_stream == nullptr) return STABLE_FLUIDS_RESULT_OK;
    auto consumer = static_cast<hipStream_t>(request->consumer_stream);
    if (consumer == c.stream) return STABLE_FLUIDS_RESULT_OK;
    hipEvent_t ready = nullptr;
    if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const bool ok = hipEventRecord(ready, c.stream) == hipSuccess && hipStreamWaitEvent(consumer, ready, 0) == hipSuccess;
    (void)hipEventDestroy(ready);
    return ok ? STABLE_FLUIDS_RESULT_OK : STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_flow_velocity_cuda(void* context, const float* x, const float* y, const float* z) {
    sf::Range r("stable.ext.set_flow_velocity");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (sf::upload(c, c.flow.vx, x) != hipSuccess || sf::upload(c, c.flow.vy, y) != hipSuccess || sf::upload(c, c.flow.vz, z) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_flow_pressure_cuda(void* context, const float* values) {
    sf::Range r("stable.ext.set_flow_pressure");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    return sf::status(sf::upload(c, c.flow.p, values));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_field_norms_cuda(void* context, StableFluidsExtNorms* out) {
    sf::Range r("stable.ext.field_norms");
    if (context == nullptr || out == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c      = *static_cast<Context*>(context);
    const int nb = static_cast<int>(std::min<uint64_t>(sf::kNormBlocks, (c.cells + 255) / 256));
    sf::k_norm_partials<<<nb, 256, 0, c.stream>>>(c.flow.div, c.flow.vx, c.flow.vy, c.flow.vz, c.flow.p, c.cells, c.norm_partials);
    sf::k_norm_final<<<1, 256, 0, c.stream>>>(c.norm_partials, nb);
    double host[4] = {};
    if (hipGetLastError() != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipMemcpyAsync(host, c.norm_partials + 4 * sf::kNormBlocks, sizeof(host), hipMemcpyDeviceToHost, c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    out->divergence_l2  = sqrt(host[0]);
    out->velocity_l2    = sqrt(host[1]);
    out->pressure_l2    = sqrt(host[2]);
    out->divergence_max = host[3];
    return STABLE_FLUIDS_RESULT_OK;
}

// The reduction of stable_fluids_ext_field_norms_cuda (same kernels, same workgroup count,
// so the same bits) over caller-given fields, with the partial sums in the static scratch
// array instead of a context buffer, and the result written to device memory on the
// caller's stream: no allocation, no synchronisation.
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_field_norms_noalloc_cuda(const float* divergence, const float* velocity_x, const float* velocity_y, const float* velocity_z, const float* pressure, uint64_t cells, void* stream, StableFluidsExtNorms* out_norms_device) {
    sf::Range r("stable.ext.field_norms_noalloc");
    if (divergence == nullptr || velocity_x == nullptr || velocity_y == nullptr || velocity_z == nullptr || pressure == nullptr || out_norms_device == nullptr || cells == 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const auto s = static_cast<hipStream_t>(stream);
    if (s != nullptr) {
        hipDevice_t stream_device = -1;
        if (hipStreamGetDevice(s, &stream_device) != hipSuccess || stream_device != device) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    // the grid-stride reduction caps its workgroups at kNormBlocks, so every cell count
    // fits the scratch; a partial count beyond it is refused all the same
    const int nb = static_cast<int>(std::min<uint64_t>(sf::kNormBlocks, (cells + 255) / 256));
    if (4 * nb + 4 > sf::kNormScratchDoubles) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    double* partials = nullptr;
    if (sf::norm_scratch(device, &partials) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    // every partial slot the final kernel reads is written by this call first: nothing
    // from an earlier call is used
    sf::k_norm_partials<<<nb, 256, 0, s>>>(divergence, velocity_x, velocity_y, velocity_z, pressure, cells, partials);
    sf::k_norm_final<<<1, 256, 0, s>>>(partials, nb);
    sf::k_norm_export<<<1, 64, 0, s>>>(partials + 4 * sf::kNormBlocks, out_norms_device);
    return hipGetLastError() == hipSuccess ? STABLE_FLUIDS_RESULT_OK : STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
}

STABLE_FLUIDS_API uint64_t stable_fluids_ext_stage_bytes_cuda(void* context, uint32_t stage) {
    if (context == nullptr) return 0;
    const auto& c = *static_cast<Context*>(context);
    switch (stage) {
    case STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS:
    case STABLE_FLUIDS_EXT_STAGE_VELOCITY_PASS: {
        const auto& fb = c.config.flow_boundary;
        const sf::StencilBC bc = stage == STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS ? sf::pressure_rules(fb) : sf::velocity_rules(fb, 0);
        if (sf::use_half(c, bc)) return 8ull * c.cells; // read b and the black cells, write the black cells
        return 12ull * c.cells;
    }
    case STABLE_FLUIDS_EXT_STAGE_FULL_PRESSURE_PASS:
    case STABLE_FLUIDS_EXT_STAGE_FULL_VELOCITY_PASS:
    case STABLE_FLUIDS_EXT_STAGE_JACOBI_PASS: return 12ull * c.cells;   // read x_old, b; write x_new
    case STABLE_FLUIDS_EXT_STAGE_ADVECT: return 24ull * c.cells;        // read v (3); write t (3)
    case STABLE_FLUIDS_EXT_STAGE_PROJECT: return 36ull * c.cells;       // read a (3), p; write v (3), div, |v|
    case STABLE_FLUIDS_EXT_STAGE_COPY: return 8ull * c.cells;           // read one field, write one
    default: return 0;
    }
}

STABLE_FLUIDS_API const float* stable_fluids_ext_get_pressure_rhs_cuda(void* context) {
    return context == nullptr ? nullptr : static_cast<Context*>(context)->flow.rhs;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_pressure_model_cuda(void* context, const float* params, uint64_t count) {
    sf::Range r("stable.ext.set_pressure_model");
    if (context == nullptr || params == nullptr || count != stable_fluids_unet_param_count()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (c.warm.net == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return sf::status(sf::unet::net_set_params(*c.warm.net, params));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_warm_start_stats_cuda(void* context, double* out3) {
    if (context == nullptr || out3 == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (c.warm.stats == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    // the one-launch fallback solve of the last step records a dependency wait that ran
    // past its bound (its result is then NaN): reported as a failure
    int sync[3] = {0, 0, 0};
    if (hipMemcpy(sync, c.warm.skip, sizeof(sync), hipMemcpyDeviceToHost) != hipSuccess || sync[2] != 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return sf::status(hipMemcpy(out3, c.warm.stats, 3 * sizeof(double), hipMemcpyDeviceToHost));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_training_pair_cuda(void* context, float* input, float* target, uint32_t coarsening, double rel_tol, uint32_t max_iterations, double* out_rel_residual, uint32_t* out_iterations) {
    sf::Range r("stable.ext.training_pair");
    if (context == nullptr || input == nullptr || target == nullptr || !(rel_tol > 0.0) || coarsening > 4) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c       = *static_cast<Context*>(context);
    const int fac = 1 << coarsening;
    if (c.config.nx % fac || c.config.ny % fac || c.config.nz % fac) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& g  = c.cg;
    if (g.x == nullptr) {
        const size_t db = c.cells * sizeof(double);
        if (!c.alloc(g.x, db) || !c.alloc(g.r, db) || !c.alloc(g.p, db) || !c.alloc(g.ap, db) || !c.alloc(g.partials, sizeof(double) * 2 * sf::warm::kPartialBlocks) || !c.alloc(g.scal, 8 * sizeof(double))) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
        if (hipMemsetAsync(g.x, 0, db, c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    const float h2       = c.config.cell_size * c.config.cell_size;
    const sf::StencilBC bc = sf::pressure_rules(c.config.flow_boundary);
    sf::warm::CgResult res{};
    const int max_it = static_cast<int>(std::min<uint32_t>(max_iterations, 1u << 30));
    if (sf::warm::cg_solve(c.stream, c.flow.rhs, h2, g.x, g.r, g.p, g.ap, g.partials, g.scal, c.dims, bc, rel_tol, max_it, &res) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (sf::warm::launch_pair(c.stream, c.flow.rhs, g.x, h2, input, target, g.partials, g.scal, c.dims, fac, bc) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (out_rel_residual) *out_rel_residual = res.rel_residual;
    if (out_iterations) *out_iterations = static_cast<uint32_t>(res.iterations);
    return STABLE_FLUIDS_RESULT_OK;
}


// ---- batched simulations (stable-fluids-3d-ext.h) ----

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_create_cuda(void* const* contexts, uint32_t count, void* stream, void** out_batch) {
    sf::Range r("stable.ext.batch_create");
    if (out_batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_batch = nullptr;
    if (contexts == nullptr || count == 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    std::unique_ptr<sf::Batch> b{new (std::nothrow) sf::Batch{}};
    if (!b) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
    for (uint32_t i = 0; i < count; ++i) {
        if (contexts[i] == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        auto* c = static_cast<Context*>(contexts[i]);
        for (auto* o : b->sims)
            if (o == c) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE; // a context appears once
        if (!sf::batch_supported(*c) || (i > 0 && !sf::batch_compatible(*b->sims[0], *c))) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        b->sims.push_back(c);
    }
    b->stream = static_cast<hipStream_t>(stream);
    if (b->stream == nullptr) {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        b->owns_stream = true;
    }
    // each simulation starts from its own smoother tiling (set_smoother_tiling)
    b->zc.resize(count);
    b->tile.resize(count);
    b->hv.assign(count, {nullptr, nullptr, nullptr, nullptr});
    for (uint32_t i = 0; i < count; ++i) {
        auto& c     = *b->sims[i];
        b->zc[i]    = c.half_zc;
        b->tile[i]  = c.half_tile;
        if (c.half[0] == nullptr || c.options.smoother != sf::kSmoothRBGS || !sf::rbgs_half_supported(c.dims, sf::velocity_rules(c.config.flow_boundary, 0))) continue;
        const uint64_t n = sf::half_buffer_floats(c.dims);
        for (float*& h : b->hv[i]) {
            void* raw = nullptr;
            if (hipMalloc(&raw, sizeof(float) * n) != hipSuccess) {
                (void)hipGetLastError();
                return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
            }
            b->owned.push_back(raw);
            h = static_cast<float*>(raw);
            if (sf::launch_fill(b->stream, h, 0.0f, n) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        }
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    for (auto* c : b->sims)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const StableFluidsResult g = sf::build_batch_graph(*b);
    if (g != STABLE_FLUIDS_RESULT_OK) return g;
    *out_batch = b.release();
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_destroy_cuda(void* batch) {
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto* b = static_cast<sf::Batch*>(batch);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_step_cuda(void* batch) {
    sf::Range r("stable.ext.batch_step");
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    return sf::status(hipGraphLaunch(b.graph_exec, b.stream));
}

STABLE_FLUIDS_API void* stable_fluids_ext_batch_get_stream_cuda(void* batch) { return batch == nullptr ? nullptr : static_cast<sf::Batch*>(batch)->stream; }

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_set_z_chunk_cuda(void* batch, uint32_t sim, uint32_t z_chunk) {
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    if (sim >= b.sims.size() || z_chunk > static_cast<uint32_t>(b.sims[sim]->dims.nz)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(b.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    b.zc[sim] = static_cast<int>(z_chunk);
    return sf::build_batch_graph(b);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_set_tiling_cuda(void* batch, uint32_t sim, uint32_t tile, uint32_t z_chunk) {
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    if (sim >= b.sims.size() || tile >= sf::kHalfTileCount || z_chunk > static_cast<uint32_t>(b.sims[sim]->dims.nz)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (hipStreamSynchronize(b.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    b.tile[sim] = static_cast<int>(tile);
    b.zc[sim]   = static_cast<int>(z_chunk);
    return sf::build_batch_graph(b);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_get_tile_cuda(void* batch, uint32_t sim, uint32_t* out_tile) {
    if (batch == nullptr || out_tile == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    if (sim >= b.sims.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_tile = static_cast<uint32_t>(b.tile[sim]);
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_info_cuda(void* batch, uint32_t sim, int32_t* out_tiling4, uint32_t* out_launches_per_step) {
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    if (sim >= b.sims.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const bool half      = sf::batch_half(b, sim, sf::pressure_rules(b.sims[sim]->config.flow_boundary), -1);
    const sf::TileGeom g = half ? sf::batch_half_plan(b, sim) : sf::plan_tiles(b.sims[sim]->dims, b.zc[sim]);
    if (out_tiling4) {
        out_tiling4[0] = g.tiles_x;
        out_tiling4[1] = g.tiles_y;
        out_tiling4[2] = g.chunks;
        out_tiling4[3] = g.zc;
    }
    if (out_launches_per_step) *out_launches_per_step = static_cast<uint32_t>(b.launches);
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_autotune_cuda(void* batch, const char* cache_path, uint32_t repeats, uint32_t* out_tuned_shapes) {
    sf::Range r("stable.ext.batch_autotune");
    if (batch == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& b = *static_cast<sf::Batch*>(batch);
    if (hipStreamSynchronize(b.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const int reps = static_cast<int>(std::min<uint32_t>(std::max<uint32_t>(repeats, 1u), 1000u));
    const int sm   = static_cast<int>(b.sims[0]->options.smoother);
    const size_t N = b.sims.size();
    auto cache     = sf::read_tiling_cache(cache_path);
    auto same      = [](const sf::Dims& a, const sf::Dims& d) { return a.nx == d.nx && a.ny == d.ny && a.nz == d.nz; };
    // classes: the simulations of one grid shape, in first-appearance order
    struct Class {
        sf::Dims d;
        std::vector<char> only;
        int members = 0, first = -1;
        bool half = false, cached = false;
        std::vector<sf::Tiling> cands;
        std::vector<double> t; // fastest pass time per candidate, this class alone
    };
    std::vector<Class> classes;
    for (size_t i = 0; i < N; ++i) {
        const sf::Dims d = b.sims[i]->dims;
        Class* k         = nullptr;
        for (auto& c : classes)
            if (same(c.d, d)) k = &c;
        if (k == nullptr) {
            classes.push_back(Class{});
            k        = &classes.back();
            k->d     = d;
            k->only.assign(N, 0);
            k->first = static_cast<int>(i);
            k->half  = sf::batch_half(b, i, sf::pressure_rules(b.sims[i]->config.flow_boundary), -1);
        }
        k->only[i] = 1;
        ++k->members;
    }
    auto assign = [&](const Class& k, const sf::Tiling& t) {
        for (size_t i = 0; i < N; ++i)
            if (k.only[i]) {
                b.tile[i] = t.tile;
                b.zc[i]   = t.zc;
            }
    };
    auto effective = [&](const Class& k, const sf::Tiling& t) { return sf::Tiling{t.tile, k.half ? sf::half_plan(k.d, t.tile, t.zc).zc : sf::plan_tiles(k.d, t.zc).zc}; };
    const std::vector<int> zc0 = b.zc, tile0 = b.tile;
    const std::vector<char> all(N, 1);
    uint32_t tuned = 0;
    bool warm      = false;
    for (auto& k : classes) {
        if (const sf::TilingEntry* e = sf::find_tiling(cache, k.d, sm, k.members)) {
            k.cached = true;
            k.cands  = {sf::Tiling{e->tile, e->zc}};
            k.t      = {0.0};
            assign(k, k.cands[0]);
            continue;
        }
        if (!warm) sf::tune_warmup([&] { return sf::time_batch_smoother(b, all, reps); }, reps);
        warm = true;
        // every (tile shape, z chunk) candidate timed on this class alone: the minimum of
        // kTuneRounds interleaved rounds, so clock ramps and stray work hit every candidate
        k.cands = sf::tiling_candidates(k.d, k.half, b.tile[k.first]);
        k.t.assign(k.cands.size(), 1e300);
        for (int round = 0; round < sf::kTuneRounds; ++round)
            for (size_t c = 0; c < k.cands.size(); ++c) {
                assign(k, k.cands[c]);
                const double t = sf::time_batch_smoother(b, k.only, reps);
                if (t > 0.0) k.t[c] = std::min(k.t[c], t);
            }
        if (*std::min_element(k.t.begin(), k.t.end()) > 1e299) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        ++tuned;
    }
    if (tuned > 0) {
        // whole-batch plans (launches of different tile shapes run one after another):
        // every class on its own fastest tiling; every black-cell class on one tile shape
        // T, each with its fastest chunk under T, its default chunk or 16 planes; the
        // starting tiling.
        auto fastest = [](const Class& k, const int tile) {
            size_t best = k.cands.size();
            for (size_t c = 0; c < k.cands.size(); ++c)
                if ((tile < 0 || k.cands[c].tile == tile) && (best == k.cands.size() || k.t[c] < k.t[best])) best = c;
            if (best == k.cands.size())
                for (size_t c = 0; c < k.cands.size(); ++c)
                    if (best == k.cands.size() || k.t[c] < k.t[best]) best = c;
            return k.cands[best];
        };
        std::vector<std::pair<std::vector<int>, std::vector<int>>> plans; // (tile, zc) per simulation
        for (int T = -1; T < sf::kHalfTileCount; ++T) {
            for (const auto& k : classes) assign(k, fastest(k, (k.half && !k.cached) ? T : -1));
            plans.emplace_back(b.tile, b.zc);
            if (T < 0) continue;
            // the same tile shape with every black-cell class on its default chunk, and on
            // 16-plane chunks: chunks picked on a class alone can leave the whole batch
            // waiting on long workgroups
            for (const int zc : {0, 16}) {
                for (const auto& k : classes)
                    if (k.half && !k.cached) assign(k, sf::Tiling{T, zc});
                plans.emplace_back(b.tile, b.zc);
            }
        }
        plans.emplace_back(tile0, zc0);
        {
            std::vector<std::pair<std::vector<int>, std::vector<int>>> distinct;
            for (const auto& q : plans)
                if (std::find(distinct.begin(), distinct.end(), q) == distinct.end()) distinct.push_back(q);
            plans.swap(distinct);
        }
        // The plans and then each timed class's z chunk are decided on the batch's real
        // step (its graph rebuilt per plan, two rounds of steps, minimum): the smoother
        // proxy above ranks single classes well but misranked whole plans in about one
        // tuning in three (the batch then ran 10% slower). The simulations' state is
        // saved first and restored after, so tuning changes no result. Without memory for
        // the copy, the proxy decides.
        sf::BatchState saved;
        const bool real = sf::save_batch_state(b, saved);
        auto timed = [&]() {
            double t = 1e300;
            for (int round = 0; round < 2; ++round) {
                const double r = real ? sf::time_batch_steps(b, 2) : sf::time_batch_smoother(b, all, reps);
                if (r > 0.0) t = std::min(t, r);
            }
            return t;
        };
        double t_best = 1e300;
        size_t q_best = 0;
        for (size_t q = 0; q < plans.size(); ++q) {
            b.tile         = plans[q].first;
            b.zc           = plans[q].second;
            const double t = timed();
            if (t < t_best) {
                t_best = t;
                q_best = q;
            }
        }
        b.tile = plans[q_best].first;
        b.zc   = plans[q_best].second;
        // each timed class's z chunk on the whole batch, its tile shape kept: a chunk length
        // that ties when the class runs alone can leave the batch waiting on its longest
        // workgroups
        for (const auto& k : classes) {
            if (k.cached) continue;
            const int tile = b.tile[k.first], zc_now = b.zc[k.first];
            int zc_best    = zc_now;
            for (const auto& cand : k.cands) {
                if (cand.tile != tile || cand.zc == zc_now) continue;
                assign(k, cand);
                const double t = timed();
                if (t < t_best) {
                    t_best  = t;
                    zc_best = cand.zc;
                }
            }
            assign(k, sf::Tiling{tile, zc_best});
        }
        if (real && !sf::restore_batch_state(b, saved)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        for (const auto& k : classes) {
            if (k.cached) continue;
            const sf::Tiling t = effective(k, sf::Tiling{b.tile[k.first], b.zc[k.first]});
            cache.push_back(sf::TilingEntry{k.d.nx, k.d.ny, k.d.nz, sm, k.members, t.tile, t.zc});
        }
    }
    sf::write_tiling_cache(cache_path, cache);
    if (out_tuned_shapes) *out_tuned_shapes = tuned;
    return sf::build_batch_graph(b);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_run_stage_cuda(void* context, uint32_t stage, uint32_t repeat) {
    sf::Range r("stable.ext.run_stage");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c        = *static_cast<Context*>(context);
    auto& f        = c.flow;
    const auto& fb = c.config.flow_boundary;
    const float h  = c.config.cell_size;
    for (uint32_t i = 0; i < repeat; ++i) {
        hipError_t e = hipSuccess;
        switch (stage) {
        case STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS:
        case STABLE_FLUIDS_EXT_STAGE_VELOCITY_PASS: {
            // the pass the step runs: a middle black-cell pass where the solve uses them
            const bool pressure = stage == STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS;
            const sf::StencilBC bc = pressure ? sf::pressure_rules(fb) : sf::velocity_rules(fb, 0);
            const float coef       = pressure ? h * h : c.config.dt * c.config.viscosity / (h * h);
            if (sf::use_half(c, bc)) {
                sf::HalfPassArgs a{};
                a.b     = pressure ? f.rhs : f.tx;
                a.coef  = coef;
                a.out_scale = 1.0f;
                a.d     = c.dims;
                a.bc    = bc;
                a.g     = sf::half_plan(c.dims, c.half_tile, c.half_zc);
                a.h_in  = c.half[i & 1];
                a.h_out = c.half[(i + 1) & 1];
                e       = sf::launch_rbgs_half(c.stream, pressure ? sf::kSolvePoisson : sf::kSolveDiffusion, sf::kHalfInHalf, sf::kHalfOutHalf, a, c.half_tile);
                break;
            }
            [[fallthrough]];
        }
        case STABLE_FLUIDS_EXT_STAGE_FULL_PRESSURE_PASS:
        case STABLE_FLUIDS_EXT_STAGE_FULL_VELOCITY_PASS: {
            const bool odd = i & 1;
            if (stage == STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS || stage == STABLE_FLUIDS_EXT_STAGE_FULL_PRESSURE_PASS) {
                e = sf::launch_smooth(c.stream, c.options.smoother, sf::kSolvePoisson, sf::SmoothPass{odd ? f.p : f.p_aux, f.div, odd ? f.p_aux : f.p, h * h, 1.0f}, c.dims, sf::pressure_rules(fb));
            } else {
                const float alpha = c.config.dt * c.config.viscosity / (h * h);
                e = sf::launch_smooth(c.stream, c.options.smoother, sf::kSolveDiffusion, sf::SmoothPass{odd ? f.ax : f.vx, f.tx, odd ? f.vx : f.ax, alpha, 1.0f}, c.dims, sf::velocity_rules(fb, 0));
            }
            break;
        }
        case STABLE_FLUIDS_EXT_STAGE_JACOBI_PASS: {
            const bool odd = i & 1;
            e = sf::launch_smooth(c.stream, sf::kSmoothJacobi, sf::kSolvePoisson, sf::SmoothPass{odd ? f.p : f.p_aux, f.div, odd ? f.p_aux : f.p, h * h, 1.0f}, c.dims, sf::pressure_rules(fb));
            break;
        }
        case STABLE_FLUIDS_EXT_STAGE_PRESSURE_SOLVE: e = sf::enqueue_pressure_solve(c); break;
        case STABLE_FLUIDS_EXT_STAGE_PROJECTION: {
            const float* u[3];
            sf::diffused_velocity(c, u);
            e = sf::enqueue_projection(c, u);
            break;
        }
        case STABLE_FLUIDS_EXT_STAGE_COPY: e = sf::launch_copy_scaled(c.stream, f.ax, f.vx, 1.0f, c.cells); break;
        case STABLE_FLUIDS_EXT_STAGE_ADVECT: e = sf::launch_advect_velocity(c.stream, f.tx, f.ty, f.tz, f.vx, f.vy, f.vz, c.config.dt, c.dims, h, fb); break;
        case STABLE_FLUIDS_EXT_STAGE_PROJECT: e = sf::launch_project(c.stream, f.vx, f.vy, f.vz, f.div, f.mag, f.ax, f.ay, f.az, f.p, c.dims, h, sf::pressure_rules(fb), sf::divergence_rules(fb)); break;
        default: return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        }
        if (e != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_smoother_tiling_cuda(void* context, uint32_t tile, uint32_t z_chunk) {
    sf::Range r("stable.ext.set_smoother_tiling");
    if (context == nullptr || tile >= sf::kHalfTileCount || z_chunk > 4096) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    c.half_tile = static_cast<int>(tile);
    c.half_zc   = static_cast<int>(z_chunk);
    return sf::rebuild_graph(c);
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_get_smoother_tiling_cuda(void* context, uint32_t* out_tile, uint32_t* out_z_chunk) {
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const auto& c = *static_cast<Context*>(context);
    if (out_tile) *out_tile = static_cast<uint32_t>(c.half_tile);
    if (out_z_chunk) *out_z_chunk = static_cast<uint32_t>(sf::half_plan(c.dims, c.half_tile, c.half_zc).zc);
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_autotune_cuda(void* context, const char* cache_path, uint32_t repeats, uint32_t* out_tile, uint32_t* out_z_chunk) {
    sf::Range r("stable.ext.autotune");
    if (context == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& c = *static_cast<Context*>(context);
    if (hipStreamSynchronize(c.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    const int reps     = static_cast<int>(std::min<uint32_t>(std::max<uint32_t>(repeats, 1u), 1000u));
    const int sm       = static_cast<int>(c.options.smoother);
    const sf::StencilBC bc = sf::pressure_rules(c.config.flow_boundary);
    StableFluidsResult res = STABLE_FLUIDS_RESULT_OK;
    if (sf::use_half(c, bc)) {
        auto cache = sf::read_tiling_cache(cache_path);
        if (const sf::TilingEntry* e = sf::find_tiling(cache, c.dims, sm, 0)) {
            c.half_tile = e->tile;
            c.half_zc   = e->zc;
        } else {
            const auto cands = sf::tiling_candidates(c.dims, true, c.half_tile);
            std::vector<double> t_min(cands.size(), 1e300);
            sf::tune_warmup([&] { return sf::time_half_pass(c, c.half_tile, c.half_zc, reps); }, reps);
            for (int round = 0; round < sf::kTuneRounds; ++round)
                for (size_t k = 0; k < cands.size(); ++k) {
                    const double t = sf::time_half_pass(c, cands[k].tile, cands[k].zc, reps);
                    if (t > 0.0) t_min[k] = std::min(t_min[k], t);
                }
            const size_t best = static_cast<size_t>(std::min_element(t_min.begin(), t_min.end()) - t_min.begin());
            if (t_min[best] > 1e299) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            const int best_tile = cands[best].tile;
            const int best_zc   = sf::half_plan(c.dims, best_tile, cands[best].zc).zc;
            c.half_tile         = best_tile;
            c.half_zc           = best_zc;
            cache.push_back(sf::TilingEntry{c.dims.nx, c.dims.ny, c.dims.nz, sm, 0, best_tile, best_zc});
            sf::write_tiling_cache(cache_path, cache);
        }
        res = sf::rebuild_graph(c);
    }
    if (out_tile) *out_tile = static_cast<uint32_t>(c.half_tile);
    if (out_z_chunk) *out_z_chunk = static_cast<uint32_t>(sf::half_plan(c.dims, c.half_tile, c.half_zc).zc);
    return res;
}

STABLE_FLUIDS_API void* stable_fluids_ext_get_stream_cuda(void* context) {
    if (context == nullptr) return nullptr;
    return static_cast<Context*>(context)->stream;
}

} // extern "C"
