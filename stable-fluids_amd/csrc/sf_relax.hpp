// sf_relax.hpp -- arithmetic and work-mapping helpers shared by the smoother kernels
// (sf_smooth.hip, sf_rbgs_half.hip): the relaxation formulas with their correctly
// rounded reciprocal divisions, ghost rules in FMA form, the XCD-aware block order and
// the tile decomposition of a pass.
#pragma once

#include "sf_kernels.hpp"

namespace sf {

namespace {

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

// x / 6, correctly rounded, as the reciprocal product plus one FMA correction (Markstein).
// An exhaustive check over every finite float shows this equals the IEEE quotient for
// |x| >= 2^-125 (below that the residual underflows); tiny, infinite and NaN inputs take
// the division instruction sequence.
__device__ __forceinline__ float div6(const float x) {
    constexpr float r = 0x1.555556p-3f; // RN(1/6)
    const float q1    = x * r;
    const float e     = __builtin_fmaf(-q1, 6.0f, x);
    float q           = __builtin_fmaf(e, r, q1);
    if (__builtin_expect(!(__builtin_fabsf(x) >= 0x1p-110f && __builtin_fabsf(x) <= 0x1.fffffep127f), 0)) q = x / 6.0f;
    return q;
}

// Division by the diffusion denominator den = 1 + 6 coef, done the same way: x * RN(1/den)
// plus one FMA correction. By Markstein's theorem the corrected quotient is the IEEE one
// whenever the reciprocal is correctly rounded and nothing under- or overflows; an
// exhaustive check over every finite float x for eleven divisors in [1, 2^20] (among them
// all-ones mantissas) finds no difference for 2^-100 <= |x| <= 2^100. Other x, and
// divisors outside [1, 2^20], take the division instruction sequence.
struct DiffusionDiv {
    float den, inv;
    bool fast;
};

__device__ __forceinline__ DiffusionDiv diffusion_div(const float coef) {
    const float den = 1.0f + 6.0f * coef;
    return DiffusionDiv{den, 1.0f / den, den >= 1.0f && den <= 0x1p20f};
}

__device__ __forceinline__ float div_by(const DiffusionDiv& dv, const float x) {
    const float q1 = x * dv.inv;
    const float e  = __builtin_fmaf(-q1, dv.den, x);
    float q        = __builtin_fmaf(e, dv.inv, q1);
    const float ax = __builtin_fabsf(x);
    if (__builtin_expect(!(dv.fast && ax >= 0x1p-100f && ax <= 0x1p100f), 0)) q = x / dv.den;
    return q;
}

// relax_value with both divisions done by the corrected reciprocal (bit-identical); the
// divisor set-up is loop-invariant and hoisted out of the plane loops
template <int kSolve>
__device__ __forceinline__ float relax_fast(const float bv, const float sum, const float coef) {
    if constexpr (kSolve == kSolvePoisson)
        return div6(sum - coef * bv);
    else
        return div_by(diffusion_div(coef), bv + coef * sum);
}

// A ghost rule in the form g(v) = k ? c : fma(a, v, c): copy (1, -0), reflect (-1, value),
// const (value). fma(1, v, -0) = v and fma(-1, v, value) = value - v exactly, so this is
// apply_ghost bit for bit; (1, -0) is also the identity for cells that are not ghosts.
struct LinGhost {
    float a, c;
    bool k;
};

__device__ __forceinline__ LinGhost lin_identity() { return LinGhost{1.0f, -0.0f, false}; }

__device__ __forceinline__ LinGhost lin_ghost(const FaceRule f) {
    if (f.mode == kGhostCopy) return lin_identity();
    if (f.mode == kGhostReflect) return LinGhost{-1.0f, f.value, false};
    return LinGhost{0.0f, f.value, true};
}

__device__ __forceinline__ float lin_apply(const LinGhost g, const float v) { return g.k ? g.c : __builtin_fmaf(g.a, v, g.c); }

// Relaxation quotients of a stage with the range guard of div6 / div_by deferred: each
// quotient is taken by the corrected reciprocal (its sign copied from the numerator, which
// makes zeros exact), the numerator's |x| bit pattern is folded into a running max and
// min, and one test per stage sends the rare stage with an out-of-range numerator back
// through the exact quotient. Results are those of relax_fast bit for bit.
template <int kSolve>
struct StageDiv {
    // |x| in [2^-110, FLT_MAX] (x / 6) or [2^-100, 2^100] (x / den), or x = 0
    static constexpr uint32_t kLo = kSolve == kSolvePoisson ? 0x08800000u : 0x0d800000u;
    static constexpr uint32_t kHi = kSolve == kSolvePoisson ? 0x7f7fffffu : 0x71800000u;
    DiffusionDiv dv;
    uint32_t hi = 0u, lo = 0xffffffffu;
    __device__ __forceinline__ explicit StageDiv(const float coef) : dv(diffusion_div(coef)) {}
    __device__ __forceinline__ explicit StageDiv(const DiffusionDiv d) : dv(d) {}
    __device__ __forceinline__ float num(const float bv, const float sum, const float coef) const {
        if constexpr (kSolve == kSolvePoisson) return sum - coef * bv;
        else return bv + coef * sum;
    }
    __device__ __forceinline__ float quot(const float x) {
        const uint32_t bits = __float_as_uint(x) & 0x7fffffffu;
        hi                  = max(hi, bits);
        lo                  = min(lo, bits - 1u);
        float q;
        if constexpr (kSolve == kSolvePoisson) {
            constexpr float r = 0x1.555556p-3f;
            const float q1    = x * r;
            q                 = __builtin_fmaf(__builtin_fmaf(-q1, 6.0f, x), r, q1);
        } else {
            const float q1 = x * dv.inv;
            q              = __builtin_fmaf(__builtin_fmaf(-q1, dv.den, x), dv.inv, q1);
        }
        return __builtin_copysignf(q, x);
    }
    // true when some quotient of this lane needs the exact path
    __device__ __forceinline__ bool bad() const {
        const bool fast = kSolve == kSolvePoisson || dv.fast;
        return !fast || hi > kHi || lo < kLo - 1u;
    }
    __device__ __forceinline__ float exact(const float x) const {
        if constexpr (kSolve == kSolvePoisson) return x / 6.0f;
        else return x / dv.den;
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs; renumbering gives each XCD (and
// its L2) a contiguous run of chunks and tiles, so z-halo planes and x-y halo columns
// are mostly shared within one L2.
__device__ __forceinline__ int xcd_block() {
    const int n = static_cast<int>(gridDim.x);
    const int w = static_cast<int>(blockIdx.x);
    return (n & 7) == 0 ? (w & 7) * (n >> 3) + (w >> 3) : w;
}

// Block w of a pass -> (tile x, tile y, z chunk).
__device__ __forceinline__ void tile_of(const TileGeom g, const int w, int& tx, int& ty, int& cz) {
    cz          = w % g.chunks;
    const int t = w / g.chunks;
    tx          = t % g.tiles_x;
    ty          = t / g.tiles_x;
}

// FaceRule of face f (0..5) by a register select chain: indexing the kernel-argument
// struct with a runtime value would make the compiler copy it to scratch memory.
__device__ __forceinline__ FaceRule face_rule(const StencilBC& bc, const int f) {
    int mode    = bc.face[0].mode;
    float value = bc.face[0].value;
#pragma unroll
    for (int k = 1; k < 6; ++k)
        if (f == k) {
            mode  = bc.face[k].mode;
            value = bc.face[k].value;
        }
    return FaceRule{mode, value};
}

} // namespace

} // namespace sf
