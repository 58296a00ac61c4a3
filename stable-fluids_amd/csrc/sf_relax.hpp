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
    double rd;    // RN64(1 / den), for StageDiv
    uint32_t tlo; // StageDiv's key of den * 2^-126 (numerators below give subnormal quotients)
};

// StageDiv's key of |x|: (bits << 1) - 2, the sign shifted out; zeros map to the largest keys
__device__ __forceinline__ uint32_t stage_key(const float x) { return (__float_as_uint(x) << 1) - 2u; }

__device__ __forceinline__ DiffusionDiv diffusion_div(const float coef) {
    const float den = 1.0f + 6.0f * coef;
    const bool fast = den >= 1.0f && den <= 0x1p20f;
    // the reciprocal is wave-uniform: scalar registers hold it, not a vector register pair
    const uint64_t r  = __builtin_bit_cast(uint64_t, 1.0 / static_cast<double>(den));
    const uint32_t rl = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(r)));
    const uint32_t rh = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(r >> 32)));
    const double rd   = __builtin_bit_cast(double, (static_cast<uint64_t>(rh) << 32) | rl);
    return DiffusionDiv{den, 1.0f / den, fast, rd, static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(stage_key(den * 0x1p-126f)))};
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

// Relaxation quotients of a stage, taken as q = RN32(RN64(x * RN64(1/d))). For floats x and
// d, x/d is never a rounding boundary of the float grid (a midpoint between consecutive
// floats) in the normal range: d times a midpoint's 25-bit odd significand has more
// significant bits than a float. Its relative distance to such a boundary is at least
// 2^-49, while the float64 product is within 2^-52 of x/d, so q is the IEEE quotient for
// every result that is normal, zero or infinite. Subnormal results can sit on a midpoint of
// the subnormal grid: a nonzero numerator below d * 2^-126 is folded into a running
// minimum of keys, and one test per stage sends such a stage back through the exact
// quotient. Results are those of relax_fast bit for bit.
template <int kSolve>
struct StageDiv {
    // key of 6 * 2^-126 (x / 6 below 2^-126)
    static constexpr uint32_t kTlo6 = (__builtin_bit_cast(uint32_t, 6.0f * 0x1p-126f) << 1) - 2u;
    DiffusionDiv dv;
    uint32_t lo = 0xffffffffu;
    __device__ __forceinline__ explicit StageDiv(const float coef) : dv(diffusion_div(coef)) {}
    __device__ __forceinline__ explicit StageDiv(const DiffusionDiv d) : dv(d) {}
    __device__ __forceinline__ float num(const float bv, const float sum, const float coef) const {
        if constexpr (kSolve == kSolvePoisson) return sum - coef * bv;
        else return bv + coef * sum;
    }
    __device__ __forceinline__ float quot(const float x) {
        lo = min(lo, stage_key(x));
        if constexpr (kSolve == kSolvePoisson) return static_cast<float>(static_cast<double>(x) * (1.0 / 6.0));
        else return static_cast<float>(static_cast<double>(x) * dv.rd);
    }
    // true when some quotient of this lane needs the exact path
    __device__ __forceinline__ bool bad() const {
        if constexpr (kSolve == kSolvePoisson) return lo < kTlo6;
        else return !dv.fast || lo < dv.tlo;
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
