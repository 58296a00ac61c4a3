// sf_storage.hpp -- element storage of smoother iterates: FP32, FP16 or BF16.
// Arithmetic is always FP32; 16-bit storage only rounds what is written to memory
// (round to nearest even).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sf {

struct StoreF32 {
    using T = float;
    static constexpr bool kNarrow = false;
    __device__ __forceinline__ static float load(const T v) { return v; }
    // aligned pair p[0], p[1]
    __device__ __forceinline__ static float2 load2(const T* p) { return *reinterpret_cast<const float2*>(p); }
    __device__ __forceinline__ static T store(const float v) { return v; }
};

struct StoreF16 {
    using T = _Float16;
    static constexpr bool kNarrow = true;
    __device__ __forceinline__ static float load(const T v) { return static_cast<float>(v); }
    __device__ __forceinline__ static float2 load2(const T* p) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
        return make_float2(static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu))), static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16))));
    }
    __device__ __forceinline__ static T store(const float v) { return static_cast<_Float16>(v); }
};

struct StoreBF16 {
    using T = uint16_t; // upper half of the FP32 bit pattern
    static constexpr bool kNarrow = true;
    __device__ __forceinline__ static float load(const T v) { return __uint_as_float(static_cast<uint32_t>(v) << 16); }
    __device__ __forceinline__ static float2 load2(const T* p) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    }
    __device__ __forceinline__ static T store(const float v) {
        const uint32_t u = __float_as_uint(v);
        if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<T>((u >> 16) | 0x40u); // quiet NaN
        return static_cast<T>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
};

// Power-of-two scale s >= m (1 for m = 0 or non-finite m); 1/s is exact.
__device__ __forceinline__ float pow2_at_least(const float m) {
    if (!(m > 0.0f) || !(m <= 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(m, &e); // m = f * 2^e, f in [0.5, 1)
    return ldexpf(1.0f, e);
}

// Scale s of a correction iterate held as e/s: the power of two at least max|r| (*slot)
// for 16-bit storage, 1 for FP32 storage.
template <class St>
__device__ __forceinline__ float correction_scale(const float* slot) {
    if constexpr (St::kNarrow) return pow2_at_least(*slot);
    else return 1.0f;
}

} // namespace sf
