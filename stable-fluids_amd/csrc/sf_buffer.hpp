// sf_buffer.hpp -- raw buffer access to one z-plane of a field (gfx9 buffer resources).
//
// A resource spans [base, base + bytes). A lane whose byte offset lies at or past `bytes`
// reads 0 and its store is dropped by the hardware, without touching memory, so the
// stencil kernels bound their lanes (and switch whole planes off by giving the resource
// 0 bytes) without branches around memory instructions. Branch-free memory keeps the
// compiler's wait counts exact, which is what lets loads stay in flight across planes.
// Offsets are 32-bit byte offsets within one plane, or within a whole field below 2 GiB
// when the plane is chosen by the scalar offset.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sf_storage.hpp"

namespace sf {

constexpr int kBufferConfig = 0x00020000; // descriptor dword 3 of a gfx9 raw (untyped) buffer
constexpr int kOutOfRange   = 0x7ffffff0; // byte offset past any plane
// Cache policy bit sc1 (device scope) of a buffer store. The smoother outputs, which the
// next pass reads from whichever XCD runs its workgroup, use it: measured 28.8 -> 28.0 us
// per black-cell RBGS pass and 33.2 -> 32.6 us per Jacobi sweep at 256^3 in the step.
constexpr int kDeviceScopeStore = 16;

using Rsrc = __amdgpu_buffer_rsrc_t;

// The base and size are wave-uniform by construction; passing them through readfirstlane
// keeps the descriptor in scalar registers even where the compiler has parked a uniform
// value in a vector register (a descriptor in vector registers turns every access into a
// loop over the lanes' distinct descriptors).
__device__ __forceinline__ Rsrc plane_rsrc(const void* base, const uint32_t bytes) {
    const uint64_t p  = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p >> 32));
    const uint32_t n  = __builtin_amdgcn_readfirstlane(bytes);
    void* q           = reinterpret_cast<void*>((static_cast<uint64_t>(hi) << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(q, 0, static_cast<int>(n), kBufferConfig);
}

// A wave-uniform pointer kept in scalar registers (see plane_rsrc), and a resource at it
// of `bytes` bytes: build the pointer once outside a loop, the resource per use.
__device__ __forceinline__ const void* uniform_ptr(const void* base) {
    const uint64_t p  = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p >> 32));
    return reinterpret_cast<const void*>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ Rsrc rsrc_at(const void* uniform, const uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(uniform), 0, static_cast<int>(__builtin_amdgcn_readfirstlane(bytes)), kBufferConfig);
}

// Aligned pair of elements at byte offset `off` (+ the wave-uniform `soff`).
template <class St>
__device__ __forceinline__ float2 buf_load2(const Rsrc r, const int off, const int soff = 0) {
    if constexpr (St::kNarrow) {
        const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0);
        if constexpr (__is_same(typename St::T, _Float16))
            return make_float2(static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu))), static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16))));
        else
            return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
        return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
    }
}

// One element at byte offset `off` (+ the wave-uniform `soff`).
template <class St>
__device__ __forceinline__ float buf_load1(const Rsrc r, const int off, const int soff = 0) {
    if constexpr (St::kNarrow) {
        const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, 0);
        return St::load(__builtin_bit_cast(typename St::T, h));
    } else {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    }
}

// Aligned pair (a, b) at byte offset `off`, rounded to the storage type. (Non-temporal
// stores, aux bit NT, measured slower for the smoother outputs: the next pass reads them.)
template <class St, int kPolicy = 0>
__device__ __forceinline__ void buf_store2(const Rsrc r, const int off, const float a, const float b, const int soff = 0) {
    if constexpr (St::kNarrow) {
        const uint32_t w = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, St::store(a))) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, St::store(b))) << 16);
        __builtin_amdgcn_raw_buffer_store_b32(w, r, off, soff, kPolicy);
    } else {
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 v = {__float_as_uint(a), __float_as_uint(b)};
        __builtin_amdgcn_raw_buffer_store_b64(v, r, off, soff, kPolicy);
    }
}

} // namespace sf
