// sf_mixed.hip -- defect and finalize passes of the mixed-precision smoother mode.
//
// A solve of K smoother passes A x = rhs (diffusion or pressure Poisson, reference
// backend-cuda.cu:354-384 and loops :628-650) runs as ceil(K/m) refinement cycles:
//   defect:     x' = x + s*e (FP32), r = rhs - A x' with the full boundary rules (FP32),
//               max|r| into a per-cycle slot;
//   correction: up to m smoother passes on A e = r from e = 0 with homogeneous boundary
//               rules, e stored FP16 or BF16 scaled by 1/s, s = 2^ceil(log2 max|r|)
//               (FP32 storage: s = 1);
//   finalize:   out = (x + s*e) * out_scale.
// The smoother passes are affine in the iterate, so in exact arithmetic m correction
// passes from zero add exactly what m passes on x would add: the cycles reproduce the K
// plain passes, and the 16-bit rounding of e is corrected by the next FP32 defect.
#include "sf_kernels.hpp"
#include "sf_storage.hpp"

namespace sf {

namespace {

constexpr int kTileX     = 64;
constexpr int kTileY     = 4;
constexpr int kElemBlock = 256;

template <class St>
__device__ __forceinline__ float refined(const float* __restrict__ x, const typename St::T* __restrict__ e, const float s, const uint64_t i) {
    const float xv = x ? x[i] : 0.0f;
    return e ? xv + s * St::load(e[i]) : xv;
}

// Largest |v| over the block, one atomic per block (non-negative floats order as their
// bit patterns, so an unsigned max is a float max).
__device__ __forceinline__ void block_max_into(float v, float* slot) {
    __shared__ float s_max[kTileX * kTileY / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int t    = threadIdx.y * blockDim.x + threadIdx.x;
    const int wave = t >> 6;
    if ((t & 63) == 0) s_max[wave] = v;
    __syncthreads();
    if (t == 0) {
        float m = s_max[0];
        for (int w = 1; w < static_cast<int>(blockDim.x * blockDim.y) / 64; ++w) m = fmaxf(m, s_max[w]);
        atomicMax(reinterpret_cast<unsigned int*>(slot), __float_as_uint(m));
    }
}

template <int kSolve, class St>
__global__ __launch_bounds__(kTileX* kTileY) void k_defect(const float* __restrict__ x_in, const typename St::T* __restrict__ e, const float* __restrict__ e_slot, const float* __restrict__ b, const float coef, float* __restrict__ x_out, float* __restrict__ r, float* __restrict__ r_slot, const Dims d, const StencilBC bc) {
    const int x      = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int y      = static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    const int z      = static_cast<int>(blockIdx.z);
    const float s    = e ? correction_scale<St>(e_slot) : 0.0f;
    float local_max  = 0.0f;
    if (x < d.nx && y < d.ny) {
        const uint64_t i   = lin(x, y, z, d);
        const float centre = refined<St>(x_in, e, s, i);
        const int n[3]     = {d.nx, d.ny, d.nz};
        float sum          = 0.0f;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int axis = f >> 1;
            int q[3]       = {x, y, z};
            float v;
            if (step_axis(q[axis], (f & 1) ? 1 : -1, n[axis], (bc.periodic >> axis) & 1))
                v = refined<St>(x_in, e, s, lin(q[0], q[1], q[2], d));
            else
                v = apply_ghost(bc.face[f], centre);
            sum += v;
        }
        float res;
        if constexpr (kSolve == kSolvePoisson)
            res = -coef * b[i] - (6.0f * centre - sum);
        else
            res = b[i] - ((1.0f + 6.0f * coef) * centre - coef * sum);
        r[i] = res;
        if (e) x_out[i] = centre;
        local_max = fabsf(res);
    }
    block_max_into(local_max, r_slot);
}

template <class St>
__global__ __launch_bounds__(kElemBlock) void k_finalize(float* __restrict__ out, const float* __restrict__ x, const typename St::T* __restrict__ e, const float* __restrict__ e_slot, const float out_scale, const uint64_t n) {
    const float s = correction_scale<St>(e_slot);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const float xv = x ? x[i] : 0.0f;
        out[i]         = (xv + s * St::load(e[i])) * out_scale;
    }
}

template <class St>
hipError_t defect(hipStream_t s, const SolveKind solve, const float* x_in, const void* e, const float* e_slot, const float* b, const float coef, float* x_out, float* r, float* r_slot, const Dims d, const StencilBC& bc) {
    const dim3 grid(static_cast<unsigned>((d.nx + kTileX - 1) / kTileX), static_cast<unsigned>((d.ny + kTileY - 1) / kTileY), static_cast<unsigned>(d.nz));
    const dim3 block(kTileX, kTileY, 1);
    const auto* et = static_cast<const typename St::T*>(e);
    if (solve == kSolvePoisson)
        k_defect<kSolvePoisson, St><<<grid, block, 0, s>>>(x_in, et, e_slot, b, coef, x_out, r, r_slot, d, bc);
    else
        k_defect<kSolveDiffusion, St><<<grid, block, 0, s>>>(x_in, et, e_slot, b, coef, x_out, r, r_slot, d, bc);
    return hipGetLastError();
}

template <class St>
hipError_t finalize(hipStream_t s, float* out, const float* x, const void* e, const float* e_slot, const float out_scale, const uint64_t n) {
    const uint64_t want = (n + kElemBlock - 1) / kElemBlock;
    const unsigned grid = static_cast<unsigned>(want < 65536 ? want : 65536);
    k_finalize<St><<<grid, kElemBlock, 0, s>>>(out, x, static_cast<const typename St::T*>(e), e_slot, out_scale, n);
    return hipGetLastError();
}

} // namespace

hipError_t launch_defect(hipStream_t s, const SolveKind solve, const StorageKind storage, const float* x_in, const void* e, const float* e_slot, const float* b, const float coef, float* x_out, float* r, float* r_slot, const Dims d, const StencilBC& bc) {
    if (e != nullptr && (e_slot == nullptr || x_out == nullptr || x_out == x_in)) return hipErrorInvalidValue;
    switch (storage) {
    case kStoreF16: return defect<StoreF16>(s, solve, x_in, e, e_slot, b, coef, x_out, r, r_slot, d, bc);
    case kStoreBF16: return defect<StoreBF16>(s, solve, x_in, e, e_slot, b, coef, x_out, r, r_slot, d, bc);
    default: return defect<StoreF32>(s, solve, x_in, e, e_slot, b, coef, x_out, r, r_slot, d, bc);
    }
}

hipError_t launch_finalize(hipStream_t s, const StorageKind storage, float* out, const float* x, const void* e, const float* e_slot, const float out_scale, const uint64_t n) {
    if (n == 0) return hipSuccess;
    if (e == nullptr || e_slot == nullptr) return hipErrorInvalidValue;
    switch (storage) {
    case kStoreF16: return finalize<StoreF16>(s, out, x, e, e_slot, out_scale, n);
    case kStoreBF16: return finalize<StoreBF16>(s, out, x, e, e_slot, out_scale, n);
    default: return finalize<StoreF32>(s, out, x, e, e_slot, out_scale, n);
    }
}

} // namespace sf
