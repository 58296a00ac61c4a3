// sf_mixed.hip -- defect and finalize passes of the mixed-precision smoother mode.
//
// A solve of K smoother passes A x = rhs (diffusion or pressure Poisson, reference
// backend-cuda.cu:354-384 and loops :628-650) runs as ceil(K/m) refinement cycles:
//   defect:     x' = x + s*e (FP32), r = rhs - A x' with the full boundary rules (FP32),
//               max|r| into a per-cycle slot (per-workgroup maxima, then one fold);
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
constexpr int kDefectZ   = 8; // planes per defect workgroup
constexpr int kElemBlock = 256;
constexpr int kReduce    = 1024;

template <class St>
__device__ __forceinline__ float refined(const float* __restrict__ x, const typename St::T* __restrict__ e, const float s, const uint64_t i) {
    const float xv = x ? x[i] : 0.0f;
    return e ? xv + s * St::load(e[i]) : xv;
}

// Largest v over the workgroup, written by thread 0 to *out. fmaxf drops NaN operands,
// so a NaN defect does not hide the finite maximum.
template <int kThreadsPerGroup>
__device__ __forceinline__ void block_max_to(float v, float* out) {
    __shared__ float s_max[kThreadsPerGroup / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) s_max[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        float m = s_max[0];
#pragma unroll
        for (int w = 1; w < kThreadsPerGroup / 64; ++w) m = fmaxf(m, s_max[w]);
        *out = m;
    }
}

// One workgroup: a 64x4 (x, y) tile marching over kDefectZ planes (the z-1, z, z+1
// values rotate through registers); its max|r| goes to partials[workgroup] and
// k_partials_max folds them into the cycle's slot.
template <int kSolve, class St>
__global__ __launch_bounds__(kTileX* kTileY) void k_defect(const float* __restrict__ x_in, const typename St::T* __restrict__ e, const float* __restrict__ e_slot, const float* __restrict__ b, const float coef, float* __restrict__ x_out, float* __restrict__ r, float* __restrict__ partials, const Dims d, const StencilBC bc) {
    const int x     = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int y     = static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    const int z0    = static_cast<int>(blockIdx.z) * kDefectZ;
    const int z1    = min(z0 + kDefectZ, d.nz);
    const float s   = e ? correction_scale<St>(e_slot) : 0.0f;
    float local_max = 0.0f;
    if (x < d.nx && y < d.ny) {
        const bool px = bc.periodic & 1, py = (bc.periodic >> 1) & 1, pz = (bc.periodic >> 2) & 1;
        const uint64_t plane = static_cast<uint64_t>(d.nx) * d.ny;
        // in-plane neighbour offsets (or -1: ghost of that face)
        int xm = x - 1, xp = x + 1, ym = y - 1, yp = y + 1;
        const bool gxm = !step_axis(xm, 0, d.nx, px), gxp = !step_axis(xp, 0, d.nx, px);
        const bool gym = !step_axis(ym, 0, d.ny, py), gyp = !step_axis(yp, 0, d.ny, py);
        const uint32_t row = static_cast<uint32_t>(y) * d.nx;
        const uint32_t o_xm = row + (gxm ? x : xm), o_xp = row + (gxp ? x : xp);
        const uint32_t o_ym = static_cast<uint32_t>(gym ? y : ym) * d.nx + x, o_yp = static_cast<uint32_t>(gyp ? y : yp) * d.nx + x;
        auto at = [&](const int z, const uint32_t o) { return refined<St>(x_in, e, s, static_cast<uint64_t>(z) * plane + o); };
        const bool gzm0 = z0 == 0 && !pz;
        float vm        = gzm0 ? 0.0f : at(z0 == 0 ? d.nz - 1 : z0 - 1, row + x);
        bool gzm        = gzm0;
        float vc        = at(z0, row + x);
        for (int z = z0; z < z1; ++z) {
            const bool gzp = z + 1 >= d.nz && !pz;
            const float vp = gzp ? 0.0f : at(z + 1 >= d.nz ? 0 : z + 1, row + x);
            float sum      = gxm ? apply_ghost(bc.face[0], vc) : at(z, o_xm);
            sum += gxp ? apply_ghost(bc.face[1], vc) : at(z, o_xp);
            sum += gym ? apply_ghost(bc.face[2], vc) : at(z, o_ym);
            sum += gyp ? apply_ghost(bc.face[3], vc) : at(z, o_yp);
            sum += gzm ? apply_ghost(bc.face[4], vc) : vm;
            sum += gzp ? apply_ghost(bc.face[5], vc) : vp;
            const uint64_t i = static_cast<uint64_t>(z) * plane + row + x;
            float res;
            if constexpr (kSolve == kSolvePoisson)
                res = -coef * b[i] - (6.0f * vc - sum);
            else
                res = b[i] - ((1.0f + 6.0f * coef) * vc - coef * sum);
            r[i] = res;
            if (e) x_out[i] = vc;
            local_max = fmaxf(local_max, fabsf(res));
            vm  = vc;
            gzm = false;
            vc  = vp;
        }
    }
    block_max_to<kTileX * kTileY>(local_max, partials + (static_cast<uint64_t>(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
}

__global__ __launch_bounds__(kReduce) void k_partials_max(const float* __restrict__ partials, const uint32_t n, float* __restrict__ slot) {
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += kReduce) m = fmaxf(m, partials[i]);
    block_max_to<kReduce>(m, slot);
}

template <class St>
__global__ __launch_bounds__(kElemBlock) void k_finalize(float* __restrict__ out, const float* __restrict__ x, const typename St::T* __restrict__ e, const float* __restrict__ e_slot, const float out_scale, const uint64_t n) {
    const float s = correction_scale<St>(e_slot);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const float xv = x ? x[i] : 0.0f;
        out[i]         = (xv + s * St::load(e[i])) * out_scale;
    }
}

dim3 defect_grid(const Dims d) {
    return dim3(static_cast<unsigned>((d.nx + kTileX - 1) / kTileX), static_cast<unsigned>((d.ny + kTileY - 1) / kTileY), static_cast<unsigned>((d.nz + kDefectZ - 1) / kDefectZ));
}

template <class St>
hipError_t defect(hipStream_t s, const SolveKind solve, const float* x_in, const void* e, const float* e_slot, const float* b, const float coef, float* x_out, float* r, float* partials, float* r_slot, const Dims d, const StencilBC& bc) {
    const dim3 grid = defect_grid(d);
    const dim3 block(kTileX, kTileY, 1);
    const auto* et = static_cast<const typename St::T*>(e);
    if (solve == kSolvePoisson)
        k_defect<kSolvePoisson, St><<<grid, block, 0, s>>>(x_in, et, e_slot, b, coef, x_out, r, partials, d, bc);
    else
        k_defect<kSolveDiffusion, St><<<grid, block, 0, s>>>(x_in, et, e_slot, b, coef, x_out, r, partials, d, bc);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    k_partials_max<<<1, kReduce, 0, s>>>(partials, grid.x * grid.y * grid.z, r_slot);
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

uint64_t defect_partials(const Dims d) {
    const dim3 g = defect_grid(d);
    return static_cast<uint64_t>(g.x) * g.y * g.z;
}

hipError_t launch_defect(hipStream_t s, const SolveKind solve, const StorageKind storage, const float* x_in, const void* e, const float* e_slot, const float* b, const float coef, float* x_out, float* r, float* partials, float* r_slot, const Dims d, const StencilBC& bc) {
    if (e != nullptr && (e_slot == nullptr || x_out == nullptr || x_out == x_in)) return hipErrorInvalidValue;
    if (partials == nullptr || r_slot == nullptr || defect_partials(d) > 0xffffffffull) return hipErrorInvalidValue;
    switch (storage) {
    case kStoreF16: return defect<StoreF16>(s, solve, x_in, e, e_slot, b, coef, x_out, r, partials, r_slot, d, bc);
    case kStoreBF16: return defect<StoreBF16>(s, solve, x_in, e, e_slot, b, coef, x_out, r, partials, r_slot, d, bc);
    default: return defect<StoreF32>(s, solve, x_in, e, e_slot, b, coef, x_out, r, partials, r_slot, d, bc);
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
