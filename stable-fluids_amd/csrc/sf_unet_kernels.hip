// sf_unet_kernels.hip -- gfx950 kernels of the pressure warm-start network.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation). A workgroup (4 waves) owns a 16x4x2 voxel tile; the input tile with its
// 1-voxel halo (18x6x4 voxels x C channels, channels-last) is staged once in LDS and
// every wave multiplies two 16-voxel rows (M) by all output channels (N) over
// K = 27 taps x C input channels, 32 at a time: A fragments are 16-byte LDS reads of 8
// consecutive channels of one shifted voxel, B fragments 16-byte reads of the BF16
// weights [tap][co][ci]. The backward-data pass is the same kernel with flipped,
// transposed weights. The weight gradient is a second MFMA kernel with K = voxels.
//
// MFMA 16x16x32 lane layout (gfx950): lane l supplies A row l%16, k = 8*(l/16)..+8 and
// B column l%16, same k; it receives C rows 4*(l/16)+i, i < 4, column l%16.
#include "sf_unet_kernels.hpp"

namespace sf::unet {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTX = 16, kTY = 4, kTZ = 2;                 // output voxels per workgroup
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2; // with the 1-voxel halo
constexpr int kHalo    = kHX * kHY * kHZ;                  // 432
constexpr int kTile    = kTX * kTY * kTZ;                  // 128
constexpr int kThreads = 256;
constexpr int kElem    = 256;
constexpr int kRingFew = 8; // operand sets in flight on grids of few tiles

__device__ __forceinline__ float bf2f(const bf16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ bf16_t f2bf(const float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<bf16_t>((u >> 16) | 0x40u);
    return static_cast<bf16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ __bf16 as_bf(const bf16_t h) { return __builtin_bit_cast(__bf16, h); }

struct Origin {
    int x0, y0, z0;
};

__device__ __forceinline__ Origin tile_origin(int t, const Dims d) {
    const int tx = (d.nx + kTX - 1) / kTX, ty = (d.ny + kTY - 1) / kTY;
    const int bx = t % tx;
    t /= tx;
    return Origin{bx * kTX, (t % ty) * kTY, (t / ty) * kTZ};
}

inline uint32_t tile_count(const Dims d) {
    return static_cast<uint32_t>(((d.nx + kTX - 1) / kTX) * ((d.ny + kTY - 1) / kTY) * ((d.nz + kTZ - 1) / kTZ));
}

// x[tile + halo][C] -> LDS [hz][hy][hx][C]; zeros outside the grid (zero padding). Four
// 16-byte loads per thread are issued before their stores (voxels outside the grid load
// voxel 0 and store zeros), so the staging waits out the load latency once per four.
// With up = 1, x holds the input at half resolution (extents d / 2) and the staged voxel
// (gx, gy, gz) of the full-resolution input is x's voxel (gx / 2, gy / 2, gz / 2): the
// nearest upsampling, fused.
template <int C>
__device__ __forceinline__ void stage_halo(const bf16_t* __restrict__ x, bf16_t* s, const Origin o, const Dims d, const int up = 0) {
    constexpr int kParts = C / 8;
    constexpr int kTotal = kHalo * kParts;
    constexpr int kU     = 4;
    const Dims ds{d.nx >> up, d.ny >> up, d.nz >> up};
    for (int c0 = static_cast<int>(threadIdx.x); c0 < kTotal; c0 += kU * kThreads) {
        uint4 v[kU];
        bool in[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c    = min(c0 + u * kThreads, kTotal - 1);
            const int vox  = c / kParts, part = c % kParts;
            const int hx = vox % kHX, hy = (vox / kHX) % kHY, hz = vox / (kHX * kHY);
            const int gx = o.x0 - 1 + hx, gy = o.y0 - 1 + hy, gz = o.z0 - 1 + hz;
            in[u]          = gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz;
            const uint64_t src = in[u] ? lin(gx >> up, gy >> up, gz >> up, ds) * C + part * 8 : 0;
            v[u]           = *reinterpret_cast<const uint4*>(x + src);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = c0 + u * kThreads;
            if (c < kTotal) *reinterpret_cast<uint4*>(s + c * 8) = in[u] ? v[u] : make_uint4(0u, 0u, 0u, 0u);
        }
    }
}

// g[tile][C] -> LDS [tz][ty][tx][C]; zeros outside the grid.
template <int C>
__device__ __forceinline__ void stage_tile(const bf16_t* __restrict__ g, bf16_t* s, const Origin o, const Dims d) {
    constexpr int kParts = C / 8;
    for (int c = threadIdx.x; c < kTile * kParts; c += kThreads) {
        const int vox = c / kParts, part = c % kParts;
        const int tx = vox % kTX, ty = (vox / kTX) % kTY, tz = vox / (kTX * kTY);
        const int gx = o.x0 + tx, gy = o.y0 + ty, gz = o.z0 + tz;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gx < d.nx && gy < d.ny && gz < d.nz) v = *reinterpret_cast<const uint4*>(g + lin(gx, gy, gz, d) * C + part * 8);
        *reinterpret_cast<uint4*>(s + vox * C + part * 8) = v;
    }
}

// NB output-channel blocks of 16 per workgroup; with NB < CO / 16 the CO / 16 / NB
// workgroups of one tile split the output channels (for grids of few tiles, where one
// workgroup per tile leaves the GPU nearly idle); every output keeps the same K order.
// R: operand sets in flight (the prefetch distance is R - 1 steps).
template <int CI, int CO, bool kF32, int NB, int R>
__global__ __launch_bounds__(kThreads) void k_conv3(const bf16_t* __restrict__ x, const bf16_t* __restrict__ w, const float* __restrict__ bias, void* __restrict__ y, const bf16_t* __restrict__ aux, const int epi, const Dims d) {
    static_assert(NB >= 1 && (CO / 16) % NB == 0, "channel split");
    constexpr int kSplit = CO / 16 / NB;
    __shared__ __attribute__((aligned(16))) bf16_t s_x[kHalo * CI];
    const Origin o = tile_origin(static_cast<int>(blockIdx.x) / kSplit, d);
    const int nb0  = (static_cast<int>(blockIdx.x) % kSplit) * NB;
    stage_halo<CI>(x, s_x, o, d, (epi & kEpiUpInput) ? 1 : 0);
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 4, m = lane & 15;
    constexpr int KSTEPS = (27 * CI + 31) / 32;
    constexpr bool kWholeSteps = (27 * CI) % 32 == 0; // no step runs past the last tap
    f32x4 acc[2][NB];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // Operands of step ks: A from the staged tile in LDS, B (the weights) from L2. They are
    // fetched R - 1 steps ahead into a ring of R register sets (compile-time slots):
    // fetched at their use, each step waited out the weights' L2 latency. The
    // accumulation order is that of the plain step loop.
    struct Ops {
        bf16x8 a[2], b[NB];
    };
    auto fetch = [&](const int ks, Ops& op) {
        const int k = ks * 32 + 8 * grp;
        const int tap = k / CI, ch = k % CI;
        if (kWholeSteps || tap < 27) {
            const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                const int rb = 2 * wave + mb, cy = rb % kTY, cz = rb / kTY;
                op.a[mb]     = *reinterpret_cast<const bf16x8*>(s_x + (((cz + dz) * kHY + (cy + dy)) * kHX + (m + dx)) * CI + ch);
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) op.b[nb] = *reinterpret_cast<const bf16x8*>(w + static_cast<size_t>(tap * CO + (nb0 + nb) * 16 + m) * CI + ch);
        } else {
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) op.a[mb] = bf16x8{};
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) op.b[nb] = bf16x8{};
        }
    };
    auto mma = [&](const Ops& op) {
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op.a[mb], op.b[nb], acc[mb][nb], 0, 0, 0);
    };
    Ops ring[R];
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
        if (i < KSTEPS) fetch(i, ring[i]);
    for (int ks = 0; ks < KSTEPS; ks += R) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (ks + j < KSTEPS) {
                if (ks + j + R - 1 < KSTEPS) fetch(ks + j + R - 1, ring[(j + R - 1) % R]);
                mma(ring[j]);
            }
        }
    }

#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        const int rb = 2 * wave + mb;
        const int gy = o.y0 + rb % kTY, gz = o.z0 + rb / kTY;
        if (gy >= d.ny || gz >= d.nz) continue;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int co    = (nb0 + nb) * 16 + m;
            const float bco = bias ? bias[co] : 0.0f;
            if ((epi & kEpiDense0) && co != 0) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gx = o.x0 + 4 * grp + i;
                if (gx >= d.nx) continue;
                const uint64_t idx = (epi & kEpiDense0) ? lin(gx, gy, gz, d) : lin(gx, gy, gz, d) * CO + co;
                float v            = acc[mb][nb][i] + bco;
                if (epi & kEpiMask) v = bf2f(aux[idx]) > 0.0f ? v : 0.0f;
                if (epi & kEpiRelu) v = fmaxf(v, 0.0f);
                if (epi & kEpiAdd) v += bf2f(aux[idx]);
                if constexpr (kF32)
                    static_cast<float*>(y)[idx] = v;
                else
                    static_cast<bf16_t*>(y)[idx] = f2bf(v);
            }
        }
    }
}

// Weight gradient: per tap, C[co][ci] += A[co][v] * B[v][ci] with K = voxels, 32 per MFMA
// (two 16-voxel rows of the tile). Workgroup (blockIdx.x, blockIdx.y) sums taps
// [TG*y, TG*y + TG) over `per_group` consecutive voxel tiles, then adds its partial sums
// to dw with FP32 atomics (order-dependent rounding).
template <int CI, int CO, int TG>
__global__ __launch_bounds__(kThreads) void k_conv3_wgrad(const bf16_t* __restrict__ x, const bf16_t* __restrict__ g, float* __restrict__ dw, const Dims d, const int per_group, const int tiles) {
    __shared__ __attribute__((aligned(16))) bf16_t s_x[kHalo * CI];
    __shared__ __attribute__((aligned(16))) bf16_t s_g[kTile * CO];
    constexpr int CB = CO / 16, IB = CI / 16;
    constexpr int NBLK = TG * CB * IB;
    constexpr int PER  = (NBLK + 3) / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 4, m = lane & 15;
    const int tap0 = static_cast<int>(blockIdx.y) * TG;
    f32x4 acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    const int t0 = static_cast<int>(blockIdx.x) * per_group;
    const int t1 = min(t0 + per_group, tiles);
    for (int t = t0; t < t1; ++t) {
        const Origin o = tile_origin(t, d);
        __syncthreads();
        stage_halo<CI>(x, s_x, o, d);
        stage_tile<CO>(g, s_g, o, d);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTile / 32; ++ks) {
            // this lane's 8 voxels: row r (y, z) of the tile, x = xb..xb+7
            const int r = 2 * ks + (grp >> 1), xb = 8 * (grp & 1);
            const int cy = r % kTY, cz = r / kTY;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int blk = wave + 4 * j;
                if (blk >= NBLK) continue;
                const int tap = tap0 + blk / (CB * IB);
                if (tap >= 27) continue;
                const int rest = blk % (CB * IB), cob = rest / IB, cib = rest % IB;
                const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
                const int co = cob * 16 + m, ci = cib * 16 + m;
                bf16x8 a, b;
                const bf16_t* ga = s_g + (r * kTX + xb) * CO + co;
                const bf16_t* xb_ = s_x + (((cz + dz) * kHY + (cy + dy)) * kHX + (xb + dx)) * CI + ci;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    a[i] = as_bf(ga[i * CO]);
                    b[i] = as_bf(xb_[i * CI]);
                }
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
            }
        }
    }
    if (t0 >= t1) return;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int blk = wave + 4 * j;
        if (blk >= NBLK) continue;
        const int tap = tap0 + blk / (CB * IB);
        if (tap >= 27) continue;
        const int rest = blk % (CB * IB), cob = rest / IB, cib = rest % IB;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = cob * 16 + 4 * grp + i, ci = cib * 16 + m;
            atomicAdd(dw + (static_cast<size_t>(tap) * CO + co) * CI + ci, acc[j][i]);
        }
    }
}

// db[co] += sum_v g[v][co]
template <int C>
__global__ __launch_bounds__(kThreads) void k_bias_grad(const bf16_t* __restrict__ g, float* __restrict__ db, const uint64_t n, const uint64_t per_group) {
    __shared__ float s[kThreads];
    constexpr int kLanes = kThreads / C;
    const int co = threadIdx.x % C, vl = threadIdx.x / C;
    const uint64_t v0 = blockIdx.x * per_group, v1 = min(v0 + per_group, n);
    float sum = 0.0f;
    for (uint64_t v = v0 + vl; v < v1; v += kLanes) sum += bf2f(g[v * C + co]);
    s[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int l = 0; l < kLanes; ++l) t += s[l * C + threadIdx.x];
        atomicAdd(db + threadIdx.x, t);
    }
}

// out[v][c] = scale * sum of the 8 children of v (in dims = 2 * out dims), 8 channels per thread
__global__ __launch_bounds__(kElem) void k_gather8(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, const int c, const float scale, const Dims od) {
    const int parts    = c / 8;
    const uint64_t n   = static_cast<uint64_t>(od.nx) * od.ny * od.nz * parts;
    const Dims id{2 * od.nx, 2 * od.ny, 2 * od.nz};
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; t < n; t += static_cast<uint64_t>(gridDim.x) * kElem) {
        const int part = static_cast<int>(t % parts);
        uint64_t v     = t / parts;
        const int x = static_cast<int>(v % od.nx), y = static_cast<int>((v / od.nx) % od.ny), z = static_cast<int>(v / (static_cast<uint64_t>(od.nx) * od.ny));
        float acc[8] = {};
        for (int k = 0; k < 8; ++k) {
            const uint4 q = *reinterpret_cast<const uint4*>(in + lin(2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2), id) * c + part * 8);
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[2 * j] += bf2f(static_cast<bf16_t>(w[j] & 0xffffu));
                acc[2 * j + 1] += bf2f(static_cast<bf16_t>(w[j] >> 16));
            }
        }
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = static_cast<uint32_t>(f2bf(acc[2 * j] * scale)) | (static_cast<uint32_t>(f2bf(acc[2 * j + 1] * scale)) << 16);
        *reinterpret_cast<uint4*>(out + (static_cast<uint64_t>(v) * c + part * 8)) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// out[v][c] = scale * in[v / 2][c] (out dims = 2 * in dims)
__global__ __launch_bounds__(kElem) void k_spread8(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, const int c, const float scale, const Dims od) {
    const int parts  = c / 8;
    const uint64_t n = static_cast<uint64_t>(od.nx) * od.ny * od.nz * parts;
    const Dims id{od.nx / 2, od.ny / 2, od.nz / 2};
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; t < n; t += static_cast<uint64_t>(gridDim.x) * kElem) {
        const int part = static_cast<int>(t % parts);
        const uint64_t v = t / parts;
        const int x = static_cast<int>(v % od.nx), y = static_cast<int>((v / od.nx) % od.ny), z = static_cast<int>(v / (static_cast<uint64_t>(od.nx) * od.ny));
        uint4 q = *reinterpret_cast<const uint4*>(in + lin(x / 2, y / 2, z / 2, id) * c + part * 8);
        if (scale != 1.0f) {
            uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                w[j] = static_cast<uint32_t>(f2bf(bf2f(static_cast<bf16_t>(w[j] & 0xffffu)) * scale)) | (static_cast<uint32_t>(f2bf(bf2f(static_cast<bf16_t>(w[j] >> 16)) * scale)) << 16);
            q = make_uint4(w[0], w[1], w[2], w[3]);
        }
        *reinterpret_cast<uint4*>(out + v * c + part * 8) = q;
    }
}

__global__ __launch_bounds__(kElem) void k_add_mask(const bf16_t* __restrict__ a, const bf16_t* __restrict__ b, const bf16_t* __restrict__ mask, bf16_t* __restrict__ out, const uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * kElem) {
        float v = bf2f(a[i]);
        if (b) v += bf2f(b[i]);
        if (mask && !(bf2f(mask[i]) > 0.0f)) v = 0.0f;
        out[i] = f2bf(v);
    }
}

__global__ __launch_bounds__(kElem) void k_weights_to_bf16(const float* __restrict__ w, bf16_t* __restrict__ wb, bf16_t* __restrict__ wt, const int cin, const int cout) {
    const int n = 27 * cin * cout;
    for (int i = blockIdx.x * kElem + threadIdx.x; i < n; i += gridDim.x * kElem) {
        const int ci = i % cin, co = (i / cin) % cout, tap = i / (cin * cout);
        const bf16_t h = f2bf(w[i]);
        wb[i]          = h;
        wt[(static_cast<size_t>(26 - tap) * cin + ci) * cout + co] = h;
    }
}

__global__ __launch_bounds__(kElem) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, const uint64_t n, const float lr, const float b1, const float b2, const float eps, const float c1, const float c2) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * kElem) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i]           = mi;
        v[i]           = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ __launch_bounds__(kElem) void k_fill_u16(bf16_t* __restrict__ p, const uint16_t value, const uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * kElem) p[i] = value;
}
__global__ __launch_bounds__(kElem) void k_fill_f32(float* __restrict__ p, const float value, const uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * kElem) p[i] = value;
}

__global__ __launch_bounds__(kElem) void k_pack_input(const float* __restrict__ in, const float* __restrict__ scale_ptr, const float in_scale, bf16_t* __restrict__ a0, const uint64_t n) {
    const float s = scale_ptr ? *scale_ptr * in_scale : in_scale;
    for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; v < n; v += static_cast<uint64_t>(gridDim.x) * kElem) {
        const uint32_t c01 = static_cast<uint32_t>(f2bf(in[2 * v] * s)) | (static_cast<uint32_t>(f2bf(in[2 * v + 1])) << 16);
        uint4* dst = reinterpret_cast<uint4*>(a0 + v * 16);
        dst[0]     = make_uint4(c01, 0u, 0u, 0u);
        dst[1]     = make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ __launch_bounds__(kElem) void k_mse_grad(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ mean, const float w, bf16_t* __restrict__ g, float* __restrict__ loss, const uint64_t n) {
    __shared__ float s[kElem / 64];
    const float inv_n = w / static_cast<float>(n);
    const float shift = mean ? *mean : 0.0f;
    float acc = 0.0f;
    for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; v < n; v += static_cast<uint64_t>(gridDim.x) * kElem) {
        const float dv = (y[v * 16] - shift) - t[v];
        acc += dv * dv;
        uint4* dst = reinterpret_cast<uint4*>(g + v * 16);
        dst[0]     = make_uint4(static_cast<uint32_t>(f2bf(2.0f * dv * inv_n)), 0u, 0u, 0u);
        dst[1]     = make_uint4(0u, 0u, 0u, 0u);
    }
    if (loss == nullptr) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (s[0] + s[1] + s[2] + s[3]) * inv_n);
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double s[kElem / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = s[0] + s[1] + s[2] + s[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kElem) void k_head_partials(const float* __restrict__ y, double* __restrict__ partials, const uint64_t n) {
    double acc = 0.0;
    for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; v < n; v += static_cast<uint64_t>(gridDim.x) * kElem) acc += y[v * 16];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(kElem) void k_head_mean(const double* __restrict__ partials, const int nb, const uint64_t n, float* __restrict__ mean) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kElem) acc += partials[b];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) *mean = static_cast<float>(t / static_cast<double>(n));
}

// Input packing with the normalisation computed in place: every block adds the per-block
// sums of squares of the restricted divergence (the same order as a one-block reduction),
// so all blocks get the same sigma; block 0 stores sigma and 1/sigma for the later stages.
__global__ __launch_bounds__(kElem) void k_pack_input_sigma(const float* __restrict__ in, const double* __restrict__ partials, const int n_partials, float* __restrict__ scal, bf16_t* __restrict__ a0, const uint64_t n) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < n_partials; b += kElem) acc += partials[b];
    const double sigma = sqrt(block_sum(acc) / static_cast<double>(n));
    const float inv    = sigma > 0.0 ? static_cast<float>(1.0 / sigma) : 0.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[0] = static_cast<float>(sigma);
        scal[1] = inv;
    }
    for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; v < n; v += static_cast<uint64_t>(gridDim.x) * kElem) {
        const uint32_t c01 = static_cast<uint32_t>(f2bf(in[2 * v] * inv)) | (static_cast<uint32_t>(f2bf(in[2 * v + 1])) << 16);
        uint4* dst = reinterpret_cast<uint4*>(a0 + v * 16);
        dst[0]     = make_uint4(c01, 0u, 0u, 0u);
        dst[1]     = make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ __launch_bounds__(kElem) void k_head_channel0(const float* __restrict__ y, float* __restrict__ out, const uint64_t n) {
    for (uint64_t v = blockIdx.x * static_cast<uint64_t>(kElem) + threadIdx.x; v < n; v += static_cast<uint64_t>(gridDim.x) * kElem) out[v] = y[v * 16];
}

inline unsigned elem_grid(const uint64_t n) {
    const uint64_t b = (n + kElem - 1) / kElem;
    return static_cast<unsigned>(b < 8192 ? (b ? b : 1) : 8192);
}

bool valid_c(const int c) { return c == 16 || c == 32 || c == 64; }

template <int CI, int CO>
hipError_t conv3_t(hipStream_t s, const bf16_t* x, const bf16_t* w, const float* bias, bf16_t* yb, float* yf, const bf16_t* aux, const int epi, const Dims d) {
    const unsigned tiles = tile_count(d);
    constexpr int kNB    = CO / 16;
    // few tiles: one workgroup per 16 output channels of a tile, and (with so few waves on
    // the GPU) the operands of kRingFew - 1 steps in flight in each
    if (tiles * kNB <= 1024) {
        if (yf) k_conv3<CI, CO, true, 1, kRingFew><<<tiles * kNB, kThreads, 0, s>>>(x, w, bias, yf, aux, epi, d);
        else k_conv3<CI, CO, false, 1, kRingFew><<<tiles * kNB, kThreads, 0, s>>>(x, w, bias, yb, aux, epi, d);
    } else if (yf) {
        k_conv3<CI, CO, true, kNB, 3><<<tiles, kThreads, 0, s>>>(x, w, bias, yf, aux, epi, d);
    } else {
        k_conv3<CI, CO, false, kNB, 3><<<tiles, kThreads, 0, s>>>(x, w, bias, yb, aux, epi, d);
    }
    return hipGetLastError();
}

template <int CI, int CO>
hipError_t wgrad_t(hipStream_t s, const bf16_t* x, const bf16_t* g, float* dw, float* db, const Dims d) {
    constexpr int kBlocksPerTap = (CI / 16) * (CO / 16);
    constexpr int TG            = kBlocksPerTap >= 64 ? 1 : (64 / kBlocksPerTap > 27 ? 27 : 64 / kBlocksPerTap);
    const int tiles             = static_cast<int>(tile_count(d));
    const int gy                = (27 + TG - 1) / TG;
    int per                     = (tiles * gy + 1023) / 1024;
    if (per < 1) per = 1;
    const int gx = (tiles + per - 1) / per;
    k_conv3_wgrad<CI, CO, TG><<<dim3(gx, gy), kThreads, 0, s>>>(x, g, dw, d, per, tiles);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (db) {
        const uint64_t n   = static_cast<uint64_t>(d.nx) * d.ny * d.nz;
        const uint64_t pg  = (n + 1023) / 1024 < 64 ? 64 : (n + 1023) / 1024;
        const unsigned bgr = static_cast<unsigned>((n + pg - 1) / pg);
        k_bias_grad<CO><<<bgr, kThreads, 0, s>>>(g, db, n, pg);
    }
    return hipGetLastError();
}

} // namespace

#define SF_CONV_CASES(FN, ...)                                      \
    switch (cin * 1000 + cout) {                                    \
    case 16016: return FN<16, 16>(__VA_ARGS__);                     \
    case 16032: return FN<16, 32>(__VA_ARGS__);                     \
    case 32016: return FN<32, 16>(__VA_ARGS__);                     \
    case 32032: return FN<32, 32>(__VA_ARGS__);                     \
    case 32064: return FN<32, 64>(__VA_ARGS__);                     \
    case 64032: return FN<64, 32>(__VA_ARGS__);                     \
    case 64064: return FN<64, 64>(__VA_ARGS__);                     \
    default: return hipErrorInvalidValue;                           \
    }

hipError_t launch_conv3(hipStream_t s, const int cin, const int cout, const bf16_t* x, const bf16_t* w, const float* bias, bf16_t* y_bf16, float* y_f32, const bf16_t* aux, const int epilogue, const Dims d) {
    if ((y_bf16 == nullptr) == (y_f32 == nullptr)) return hipErrorInvalidValue;
    if ((epilogue & (kEpiAdd | kEpiMask)) && aux == nullptr) return hipErrorInvalidValue;
    if ((epilogue & kEpiDense0) && y_f32 == nullptr) return hipErrorInvalidValue;
    if ((epilogue & kEpiUpInput) && ((d.nx | d.ny | d.nz) & 1)) return hipErrorInvalidValue;
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0) return hipErrorInvalidValue;
    SF_CONV_CASES(conv3_t, s, x, w, bias, y_bf16, y_f32, aux, epilogue, d)
}

hipError_t launch_conv3_wgrad(hipStream_t s, const int cin, const int cout, const bf16_t* x, const bf16_t* g, float* dw, float* db, const Dims d) {
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0) return hipErrorInvalidValue;
    SF_CONV_CASES(wgrad_t, s, x, g, dw, db, d)
}

#undef SF_CONV_CASES

hipError_t launch_pool2(hipStream_t s, const int c, const bf16_t* in, bf16_t* out, const Dims od) {
    if (!valid_c(c)) return hipErrorInvalidValue;
    k_gather8<<<elem_grid(static_cast<uint64_t>(od.nx) * od.ny * od.nz * (c / 8)), kElem, 0, s>>>(in, out, c, 0.125f, od);
    return hipGetLastError();
}
hipError_t launch_up2_adjoint(hipStream_t s, const int c, const bf16_t* g_out, bf16_t* g_in, const Dims id) {
    if (!valid_c(c)) return hipErrorInvalidValue;
    k_gather8<<<elem_grid(static_cast<uint64_t>(id.nx) * id.ny * id.nz * (c / 8)), kElem, 0, s>>>(g_out, g_in, c, 1.0f, id);
    return hipGetLastError();
}
hipError_t launch_up2(hipStream_t s, const int c, const bf16_t* in, bf16_t* out, const Dims od) {
    if (!valid_c(c)) return hipErrorInvalidValue;
    k_spread8<<<elem_grid(static_cast<uint64_t>(od.nx) * od.ny * od.nz * (c / 8)), kElem, 0, s>>>(in, out, c, 1.0f, od);
    return hipGetLastError();
}
hipError_t launch_pool2_adjoint(hipStream_t s, const int c, const bf16_t* g_out, bf16_t* g_in, const Dims id) {
    if (!valid_c(c)) return hipErrorInvalidValue;
    k_spread8<<<elem_grid(static_cast<uint64_t>(id.nx) * id.ny * id.nz * (c / 8)), kElem, 0, s>>>(g_out, g_in, c, 0.125f, id);
    return hipGetLastError();
}
hipError_t launch_add_mask(hipStream_t s, const bf16_t* a, const bf16_t* b, const bf16_t* mask, bf16_t* out, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_add_mask<<<elem_grid(n), kElem, 0, s>>>(a, b, mask, out, n);
    return hipGetLastError();
}
hipError_t launch_weights_to_bf16(hipStream_t s, const float* w, bf16_t* wb, bf16_t* wt, const int cin, const int cout) {
    k_weights_to_bf16<<<elem_grid(static_cast<uint64_t>(27) * cin * cout), kElem, 0, s>>>(w, wb, wt, cin, cout);
    return hipGetLastError();
}
hipError_t launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, const uint64_t n, const float lr, const float beta1, const float beta2, const float eps, const int t) {
    const float c1 = 1.0f - powf(beta1, static_cast<float>(t));
    const float c2 = 1.0f - powf(beta2, static_cast<float>(t));
    k_adam<<<elem_grid(n), kElem, 0, s>>>(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t launch_fill_u16(hipStream_t s, bf16_t* p, const uint16_t value, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_fill_u16<<<elem_grid(n), kElem, 0, s>>>(p, value, n);
    return hipGetLastError();
}
hipError_t launch_fill_f32(hipStream_t s, float* p, const float value, const uint64_t n) {
    if (n == 0) return hipSuccess;
    k_fill_f32<<<elem_grid(n), kElem, 0, s>>>(p, value, n);
    return hipGetLastError();
}
hipError_t launch_pack_input_sigma(hipStream_t s, const float* in, const double* partials, const int n_partials, float* scal, bf16_t* a0, const uint64_t n) {
    k_pack_input_sigma<<<elem_grid(n), kElem, 0, s>>>(in, partials, n_partials, scal, a0, n);
    return hipGetLastError();
}

hipError_t launch_pack_input(hipStream_t s, const float* in, const float* scale_ptr, const float in_scale, bf16_t* a0, const uint64_t n) {
    k_pack_input<<<elem_grid(n), kElem, 0, s>>>(in, scale_ptr, in_scale, a0, n);
    return hipGetLastError();
}
hipError_t launch_mse_grad(hipStream_t s, const float* y, const float* t, const float* mean, const float w, bf16_t* g, float* loss, const uint64_t n) {
    k_mse_grad<<<elem_grid(n), kElem, 0, s>>>(y, t, mean, w, g, loss, n);
    return hipGetLastError();
}
hipError_t launch_head_mean(hipStream_t s, const float* y, double* partials, float* mean, const uint64_t n) {
    const uint64_t b  = (n + kElem - 1) / kElem;
    const unsigned nb = static_cast<unsigned>(b < 1024 ? (b ? b : 1) : 1024);
    k_head_partials<<<nb, kElem, 0, s>>>(y, partials, n);
    k_head_mean<<<1, kElem, 0, s>>>(partials, static_cast<int>(nb), n, mean);
    return hipGetLastError();
}
hipError_t launch_head_channel0(hipStream_t s, const float* y, float* out, const uint64_t n) {
    k_head_channel0<<<elem_grid(n), kElem, 0, s>>>(y, out, n);
    return hipGetLastError();
}

} // namespace sf::unet
