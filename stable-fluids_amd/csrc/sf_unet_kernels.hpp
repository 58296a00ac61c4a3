// sf_unet_kernels.hpp -- launchers of the warm-start network kernels (sf_unet_kernels.hip).
//
// Tensors are channels-last ([z][y][x][c], c fastest) BF16 bit patterns (uint16_t) with
// C in {16, 32, 64}; weights of a 3x3x3 convolution are [27][cout][cin] with tap =
// 9*dz + 3*dy + dx (offsets dz-1, dy-1, dx-1), zero padding at the grid faces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sf_device.hpp"

namespace sf::unet {

using bf16_t = uint16_t;

enum ConvEpilogue : int {
    kEpiNone = 0,
    kEpiRelu = 1, // y = max(acc + bias, 0)
    kEpiAdd  = 2, // y = (...) + aux   (after the ReLU when both are set)
    kEpiMask = 4, // y = (acc + bias) * (aux > 0)
    kEpiUpInput = 8,  // x holds the input at half resolution: nearest upsampling fused into the staging
    kEpiDense0  = 16, // FP32 output of channel 0 only, dense (y[voxel])
};

// y[v][co] = epilogue(sum_tap sum_ci w[tap][co][ci] * x[v + off(tap)][ci] + bias[co]).
// bias may be null (zero); y is BF16 (y_f32 == nullptr) or FP32 (y_bf16 == nullptr).
hipError_t launch_conv3(hipStream_t s, int cin, int cout, const bf16_t* x, const bf16_t* w, const float* bias, bf16_t* y_bf16, float* y_f32, const bf16_t* aux, int epilogue, Dims d);
// dw[tap][co][ci] += sum_v g[v][co] * x[v + off(tap)][ci]; db[co] += sum_v g[v][co] (db may be null).
hipError_t launch_conv3_wgrad(hipStream_t s, int cin, int cout, const bf16_t* x, const bf16_t* g, float* dw, float* db, Dims d);

// 2x2x2 average (out dims = in dims / 2) and nearest 2x upsampling (out = 2 * in), and
// their adjoints; c channels.
hipError_t launch_pool2(hipStream_t s, int c, const bf16_t* in, bf16_t* out, Dims out_dims);
hipError_t launch_up2(hipStream_t s, int c, const bf16_t* in, bf16_t* out, Dims out_dims);
hipError_t launch_pool2_adjoint(hipStream_t s, int c, const bf16_t* g_out, bf16_t* g_in, Dims in_dims);
hipError_t launch_up2_adjoint(hipStream_t s, int c, const bf16_t* g_out, bf16_t* g_in, Dims in_dims);
// out = a + b (b may be null: copy), then optionally * (mask > 0); n elements.
hipError_t launch_add_mask(hipStream_t s, const bf16_t* a, const bf16_t* b, const bf16_t* mask, bf16_t* out, uint64_t n);

// FP32 master parameters -> BF16 forward weights [27][co][ci] and flipped transposed
// weights wt[26 - tap][ci][co] (backward data).
hipError_t launch_weights_to_bf16(hipStream_t s, const float* w, bf16_t* wb, bf16_t* wt, int cin, int cout);
// Adam on n FP32 parameters: m, v moments, step t (1-based, bias correction).
hipError_t launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, uint64_t n, float lr, float beta1, float beta2, float eps, int t);
hipError_t launch_fill_u16(hipStream_t s, bf16_t* p, uint16_t value, uint64_t n);
hipError_t launch_fill_f32(hipStream_t s, float* p, float value, uint64_t n);

// Network input: a0[v][0] = in[v][0] * in_scale (*scale_ptr when given), a0[v][1] =
// in[v][1], channels 2..15 zero. in is [v][2] FP32.
hipError_t launch_pack_input(hipStream_t s, const float* in, const float* scale_ptr, float in_scale, bf16_t* a0, uint64_t n);
// The same with scale 1/sigma for the pair-channel input in, sigma = sqrt(sum of the
// n_partials per-block sums / n); sigma and 1/sigma are stored in scal[0], scal[1].
hipError_t launch_pack_input_sigma(hipStream_t s, const float* in, const double* partials, int n_partials, float* scal, bf16_t* a0, uint64_t n);
// Loss of channel 0 of the FP32 head output y[v][16] against target t[v]:
// loss += w sum (y0 - t)^2 / n (into *loss, FP32 atomic, may be null); g[v][0] =
// 2 w (y0 - t) / n as BF16, channels 1..15 zero.
// With `mean` (device scalar, from launch_head_mean) y0 is replaced by y0 - *mean.
hipError_t launch_mse_grad(hipStream_t s, const float* y, const float* t, const float* mean, float w, bf16_t* g, float* loss, uint64_t n);
// *mean = mean over v of y[v][0] (FP64 sums; partials: 1024 doubles).
hipError_t launch_head_mean(hipStream_t s, const float* y, double* partials, float* mean, uint64_t n);
// out[v] = y[v][0] (channel 0 of the FP32 head output).
hipError_t launch_head_channel0(hipStream_t s, const float* y, float* out, uint64_t n);

} // namespace sf::unet
