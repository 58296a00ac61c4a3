/* stable-fluids-unet.h -- pressure warm-start network of libstable-fluids.so (gfx950).
 *
 * No counterpart in the reference: the reference projection always restarts pressure
 * from zero (backend-cuda.cu:644). This network predicts the pressure from the
 * divergence; stable-fluids-3d-ext.h uses it as the first pressure iterate.
 *
 * Network: 3-level U-Net, 3x3x3 convolutions with zero padding, ReLU, 2x2x2 average
 * pooling, nearest 2x upsampling, additive skips:
 *   in(2: divergence, wall mask) -> 16 -> 16 [skip1] -> pool -> 32 -> 32 [skip2] -> pool
 *   -> 64 -> 64 -> up -> 32 (+skip2) -> 32 -> up -> 16 (+skip1) -> 16 -> 1 (no ReLU).
 * BF16 weights and activations, FP32 accumulation (MFMA), FP32 master parameters.
 * Parameter layout (FP32, host): the 11 layers in order, each weights [27][cout][cin]
 * (tap = 9*dz + 3*dy + dx, offsets -1..1) followed by bias [cout]; 319,569 values.
 * Grid extents must be multiples of 4. Tensors are packed [z][y][x] (x fastest).
 */
#ifndef STABLE_FLUIDS_UNET_H
#define STABLE_FLUIDS_UNET_H

#include <stdint.h>

#include "stable-fluids-3d.h"

#ifdef __cplusplus
extern "C" {
#endif

STABLE_FLUIDS_API uint64_t stable_fluids_unet_param_count(void);

/* Allocates a network for nx*ny*nz inputs (extents multiples of 4) on `stream`
 * (NULL: a new stream owned by the network). Parameters start at zero. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_create_cuda(int32_t nx, int32_t ny, int32_t nz, void* stream, void** out_net);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_destroy_cuda(void* net);

/* Host FP32 parameters in the layout above (count must equal the parameter count).
 * Synchronous. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_set_params_cuda(void* net, const float* params, uint64_t count);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_get_params_cuda(void* net, float* params, uint64_t count);
/* Gradient of the last backward call, same layout. Synchronous. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_get_grads_cuda(void* net, float* grads, uint64_t count);

/* output[v] = net(input)[v]; input is device [v][2] FP32 (divergence, wall mask),
 * output device [v] FP32. Asynchronous on the network stream. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_forward_cuda(void* net, const float* input, float* output);
/* Forward, loss L = weight * mean_v (q[v] - target[v])^2 added to *loss (device FP32, may
 * be NULL) and the parameter gradient dL/dparams (replaces the previous one), q =
 * net(input). flags bit 0 (STABLE_FLUIDS_UNET_LOSS_CENTERED): q is replaced by
 * q - mean(q), for zero-mean targets defined up to a constant (closed boxes).
 * Asynchronous. */
#define STABLE_FLUIDS_UNET_LOSS_CENTERED 1u
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_backward_cuda(void* net, const float* input, const float* target, float weight, uint32_t flags, float* loss);
/* One Adam step on the parameters with the last gradient. Asynchronous. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_adam_step_cuda(void* net, float lr, float beta1, float beta2, float eps);

/* Single-layer kernels for gradient checks (asynchronous on `stream`, device buffers).
 * x: [v][cin] BF16 bits; w: [27][cout][cin] BF16 bits; bias: [cout] FP32 or NULL;
 * y: [v][cout] FP32 = conv(x) + bias (ReLU when relu != 0); cin, cout in {16, 32, 64}. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_conv3_cuda(void* stream, int32_t cin, int32_t cout, const uint16_t* x, const uint16_t* w, const float* bias, float* y, int32_t relu, int32_t nx, int32_t ny, int32_t nz);
/* dw[27][cout][cin] += sum_v g[v][co] x[v + off(tap)][ci]; db[cout] += sum_v g[v][co]
 * (db may be NULL). x: [v][cin], g: [v][cout], BF16 bits. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_conv3_wgrad_cuda(void* stream, int32_t cin, int32_t cout, const uint16_t* x, const uint16_t* g, float* dw, float* db, int32_t nx, int32_t ny, int32_t nz);

#ifdef __cplusplus
}
#endif

#endif /* STABLE_FLUIDS_UNET_H */
