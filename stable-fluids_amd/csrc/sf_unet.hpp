// sf_unet.hpp -- the warm-start network as the step runtime uses it (sf_unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sf_unet_kernels.hpp"

namespace sf::unet {

struct Net;

// nx, ny, nz multiples of 4; stream nullptr: the network owns a new stream.
hipError_t net_create(int nx, int ny, int nz, hipStream_t stream, Net** out);
void net_destroy(Net* n);
hipStream_t net_stream(const Net& n);
Dims net_dims(const Net& n);
// Packed input tensor [v][16] BF16 (channel 0 divergence, 1 wall mask, rest zero),
// written by the caller before net_forward_packed.
bf16_t* net_input(Net& n);
// Forward pass on the network stream; the result is net_output() [v] FP32.
hipError_t net_forward_packed(Net& n, bool train);
const float* net_output(const Net& n);
// Host FP32 parameters in the public layout (synchronous).
hipError_t net_set_params(Net& n, const float* host);

} // namespace sf::unet
