// sf_unet.hip -- host object and C ABI of the pressure warm-start network
// (include/stable-fluids-unet.h). No counterpart in the reference.
//
// Layers (cin -> cout as run; the 2-channel input and the 1-channel output are padded to
// 16 channels with zero weights so every convolution runs on the MFMA kernel):
//   L1 2(16)->16, L2 16->16 | pool | L3 16->32, L4 32->32 | pool | L5 32->64, L6 64->64
//   | up | L7 64->32 (+a4), L8 32->32 | up | L9 32->16 (+a2), L10 16->16, L11 16->1(16).
// Master parameters are FP32 in the padded layout; BF16 copies (forward [tap][co][ci]
// and flipped transposed for the backward-data pass) are refreshed after every update.
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/stable-fluids-unet.h"
#include "sf_unet.hpp"
#include "sf_unet_kernels.hpp"

namespace sf::unet {

namespace {

constexpr int kLayers = 11;
// real and padded channel counts per layer
constexpr int kCinReal[kLayers]  = {2, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
constexpr int kCoutReal[kLayers] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 1};
constexpr int kCin[kLayers]      = {16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
constexpr int kCout[kLayers]     = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16};

uint64_t public_count() {
    uint64_t n = 0;
    for (int l = 0; l < kLayers; ++l) n += 27ull * kCinReal[l] * kCoutReal[l] + kCoutReal[l];
    return n;
}

} // namespace

struct Net {
    Dims d[3]{};           // levels: full, /2, /4
    uint64_t vox[3]{};
    hipStream_t stream = nullptr;
    bool owns_stream   = false;
    size_t w_off[kLayers]{}, b_off[kLayers]{}, padded = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16_t *wb[kLayers]{}, *wt[kLayers]{};
    int adam_t = 0;
    // activations (forward) and gradients (backward), channels-last BF16
    bf16_t *a0, *a1, *a2, *p1, *a3, *a4, *p2, *a5, *a6, *u1, *r7, *s7, *a8, *u2, *r9, *s9, *a10;
    float* y = nullptr;    // head output [v][16] FP32
    float* q = nullptr;    // channel 0 of y
    double* partials = nullptr; // 1024 block sums
    float* mean = nullptr;      // mean of q (centred loss)
    bf16_t *gy, *g10, *ds9, *g9, *du2, *da8, *g8, *ds7, *g7, *du1, *da6, *g6, *g5, *dp2, *t4, *g4, *g3, *dp1, *t2, *g2, *g1;
    std::vector<void*> allocations;

    ~Net() {
        for (void* p : allocations) (void)hipFree(p);
        if (owns_stream && stream) (void)hipStreamDestroy(stream);
    }
    template <class T>
    bool alloc(T*& p, const size_t bytes) {
        void* raw = nullptr;
        if (hipMalloc(&raw, bytes < 16 ? 16 : bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        allocations.push_back(raw);
        p = static_cast<T*>(raw);
        return true;
    }
    bool act(bf16_t*& p, const int level, const int c) { return alloc(p, vox[level] * c * sizeof(bf16_t)); }
};

#define SF_TRY(expr)                                      \
    do {                                                  \
        if ((expr) != hipSuccess) return hipErrorUnknown; \
    } while (0)

namespace {

hipError_t refresh_bf16(Net& n) {
    for (int l = 0; l < kLayers; ++l) SF_TRY(launch_weights_to_bf16(n.stream, n.params + n.w_off[l], n.wb[l], n.wt[l], kCin[l], kCout[l]));
    return hipSuccess;
}

hipError_t conv(Net& n, const int l, const bf16_t* x, bf16_t* y, const int epi, const bf16_t* aux, const int level) {
    return launch_conv3(n.stream, kCin[l], kCout[l], x, n.wb[l], n.params + n.b_off[l], y, nullptr, aux, epi, n.d[level]);
}
// backward data of layer l: dx = conv(g; flipped transposed w), optional mask
hipError_t dgrad(Net& n, const int l, const bf16_t* g, bf16_t* dx, const bf16_t* mask, const int level) {
    return launch_conv3(n.stream, kCout[l], kCin[l], g, n.wt[l], nullptr, dx, nullptr, mask, mask ? kEpiMask : kEpiNone, n.d[level]);
}
hipError_t wgrad(Net& n, const int l, const bf16_t* x, const bf16_t* g, const int level) {
    return launch_conv3_wgrad(n.stream, kCin[l], kCout[l], x, g, n.grads + n.w_off[l], n.grads + n.b_off[l], n.d[level]);
}

} // namespace

hipError_t net_create(const int nx, const int ny, const int nz, hipStream_t stream, Net** out) {
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 4 || ny % 4 || nz % 4) return hipErrorInvalidValue;
    std::unique_ptr<Net> n{new (std::nothrow) Net{}};
    if (!n) return hipErrorOutOfMemory;
    for (int lv = 0; lv < 3; ++lv) {
        n->d[lv]   = Dims{nx >> lv, ny >> lv, nz >> lv};
        n->vox[lv] = static_cast<uint64_t>(n->d[lv].nx) * n->d[lv].ny * n->d[lv].nz;
    }
    n->stream = stream;
    if (n->stream == nullptr) {
        if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return hipErrorUnknown;
        n->owns_stream = true;
    }
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        n->w_off[l] = off;
        off += 27ull * kCin[l] * kCout[l];
        n->b_off[l] = off;
        off += kCout[l];
    }
    n->padded = off;
    Net& m    = *n;
    bool ok   = m.alloc(m.params, off * 4) && m.alloc(m.grads, off * 4) && m.alloc(m.adam_m, off * 4) && m.alloc(m.adam_v, off * 4);
    for (int l = 0; ok && l < kLayers; ++l) ok = m.alloc(m.wb[l], 27ull * kCin[l] * kCout[l] * 2) && m.alloc(m.wt[l], 27ull * kCin[l] * kCout[l] * 2);
    ok = ok && m.act(m.a0, 0, 16) && m.act(m.a1, 0, 16) && m.act(m.a2, 0, 16) && m.act(m.p1, 1, 16) && m.act(m.a3, 1, 32) && m.act(m.a4, 1, 32) && m.act(m.p2, 2, 32) && m.act(m.a5, 2, 64) && m.act(m.a6, 2, 64) &&
         m.act(m.u1, 1, 64) && m.act(m.r7, 1, 32) && m.act(m.s7, 1, 32) && m.act(m.a8, 1, 32) && m.act(m.u2, 0, 32) && m.act(m.r9, 0, 16) && m.act(m.s9, 0, 16) && m.act(m.a10, 0, 16) &&
         m.alloc(m.y, m.vox[0] * 16 * 4) && m.alloc(m.q, m.vox[0] * 4) && m.alloc(m.partials, 1024 * sizeof(double)) && m.alloc(m.mean, 16);
    ok = ok && m.act(m.gy, 0, 16) && m.act(m.g10, 0, 16) && m.act(m.ds9, 0, 16) && m.act(m.g9, 0, 16) && m.act(m.du2, 0, 32) && m.act(m.da8, 1, 32) && m.act(m.g8, 1, 32) && m.act(m.ds7, 1, 32) &&
         m.act(m.g7, 1, 32) && m.act(m.du1, 1, 64) && m.act(m.da6, 2, 64) && m.act(m.g6, 2, 64) && m.act(m.g5, 2, 64) && m.act(m.dp2, 2, 32) && m.act(m.t4, 1, 32) && m.act(m.g4, 1, 32) &&
         m.act(m.g3, 1, 32) && m.act(m.dp1, 1, 16) && m.act(m.t2, 0, 16) && m.act(m.g2, 0, 16) && m.act(m.g1, 0, 16);
    if (!ok) return hipErrorOutOfMemory;
    SF_TRY(launch_fill_f32(m.stream, m.params, 0.0f, off));
    SF_TRY(launch_fill_f32(m.stream, m.grads, 0.0f, off));
    SF_TRY(launch_fill_f32(m.stream, m.adam_m, 0.0f, off));
    SF_TRY(launch_fill_f32(m.stream, m.adam_v, 0.0f, off));
    SF_TRY(refresh_bf16(m));
    if (hipStreamSynchronize(m.stream) != hipSuccess) return hipErrorUnknown;
    *out = n.release();
    return hipSuccess;
}

void net_destroy(Net* n) { delete n; }
hipStream_t net_stream(const Net& n) { return n.stream; }
Dims net_dims(const Net& n) { return n.d[0]; }

// Forward pass from the packed input a0 (already written) to q. With `train`, the
// ReLU outputs before each skip addition are kept for the backward pass.
hipError_t net_forward_packed(Net& n, const bool train) {
    SF_TRY(conv(n, 0, n.a0, n.a1, kEpiRelu, nullptr, 0));
    SF_TRY(conv(n, 1, n.a1, n.a2, kEpiRelu, nullptr, 0));
    SF_TRY(launch_pool2(n.stream, 16, n.a2, n.p1, n.d[1]));
    SF_TRY(conv(n, 2, n.p1, n.a3, kEpiRelu, nullptr, 1));
    SF_TRY(conv(n, 3, n.a3, n.a4, kEpiRelu, nullptr, 1));
    SF_TRY(launch_pool2(n.stream, 32, n.a4, n.p2, n.d[2]));
    SF_TRY(conv(n, 4, n.p2, n.a5, kEpiRelu, nullptr, 2));
    SF_TRY(conv(n, 5, n.a5, n.a6, kEpiRelu, nullptr, 2));
    // training keeps the upsampled inputs and the 16-channel head output for the backward
    // pass; inference stages the upsampled inputs straight from the coarse tensors and
    // writes only the head's channel 0 (the same values)
    if (train) {
        SF_TRY(launch_up2(n.stream, 64, n.a6, n.u1, n.d[1]));
        SF_TRY(conv(n, 6, n.u1, n.r7, kEpiRelu, nullptr, 1));
        SF_TRY(launch_add_mask(n.stream, n.r7, n.a4, nullptr, n.s7, n.vox[1] * 32));
    } else {
        SF_TRY(conv(n, 6, n.a6, n.s7, kEpiRelu | kEpiAdd | kEpiUpInput, n.a4, 1));
    }
    SF_TRY(conv(n, 7, n.s7, n.a8, kEpiRelu, nullptr, 1));
    if (train) {
        SF_TRY(launch_up2(n.stream, 32, n.a8, n.u2, n.d[0]));
        SF_TRY(conv(n, 8, n.u2, n.r9, kEpiRelu, nullptr, 0));
        SF_TRY(launch_add_mask(n.stream, n.r9, n.a2, nullptr, n.s9, n.vox[0] * 16));
    } else {
        SF_TRY(conv(n, 8, n.a8, n.s9, kEpiRelu | kEpiAdd | kEpiUpInput, n.a2, 0));
    }
    SF_TRY(conv(n, 9, n.s9, n.a10, kEpiRelu, nullptr, 0));
    if (!train) return launch_conv3(n.stream, kCin[10], kCout[10], n.a10, n.wb[10], n.params + n.b_off[10], nullptr, n.q, nullptr, kEpiDense0, n.d[0]);
    SF_TRY(launch_conv3(n.stream, kCin[10], kCout[10], n.a10, n.wb[10], n.params + n.b_off[10], nullptr, n.y, nullptr, kEpiNone, n.d[0]));
    return launch_head_channel0(n.stream, n.y, n.q, n.vox[0]);
}

bf16_t* net_input(Net& n) { return n.a0; }
const float* net_output(const Net& n) { return n.q; }

hipError_t net_backward(Net& n, const float* target, const float weight, const bool centered, float* loss) {
    SF_TRY(launch_fill_f32(n.stream, n.grads, 0.0f, n.padded));
    if (centered) SF_TRY(launch_head_mean(n.stream, n.y, n.partials, n.mean, n.vox[0]));
    SF_TRY(launch_mse_grad(n.stream, n.y, target, centered ? n.mean : nullptr, weight, n.gy, loss, n.vox[0]));
    SF_TRY(wgrad(n, 10, n.a10, n.gy, 0));
    SF_TRY(dgrad(n, 10, n.gy, n.g10, n.a10, 0));
    SF_TRY(wgrad(n, 9, n.s9, n.g10, 0));
    SF_TRY(dgrad(n, 9, n.g10, n.ds9, nullptr, 0));
    SF_TRY(launch_add_mask(n.stream, n.ds9, nullptr, n.r9, n.g9, n.vox[0] * 16));
    SF_TRY(wgrad(n, 8, n.u2, n.g9, 0));
    SF_TRY(dgrad(n, 8, n.g9, n.du2, nullptr, 0));
    SF_TRY(launch_up2_adjoint(n.stream, 32, n.du2, n.da8, n.d[1]));
    SF_TRY(launch_add_mask(n.stream, n.da8, nullptr, n.a8, n.g8, n.vox[1] * 32));
    SF_TRY(wgrad(n, 7, n.s7, n.g8, 1));
    SF_TRY(dgrad(n, 7, n.g8, n.ds7, nullptr, 1));
    SF_TRY(launch_add_mask(n.stream, n.ds7, nullptr, n.r7, n.g7, n.vox[1] * 32));
    SF_TRY(wgrad(n, 6, n.u1, n.g7, 1));
    SF_TRY(dgrad(n, 6, n.g7, n.du1, nullptr, 1));
    SF_TRY(launch_up2_adjoint(n.stream, 64, n.du1, n.da6, n.d[2]));
    SF_TRY(launch_add_mask(n.stream, n.da6, nullptr, n.a6, n.g6, n.vox[2] * 64));
    SF_TRY(wgrad(n, 5, n.a5, n.g6, 2));
    SF_TRY(dgrad(n, 5, n.g6, n.g5, n.a5, 2));
    SF_TRY(wgrad(n, 4, n.p2, n.g5, 2));
    SF_TRY(dgrad(n, 4, n.g5, n.dp2, nullptr, 2));
    SF_TRY(launch_pool2_adjoint(n.stream, 32, n.dp2, n.t4, n.d[1]));
    SF_TRY(launch_add_mask(n.stream, n.ds7, n.t4, n.a4, n.g4, n.vox[1] * 32));
    SF_TRY(wgrad(n, 3, n.a3, n.g4, 1));
    SF_TRY(dgrad(n, 3, n.g4, n.g3, n.a3, 1));
    SF_TRY(wgrad(n, 2, n.p1, n.g3, 1));
    SF_TRY(dgrad(n, 2, n.g3, n.dp1, nullptr, 1));
    SF_TRY(launch_pool2_adjoint(n.stream, 16, n.dp1, n.t2, n.d[0]));
    SF_TRY(launch_add_mask(n.stream, n.ds9, n.t2, n.a2, n.g2, n.vox[0] * 16));
    SF_TRY(wgrad(n, 1, n.a1, n.g2, 0));
    SF_TRY(dgrad(n, 1, n.g2, n.g1, n.a1, 0));
    return wgrad(n, 0, n.a0, n.g1, 0);
}

hipError_t net_adam(Net& n, const float lr, const float b1, const float b2, const float eps) {
    ++n.adam_t;
    SF_TRY(launch_adam(n.stream, n.params, n.grads, n.adam_m, n.adam_v, n.padded, lr, b1, b2, eps, n.adam_t));
    return refresh_bf16(n);
}

// public (unpadded) <-> padded parameter layouts, on the host
namespace {
hipError_t copy_params(Net& n, float* dev, float* host, const bool to_device) {
    std::vector<float> pad(n.padded, 0.0f);
    if (!to_device && hipMemcpy(pad.data(), dev, n.padded * 4, hipMemcpyDeviceToHost) != hipSuccess) return hipErrorUnknown;
    size_t pub = 0;
    for (int l = 0; l < kLayers; ++l) {
        for (int tap = 0; tap < 27; ++tap)
            for (int co = 0; co < kCoutReal[l]; ++co)
                for (int ci = 0; ci < kCinReal[l]; ++ci, ++pub) {
                    float& slot = pad[n.w_off[l] + (static_cast<size_t>(tap) * kCout[l] + co) * kCin[l] + ci];
                    if (to_device) slot = host[pub];
                    else host[pub] = slot;
                }
        for (int co = 0; co < kCoutReal[l]; ++co, ++pub) {
            float& slot = pad[n.b_off[l] + co];
            if (to_device) slot = host[pub];
            else host[pub] = slot;
        }
    }
    if (to_device && hipMemcpy(dev, pad.data(), n.padded * 4, hipMemcpyHostToDevice) != hipSuccess) return hipErrorUnknown;
    return hipSuccess;
}
} // namespace

hipError_t net_set_params(Net& n, const float* host) {
    if (hipStreamSynchronize(n.stream) != hipSuccess) return hipErrorUnknown;
    SF_TRY(copy_params(n, n.params, const_cast<float*>(host), true));
    SF_TRY(refresh_bf16(n));
    return hipStreamSynchronize(n.stream);
}

#undef SF_TRY

} // namespace sf::unet

// ---------------- C ABI ----------------

namespace {

using sf::unet::Net;

StableFluidsResult res(const hipError_t e) { return e == hipSuccess ? STABLE_FLUIDS_RESULT_OK : (e == hipErrorOutOfMemory ? STABLE_FLUIDS_RESULT_OUT_OF_MEMORY : STABLE_FLUIDS_RESULT_BACKEND_FAILURE); }

bool valid_c(const int c) { return c == 16 || c == 32 || c == 64; }

} // namespace

extern "C" {

STABLE_FLUIDS_API uint64_t stable_fluids_unet_param_count(void) { return sf::unet::public_count(); }

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_create_cuda(int32_t nx, int32_t ny, int32_t nz, void* stream, void** out_net) {
    if (out_net == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_net = nullptr;
    Net* n   = nullptr;
    const hipError_t e = sf::unet::net_create(nx, ny, nz, static_cast<hipStream_t>(stream), &n);
    if (e != hipSuccess) return res(e);
    *out_net = n;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_destroy_cuda(void* net) {
    if (net == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto* n = static_cast<Net*>(net);
    (void)hipStreamSynchronize(sf::unet::net_stream(*n));
    sf::unet::net_destroy(n);
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_set_params_cuda(void* net, const float* params, uint64_t count) {
    if (net == nullptr || params == nullptr || count != sf::unet::public_count()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::net_set_params(*static_cast<Net*>(net), params));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_get_params_cuda(void* net, float* params, uint64_t count) {
    if (net == nullptr || params == nullptr || count != sf::unet::public_count()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& n = *static_cast<Net*>(net);
    if (hipStreamSynchronize(n.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::copy_params(n, n.params, params, false));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_get_grads_cuda(void* net, float* grads, uint64_t count) {
    if (net == nullptr || grads == nullptr || count != sf::unet::public_count()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& n = *static_cast<Net*>(net);
    if (hipStreamSynchronize(n.stream) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::copy_params(n, n.grads, grads, false));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_forward_cuda(void* net, const float* input, float* output) {
    if (net == nullptr || input == nullptr || output == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& n = *static_cast<Net*>(net);
    if (sf::unet::launch_pack_input(n.stream, input, nullptr, 1.0f, n.a0, n.vox[0]) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (sf::unet::net_forward_packed(n, false) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(hipMemcpyAsync(output, n.q, n.vox[0] * 4, hipMemcpyDeviceToDevice, n.stream));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_backward_cuda(void* net, const float* input, const float* target, float weight, uint32_t flags, float* loss) {
    if (net == nullptr || input == nullptr || target == nullptr || !(weight >= 0.0f) || (flags & ~STABLE_FLUIDS_UNET_LOSS_CENTERED)) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& n = *static_cast<Net*>(net);
    if (sf::unet::launch_pack_input(n.stream, input, nullptr, 1.0f, n.a0, n.vox[0]) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (sf::unet::net_forward_packed(n, true) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::net_backward(n, target, weight, (flags & STABLE_FLUIDS_UNET_LOSS_CENTERED) != 0, loss));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_adam_step_cuda(void* net, float lr, float beta1, float beta2, float eps) {
    if (net == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::net_adam(*static_cast<Net*>(net), lr, beta1, beta2, eps));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_conv3_cuda(void* stream, int32_t cin, int32_t cout, const uint16_t* x, const uint16_t* w, const float* bias, float* y, int32_t relu, int32_t nx, int32_t ny, int32_t nz) {
    if (!valid_c(cin) || !valid_c(cout) || x == nullptr || w == nullptr || y == nullptr || nx <= 0 || ny <= 0 || nz <= 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::launch_conv3(static_cast<hipStream_t>(stream), cin, cout, x, w, bias, nullptr, y, nullptr, relu ? sf::unet::kEpiRelu : sf::unet::kEpiNone, sf::Dims{nx, ny, nz}));
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_unet_conv3_wgrad_cuda(void* stream, int32_t cin, int32_t cout, const uint16_t* x, const uint16_t* g, float* dw, float* db, int32_t nx, int32_t ny, int32_t nz) {
    if (!valid_c(cin) || !valid_c(cout) || x == nullptr || g == nullptr || dw == nullptr || nx <= 0 || ny <= 0 || nz <= 0) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    return res(sf::unet::launch_conv3_wgrad(static_cast<hipStream_t>(stream), cin, cout, x, g, dw, db, sf::Dims{nx, ny, nz}));
}

} // extern "C"
