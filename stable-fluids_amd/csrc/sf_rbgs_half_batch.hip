// sf_rbgs_half_batch.hip -- batched launches of the black-cell RBGS passes
// (sf_rbgs_half.hpp): one launch runs the same pass for many simulations.
#include "sf_rbgs_half.hpp"

namespace sf {

hipError_t launch_rbgs_half_batched(hipStream_t s, const SolveKind solve, const int in_mode, const int out_mode, const int tile, const HalfPassArgs* descs, const BatchItem* items, const uint32_t n_items) {
    if (n_items == 0) return hipSuccess;
    half_dispatch(solve, in_mode, out_mode, tile, [&]<int SO, int IN, int OUT, int TXP, int W, int RW>() { k_rbgs_half_batched<SO, IN, OUT, TXP, W, RW><<<n_items, 64 * W, 0, s>>>(descs, items); });
    return hipGetLastError();
}

} // namespace sf
