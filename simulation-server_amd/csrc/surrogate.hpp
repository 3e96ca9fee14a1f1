// surrogate.hpp -- the learned spring surrogate shared between the MLP
// kernels (surrogate.hip) and the simulation (sim.hip, step.hip).
#pragma once

#include "bcs_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bcs {

// 294 -> 256 -> 256 -> 256 -> 126, GELU between layers.  The input is padded
// to 320 columns (a multiple of the 32-deep MFMA step) and the output to 128
// rows of the last weight matrix (zero rows).
constexpr int kSurrIn = 294, kSurrInPad = 320, kSurrHidden = 256, kSurrOut = 126, kSurrOutPad = 128;
constexpr int kSurrPpc = 42;  // particles of the cells the surrogate serves

// float -> bf16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Device parameters of one MLP: bf16 copies of the weight matrices ([out][in],
// row-major) for the GEMMs, fp32 biases, and the feature/output scales.
struct SurrNet {
    const uint16_t* W[4];
    const float* b[4];
    float pos_scale, vel_scale, out_scale;
};

// One inference pass over n rows of bf16 features (ld kSurrInPad) into
// out (fp32, ld kSurrOutPad, in the network's normalised units), with two
// bf16 activation buffers of n x 256.
void surrogate_infer(const SurrNet& net, const uint16_t* X, int32_t n, uint16_t* act0, uint16_t* act1, float* out,
                     hipStream_t st);

}  // namespace bcs

struct bcs_surrogate;
// The device network of a surrogate (valid until its weights change).
int bcs_surrogate_net(bcs_surrogate* s, bcs::SurrNet* out, int* device);
