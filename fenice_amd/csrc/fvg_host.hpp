// Host-side pieces shared by the contexts of the C ABI (fvg_api.hip) and the 1-D ensemble (fvg_batch.hip).
#pragma once

#include "../../include/fvg.h"
#include "fvg_kernels.hpp"

namespace fvg {

// Blocks of the first pass of the conserved totals (k_totals_partial): the sums are a fixed-order tree over this
// many per-block partial sums, so every caller that wants the same bits launches the same count.
constexpr int kTotalsBlocks = 1024;

// Mesh, numerics and viscous scalars of the kernels for a finalized configuration (buffers are left to the caller).
void fill_kernel_params(const fvg_config& cfg, KP& p);

// The network's parameter vector (W1 32x15, b1, W2 32x32, b2, w3, b3; 1601 floats) as the device reads it.
MlpDev pack_mlp(const float* params, float threshold);

}   // namespace fvg
