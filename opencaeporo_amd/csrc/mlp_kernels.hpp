// The learned form's fused inference, for the library's own callers
// (mlp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocpg {

// The network on the cells [0, nb) of a SoA cell state (P, T, Ni with
// leading dimension ld): features P / 5000, T / 600, z_i = Ni_i / sum |Ni|
// over the NC hydrocarbon components; wb = W1, b1, .., W4, b4 on the
// device (W_l (in x out) row-major, hidden width 64); Y (nb x ldY floats):
// ln K_1..ln K_NC and the stability logit per cell.
hipError_t launchMlpForwardCells(hipStream_t s, int64_t nb, int64_t ld, int NC, const double* P, const double* T,
                                 const double* Ni, const float* const wb[8], float* Y, int ldY);

}  // namespace ocpg
