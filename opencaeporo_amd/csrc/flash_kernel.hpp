// Launchers of the flash kernels (flash_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.hpp"

namespace ocpg {

constexpr int kFlashBlock = 128;

bool flashSupportsNC(int nc);
// FlashFIM over every cell (zin == NULL) or InitFlashFIM from the feed zin
// (SoA, leading dimension B.ld).
hipError_t launchFlash(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                       const double* zin, hipStream_t s);

}  // namespace ocpg

namespace ocpg {
template <int NC>
hipError_t launchFlashT(const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                        const double* zin, hipStream_t s);
}  // namespace ocpg
