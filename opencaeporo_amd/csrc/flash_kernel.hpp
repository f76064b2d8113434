// Launchers of the flash kernels (flash_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.hpp"

namespace ocpg {

constexpr int kFlashBlock = 256;

bool flashSupportsNC(int nc);
hipError_t launchFlashFim(int NC, const DevEos* eos, const DevBulk& B, DevFlashCounters* ctr, hipStream_t s);
hipError_t launchInitFlash(int NC, const DevEos* eos, const DevBulk& B, const double* z, int64_t zld,
                           hipStream_t s);

}  // namespace ocpg

namespace ocpg {
template <int NC>
hipError_t launchFlashT(const DevEos* eos, const DevBulk& B, DevFlashCounters* ctr, hipStream_t s);
template <int NC>
hipError_t launchInitFlashT(const DevEos* eos, const DevBulk& B, const double* z, int64_t zld, hipStream_t s);
}  // namespace ocpg
