// Launchers of the flash kernels (flash_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.hpp"
#include "fim_kernels.hpp"
#include "well_state.hpp"

namespace ocpg {

constexpr int kFlashBlock = 128;

bool flashSupportsNC(int nc);
// FlashFIM over every cell (zin == NULL) or InitFlashFIM from the feed zin
// (SoA, leading dimension B.ld).  With ev non-NULL, ev[0..5] are recorded
// around the stability, split and finishing stages (begin, end pairs).
hipError_t launchFlash(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                       const double* zin, hipStream_t s, hipEvent_t* ev = nullptr);
// Well::CalTrans + CalFlux + CalProdWeight and the well rows of CalRes.
hipError_t launchWellFlux(int NC, const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s);
// The well blocks of AssembleMatWellsNew (after the bulk assembly).
hipError_t launchWellAssemble(int NC, const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s);

}  // namespace ocpg

namespace ocpg {
template <int NC>
hipError_t launchFlashT(const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                        const double* zin, hipStream_t s, hipEvent_t* ev);
template <int NC>
hipError_t launchWellFluxT(const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s);
template <int NC>
hipError_t launchWellAssembleT(const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s);
}  // namespace ocpg
