// One instantiation of the flash kernels; compiled once per supported
// component count with -DOCPG_NC=<NC> so the instantiations build in parallel.
#include "flash_impl.hpp"
#include "well_kernels.hpp"

#ifndef OCPG_NC
#error "compile with -DOCPG_NC=<number of hydrocarbon components>"
#endif

namespace ocpg {
template hipError_t launchFlashT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, DevFlashCounters*,
                                           const double*, hipStream_t, hipEvent_t*, const FlashSide*);
template hipError_t launchFlashSideFinishT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, hipStream_t);
template hipError_t launchWellFluxT<OCPG_NC>(const DevEos*, const DevBulk&, FimDevice&, WellDevice&, hipStream_t);
template hipError_t launchWellAssembleT<OCPG_NC>(const DevBulk&, FimDevice&, WellDevice&, double, hipStream_t);
}  // namespace ocpg
