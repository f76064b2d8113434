// One instantiation of the flash kernels; compiled once per supported
// component count with -DOCPG_NC=<NC> so the instantiations build in parallel.
#include "flash_impl.hpp"

#ifndef OCPG_NC
#error "compile with -DOCPG_NC=<number of hydrocarbon components>"
#endif

namespace ocpg {
template hipError_t launchFlashT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, DevFlashCounters*,
                                           const double*, hipStream_t, hipEvent_t*, const FlashSide*, bool);
template hipError_t launchFlashSideFinishT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, hipStream_t);
}  // namespace ocpg
