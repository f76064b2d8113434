// One instantiation of the IsoT_FIMn finishing stage; compiled once per
// supported component count with -DOCPG_NC=<NC>.
#include "flash_finish_n.hpp"

#ifndef OCPG_NC
#error "compile with -DOCPG_NC=<number of hydrocarbon components>"
#endif

namespace ocpg {
template hipError_t launchFlashFinishNT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, int, hipStream_t);
template hipError_t launchFlashSideFinishNT<OCPG_NC>(const DevEos*, const DevBulk&, const FlashWork&, hipStream_t);
}  // namespace ocpg
