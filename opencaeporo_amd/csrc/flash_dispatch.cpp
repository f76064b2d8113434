// Dispatch from a runtime component count to the compiled flash
// instantiations.  OCPG_NC_LIST is the build's list, e.g. X(6) X(10).
#include "flash_kernel.hpp"

#ifndef OCPG_NC_LIST
#define OCPG_NC_LIST(X) X(6)
#endif

namespace ocpg {

bool flashSupportsNC(int nc)
{
#define X(N) if (nc == N) return true;
    OCPG_NC_LIST(X)
#undef X
    return false;
}

hipError_t launchFlash(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                       const double* zin, hipStream_t s, hipEvent_t* ev, const FlashSide* side)
{
#define X(N) \
    if (NC == N) return launchFlashT<N>(eos, B, W, ctr, zin, s, ev, side);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchFlashTailFinish(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchFlashTailFinishT<N>(eos, B, W, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

void flashTailListsOf(const FlashWork& W, const int32_t** lists, const int32_t** counts)
{
    lists[0] = W.list[LIST_NP1T];
    lists[1] = W.list[LIST_NP2T];
    counts[0] = W.count + LIST_NP1T;
    counts[1] = W.count + LIST_NP2T;
}

hipError_t launchWellFlux(int NC, const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchWellFluxT<N>(eos, B, F, W, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchWellAssemble(int NC, const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchWellAssembleT<N>(B, F, W, dt, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace ocpg
