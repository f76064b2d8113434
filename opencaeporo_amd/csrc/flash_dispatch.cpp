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
                       const double* zin, hipStream_t s, hipEvent_t* ev, const FlashSide* side, bool finish)
{
#define X(N) \
    if (NC == N) return launchFlashT<N>(eos, B, W, ctr, zin, s, ev, side, finish);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchFlashFinishN(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, int init,
                              hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchFlashFinishNT<N>(eos, B, W, init, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}


hipError_t launchFlashSideFinishN(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchFlashSideFinishNT<N>(eos, B, W, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchFlashSideFinish(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchFlashSideFinishT<N>(eos, B, W, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

CellLists flashSideCells(const FlashWork& W)
{
    const int ids[kSideLists] = {LIST_NP1B, LIST_NP2B, LIST_NP1C, LIST_NP2C, LIST_NP1D, LIST_NP2D};
    CellLists L;
    for (int q = 0; q < kSideLists; q++) {
        L.list[q] = W.list[ids[q]];
        L.count[q] = W.count + ids[q];
    }
    return L;
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

hipError_t launchWellPrepare(int NC, const DevEos* eos, const DevBulk& B, WellDevice& W, int what, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchWellPrepareT<N>(eos, B, W, what, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchWellCheck(int NC, const DevBulk& B, WellDevice& W, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchWellCheckT<N>(B, W, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchPhaseDensity(int NC, const DevEos* eos, double P, double TF, const double* zi, int phase, double* out,
                              hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchPhaseDensityT<N>(eos, P, TF, zi, phase, out, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace ocpg
