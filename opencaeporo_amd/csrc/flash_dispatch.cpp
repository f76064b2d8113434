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

hipError_t launchFlashFim(int NC, const DevEos* eos, const DevBulk& B, DevFlashCounters* ctr, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchFlashT<N>(eos, B, ctr, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchInitFlash(int NC, const DevEos* eos, const DevBulk& B, const double* z, int64_t zld,
                           hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchInitFlashT<N>(eos, B, z, zld, s);
    OCPG_NC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace ocpg
