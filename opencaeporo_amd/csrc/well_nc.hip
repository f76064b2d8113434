// One instantiation of the well kernels (well_kernels.hpp); compiled once per
// supported component count with -DOCPG_NC=<NC>, apart from the flash
// kernels so that the two build in parallel.
#include "well_kernels.hpp"

#ifndef OCPG_NC
#error "compile with -DOCPG_NC=<number of hydrocarbon components>"
#endif

namespace ocpg {
template hipError_t launchWellFluxT<OCPG_NC>(const DevEos*, const DevBulk&, FimDevice&, WellDevice&, hipStream_t);
template hipError_t launchWellAssembleT<OCPG_NC>(const DevBulk&, FimDevice&, WellDevice&, double, hipStream_t);
template hipError_t launchWellPrepareT<OCPG_NC>(const DevEos*, const DevBulk&, WellDevice&, int, hipStream_t);
template hipError_t launchWellCheckT<OCPG_NC>(const DevBulk&, WellDevice&, hipStream_t);
template hipError_t launchPhaseDensityT<OCPG_NC>(const DevEos*, double, double, const double*, int, double*,
                                                  hipStream_t);
}  // namespace ocpg
