// Launchers of the flash kernels (flash_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.hpp"
#include "fim_kernels.hpp"
#include "well_state.hpp"

namespace ocpg {

constexpr int kFlashBlock = 128;

bool flashSupportsNC(int nc);

// Side streams and events of the overlapped flash (launchFlash with side
// non-NULL).  Streams at the default priority.
struct FlashSide {
    hipStream_t sB, sC, sD;
    hipEvent_t classified, sliced, newton;  // recorded on the main stream
    hipEvent_t doneB, doneC, doneD;         // recorded on sB / sC / sD: the caller waits for all
};

// FlashFIM over every cell (zin == NULL) or InitFlashFIM from the feed zin
// (SoA, leading dimension B.ld).  With ev non-NULL, begin/end event pairs
// are recorded: ev[0..1] around the stability test (side stream B's work
// with side non-NULL), ev[2..3] the split, ev[4..5] the finishing stage and,
// with side non-NULL, ev[6..7] side stream C's work, ev[8..9] side stream
// D's.  With side non-NULL the slow cells continue on the side streams (see
// flash_impl.hpp): the caller makes s wait for side->doneB, doneC and doneD
// and calls launchFlashSideFinish before anything reads those cells.
// With finish false the finishing stage of the main stream's cells is left
// out: the IsoT_FIMn flash runs its own (launchFlashFinishN) on the same
// lists, and after the join launchFlashSideFinishN for the side streams'.
hipError_t launchFlash(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                       const double* zin, hipStream_t s, hipEvent_t* ev = nullptr, const FlashSide* side = nullptr,
                       bool finish = true);
// IsoT_FIMn's finishing stage (FlashFIMn / InitFlashFIMn after the phase
// equilibrium, flash_finish_n.hpp) for the cells the serial flash finished.
hipError_t launchFlashFinishN(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, int init,
                              hipStream_t s);
hipError_t launchFlashSideFinishN(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s);
// After the join: the finishing stage of the side streams' cells, and
// those cells.
hipError_t launchFlashSideFinish(int NC, const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s);
CellLists flashSideCells(const FlashWork& W);
// Well::CalTrans + CalFlux + CalProdWeight and the well rows of CalRes.
hipError_t launchWellFlux(int NC, const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s);
// The well blocks of AssembleMatWellsNew (after the bulk assembly).
hipError_t launchWellAssemble(int NC, const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s);
// AllWells::PrepareWell (what = 0) or the well part of ResetToLastTimeStep
// (what = 1); AllWells::CheckP into W.status[0]; MixtureComp::XiPhase /
// RhoPhase of one feed into out[0], out[1] (device memory).
hipError_t launchWellPrepare(int NC, const DevEos* eos, const DevBulk& B, WellDevice& W, int what, hipStream_t s);
hipError_t launchWellCheck(int NC, const DevBulk& B, WellDevice& W, hipStream_t s);
hipError_t launchPhaseDensity(int NC, const DevEos* eos, double P, double TF, const double* zi, int phase, double* out,
                              hipStream_t s);

}  // namespace ocpg

namespace ocpg {
template <int NC>
hipError_t launchFlashT(const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                        const double* zin, hipStream_t s, hipEvent_t* ev, const FlashSide* side, bool finish);
template <int NC>
hipError_t launchFlashFinishNT(const DevEos* eos, const DevBulk& B, const FlashWork& W, int init, hipStream_t s);
template <int NC>
hipError_t launchFlashSideFinishNT(const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s);
template <int NC>
hipError_t launchFlashSideFinishT(const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s);

template <int NC>
hipError_t launchWellFluxT(const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s);
template <int NC>
hipError_t launchWellAssembleT(const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s);
template <int NC>
hipError_t launchWellPrepareT(const DevEos* eos, const DevBulk& B, WellDevice& W, int what, hipStream_t s);
template <int NC>
hipError_t launchWellCheckT(const DevBulk& B, WellDevice& W, hipStream_t s);
template <int NC>
hipError_t launchPhaseDensityT(const DevEos* eos, double P, double TF, const double* zi, int phase, double* out,
                               hipStream_t s);
}  // namespace ocpg
