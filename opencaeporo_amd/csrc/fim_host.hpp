// Host-side copies the library keeps to rebuild the BSR pattern: the
// connection list and the wells as the caller last described them.
#pragma once

#include <stdint.h>

#include <vector>

namespace ocpg {

// Timer slots of ocpg_timing_get (names in ocpgpu_capi.cpp).
enum TimerSlot {
    T_FLASH = 0, T_KRPC, T_RES, T_ASM, T_INIT, T_RESTORE, T_STAB, T_SPLIT, T_FINISH, T_TAIL, T_FIXUP, T_UPDASM,
    T_TAILNR, T_PACK, kTimerSlots
};

struct WellsHost {
    int nw = 0;
    std::vector<int32_t> type, open, mode, unweight, perfStart, perfCell;
    std::vector<double> bhp, ctrlBHP, ctrlRate, injZi, prodPW, perfWI, perfMult, perfDG, perfXi;
    std::vector<int32_t> hasPW;  // per well: production phase weights given, with a positive sum
    double Psurf = 0, Tsurf = 0;
};

struct FimHost {
    std::vector<int32_t> bId, eId;
    WellsHost wells;
};

}  // namespace ocpg
