// Host-side copies the library keeps to rebuild the BSR pattern: the
// connection list and the wells as the caller last described them.
#pragma once

#include <stdint.h>

#include <map>
#include <vector>

namespace ocpg {

// Timer slots of ocpg_timing_get (names in ocpgpu_capi.cpp).
enum TimerSlot {
    T_FLASH = 0, T_KRPC, T_RES, T_ASM, T_INIT, T_RESTORE, T_STAB, T_SPLIT, T_FINISH, T_TAIL, T_FIXUP, T_UPDASM,
    T_TAILNR, T_PACK, T_MLP, kTimerSlots
};

struct WellsHost {
    int nw = 0;
    std::vector<int32_t> type, open, mode, unweight, perfStart, perfCell;
    std::vector<double> bhp, ctrlBHP, ctrlRate, injZi, prodPW, perfWI, perfMult, perfDG, perfXi;
    std::vector<int32_t> hasPW;  // per well: production phase weights given, with a positive sum
    double Psurf = 0, Tsurf = 0;
};

// A serving pool (ocpg_pool_init): realisations placed at reserved ranges of
// one context.  The realisation at cell base b owns cells [b, b + nb),
// connection slots [b * maxNb, ...) and BSR blocks [b * (1 + maxNb), ...),
// so placing one never moves another.
struct PoolSlot {
    int64_t nb = 0, nconn = 0, nnzb = 0;
    std::vector<int32_t> IA;  // the realisation's own row starts (nb + 1, from 0)
    bool active = false;
    uint64_t offAt = 0;       // iteration launches queued when it was last switched off
};
struct PoolHost {
    int64_t capacity = 0;   // 0: not a pool
    int maxNb = 0;          // most neighbours of a cell
    int connPerCell = 0;    // connection slots reserved per cell: ceil(maxNb / 2)
    std::map<int64_t, PoolSlot> slots;  // by cell base
};

struct FimHost {
    std::vector<int32_t> bId, eId;
    WellsHost wells;
    PoolHost pool;
};

}  // namespace ocpg
