// Device-side wells (AllWells / Well / WellPerf of the reference, include/
// Well.hpp, include/WellPerf.hpp) and their BSR slots.  Per-well arrays have
// nw entries, per-perforation arrays nperf entries in well-major order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocpg {

struct WellDevice {
    int nw = 0;       // wells
    int nwOpen = 0;   // open wells (each owns one BSR row after the bulk rows)
    int nperf = 0;    // perforations of all wells
    int npcell = 0;   // distinct cells perforated by open wells
    double Psurf = 0, Tsurf = 0;
    // description and per-step state
    int32_t *type = nullptr, *open = nullptr, *mode = nullptr, *unweight = nullptr;
    int32_t *perfStart = nullptr, *perfCell = nullptr;
    int32_t* row = nullptr;  // per well: BSR row - nb of an open well, -1 when shut
    double *bhp = nullptr, *ctrlBHP = nullptr, *ctrlRate = nullptr;
    double *injZi = nullptr, *prodPW = nullptr;  // nw*nc, nw*3
    double *perfWI = nullptr, *perfMult = nullptr, *perfDG = nullptr, *perfXi = nullptr;
    double* lbhp = nullptr;  // nw: Well::lbhp, BHP at the last time step
    // AllWells::PrepareWell / CheckP inputs and state (ocpg_set_well_setup)
    bool hasSetup = false;
    int32_t* initMode = nullptr;   // nw: WellOpt::initOptMode (0 BHP, 1 rate)
    int32_t* injPhase = nullptr;   // nw: WellOpt::injProdPhase of injectors (1 gas, 2 water)
    double* wellDepth = nullptr;   // nw: Well::depth
    double* injTemp = nullptr;     // nw: WellOpt::injTemp (F)
    int32_t* perfState = nullptr;  // nperf: WellPerf::state (1 open, 0 closed)
    double* perfP = nullptr;       // nperf: WellPerf::P as the last CaldG / CalFlux left it
    int32_t* status = nullptr;     // [0]: AllWells::CheckP of the last ocpg_check_wells
    double* perfTmp = nullptr;     // nperf: per-interval pressure increments of CaldG
    // Well::CalTrans / CalFlux / CalProdWeight results
    double* perfTransj = nullptr;  // nperf*3
    double* perfQi = nullptr;      // nperf*nc
    double* wellQi = nullptr;      // nw*nc
    double* prodWeight = nullptr;  // nw
    double* wellRelRes = nullptr;  // nw: |res / maxRate| of an open rate well, else -1
    // cells -> their perforations (open wells, in well then perforation order)
    int32_t* cellPerfPtr = nullptr;  // nb+1
    int32_t* cellPerfIdx = nullptr;
    int32_t* pcell = nullptr;        // npcell perforated cells
    // assembly scratch and slots
    double* perfBlock = nullptr;  // nperf*bsize: dt (dQdXpB + dQdXsB dSec) of each perforation
    double* perfDQW = nullptr;    // nperf*nc: column 0 of dQdXpW
    // IsoT_FIMn (fimn set): each perforation's right-hand-side term
    // -dQdXsB res_n (nperf*ncol; zero where the cell has one hydrocarbon phase)
    int fimn = 0;
    double* perfRhs = nullptr;
    int32_t* slotBW = nullptr;    // per perforation: slot of block (cell, well) in the cell's row
    int32_t* slotWB = nullptr;    // per perforation: slot of block (well, cell) in the well's row
    void* mem = nullptr;          // one allocation holding every array above
    size_t bytes = 0;
};

}  // namespace ocpg
