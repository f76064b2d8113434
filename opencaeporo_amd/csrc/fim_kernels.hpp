// FIM property, residual and Jacobian kernels (fim_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.hpp"
#include "well_state.hpp"

namespace ocpg {

// Static connection / BSR structure on the device.
struct FimDevice {
    int64_t nb = 0;        // bulk rows
    int64_t nconn = 0;
    int64_t nnzb = 0;      // blocks in the BSR (bulk part)
    int ncol = 0;
    // connection list (BulkConn order)
    int32_t* bId = nullptr;
    int32_t* eId = nullptr;
    double* area = nullptr;
    // BSR pattern of AssembleMatBulksNew: row n = [n, neighbours ascending]
    int32_t* IA = nullptr;      // nb + 1
    int32_t* JA = nullptr;      // nnzb
    int32_t* slotConn = nullptr;    // connection of each off-diagonal slot (-1 on the diagonal)
    int32_t* slotMirror = nullptr;  // index of slot (JA[k], row) for slot k = (row, JA[k])
    int32_t* bulkEnd = nullptr;     // per bulk row: end of its bulk columns (well columns follow)
    // per-connection upwinding written by CalRes and read by the assembly
    int32_t* upblock = nullptr;     // nconn * np
    double* upRho = nullptr;        // nconn * np
    double* upVel = nullptr;        // nconn * np
    // outputs
    int64_t nrow = 0;               // rows: bulk cells, then open wells
    double* res = nullptr;          // nrow * ncol, cell-major (LinearSystem rhs layout), negated
    double* val = nullptr;          // nnzb * ncol * ncol, row-major blocks
    // IsoT_FIMn (fimn set): the right-hand side AssembleMat builds, nrow *
    // ncol: the bulk and perforation terms of the flash's correction, then
    // + res (LinearSystem::AssembleRhsAccumulate)
    int fimn = 0;
    double* rhsN = nullptr;
    // per-cell records the assembly reads, packed (AoS) by launchPackRecords
    double* rec = nullptr;          // nb * recStride: a connection side's cell values
    double* own = nullptr;          // nb * ownStride: a row's own dSec_dPri, dPcj_dS, accumulation terms
    double* normPart = nullptr;     // per-block partial maxima of CalRes norms
    int64_t* normIdx = nullptr;
    int normBlocks = 0;
    size_t bytes = 0;
    size_t patternBytes = 0;        // part of bytes held by the pattern arrays, res and val
};

void fimFree(FimDevice& f);
size_t fimBytes(const FimDevice& f);

hipError_t launchKrPcRock(const DevBulk& B, const DevSat* sat, const DevRock& rock, bool krpc, bool rockOn,
                          hipStream_t s);
// CalRes over every cell; with norms, also the per-block maxima of the
// relative residuals (normPart / normIdx).
hipError_t launchCalRes(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, hipStream_t s,
                        bool norms = true);
hipError_t launchResNorm(int NC, const DevBulk& B, FimDevice& F, hipStream_t s);
hipError_t launchAssemble(int NC, const DevBulk& B, FimDevice& F, double dt, hipStream_t s);
// IsoT_FIMn::AssembleMatBulksNew (F.fimn): the blocks, and the bulk terms of
// the right-hand side into F.rhsN; then, after the wells' blocks,
// launchRhsAccumulate adds res (rhsN = rhsN + res, row by row).
hipError_t launchAssembleN(int NC, const DevBulk& B, FimDevice& F, double dt, hipStream_t s);
hipError_t launchRhsAccumulate(FimDevice& F, hipStream_t s);
// Doubles per cell of the packed records (FimDevice::rec, ::own), and their
// packing from the SoA state (every cell, or the cells of two lists).
int64_t recordDoubles(int NC);
int64_t ownDoubles(int NC);
hipError_t launchPackRecords(int NC, const DevBulk& B, FimDevice& F, hipStream_t s);
// CalRes (no norms) and the record packing of every cell in one launch.
hipError_t launchCalResPack(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, hipStream_t s);
hipError_t launchPackRecordsList(int NC, const DevBulk& B, FimDevice& F, const CellLists& L, hipStream_t s);

// IsoT_FIM::GetSolution on the device (fim_kernels.hip) and the Newton-
// change bookkeeping of the next PassFlashValue / Bulk::CalNRdSmax.
struct NewtonDevice {
    int64_t nb = 0, ld = 0;
    double* dSNR = nullptr;     // np per cell (SoA): S before the update; after ocpg_nr_change, its change
    double* dSNRP = nullptr;    // np per cell: the update's saturation changes
    uint8_t* pexOld = nullptr;  // np per cell: phaseExist before the flash that follows
    double* part = nullptr;     // per block: 2 (value, magnitude) pairs of first maxima
    int64_t* partIdx = nullptr; // per block: their indices
    int blocks = 0;
    double* u = nullptr;        // staging of a solution vector from the host
    int64_t uLen = 0;
    bool valid = false;         // a GetSolution happened since the last flash bookkeeping
    int32_t* chk = nullptr;     // flags of ocpg_check_bulk
    double* chg = nullptr;      // per block: 4 maxima of ocpg_max_change
    void* mem = nullptr;
};
hipError_t launchGetSolution(int NC, const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                             const double* u, double dSmaxlim, hipStream_t s);
// IsoT_FIMn::GetSolution (src/OCPFluidMethod.cpp:3847-3990); scm: the
// critical saturation of each phase (Bulk::satcm).
hipError_t launchGetSolutionN(int NC, const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                              const double* u, double dSmaxlim, const double* scm, hipStream_t s);
hipError_t launchNRChange(const DevBulk& B, NewtonDevice& N, hipStream_t s);
// Bulk::CheckNi / CheckP into *flag (bit 1 / bit 0; *flag zeroed by the
// caller) and Bulk::CalMaxChange's per-block maxima against the last-step
// state L (dP, relative dN, dS, relative volume error).
hipError_t launchBulkCheck(const DevBulk& B, int nc, int32_t* flag, hipStream_t s);
hipError_t launchMaxChange(const DevBulk& B, const DevBulk& L, int nc, double* part, int blocks, hipStream_t s);

// The overlapped Newton iteration's recomputation (fim_kernels.hip): rows
// touched by the cells the flash finished on its side streams.
struct FixupRows {
    uint32_t* mark = nullptr;     // nb flags, all zero between calls
    int32_t* resRows = nullptr;   // nb: rows whose CalRes is recomputed
    int32_t* asmRows = nullptr;   // nb: rows whose assembly is recomputed
    int32_t* counts = nullptr;    // [0] resRows, [1] asmRows
    void* mem = nullptr;
    int64_t nb = 0;
};
// kr/Pc + rock of the cells in L.
hipError_t launchKrPcRockList(const DevBulk& B, const DevSat* sat, const DevRock& rock, const CellLists& L,
                              hipStream_t s);
// The rows to recompute for the side cells in L (and the perforated cells).
hipError_t launchFixupRows(const FimDevice& F, const WellDevice& W, const CellLists& L, const FixupRows& R,
                           hipStream_t s);
hipError_t launchCalResRows(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt,
                            const int32_t* rows, const int32_t* count, hipStream_t s);
hipError_t launchAssembleRows(int NC, const DevBulk& B, FimDevice& F, double dt, const int32_t* rows,
                              const int32_t* count, hipStream_t s);

// Serving pool: mask[base, base + n) = v; and a copy of fields (each
// rows x ld entries of elem bytes, the SoA layout) for the cells whose mask
// is set.
constexpr int kMaxMaskedCopy = 12;
struct MaskedCopy {
    const char* src[kMaxMaskedCopy];
    char* dst[kMaxMaskedCopy];
    int elem[kMaxMaskedCopy];
    int rows[kMaxMaskedCopy];
    int nfields;
};
hipError_t launchMaskFill(uint8_t* mask, int64_t base, int64_t n, uint8_t v, hipStream_t s);
// A realisation's pattern moved by dCell cells (its connection slots by
// dConn, its blocks by dBlock): IA and bulkEnd (nb), JA, slotConn (where
// >= 0) and slotMirror (nnzb), bId and eId (nconn).
hipError_t launchPoolRebase(int32_t* IA, int32_t* bulkEnd, int64_t nb, int32_t* JA, int32_t* slotConn,
                            int32_t* slotMirror, int64_t nnzb, int32_t* bId, int32_t* eId, int64_t nconn,
                            int64_t dCell, int64_t dConn, int64_t dBlock, hipStream_t s);
// One strided piece of device memory a realisation image holds: rows rows
// of width bytes, pitch bytes apart.
struct ImageSeg {
    char* dev;
    size_t pitch, width, rows;
};
hipError_t launchMaskedCopy(const uint8_t* mask, int64_t nb, int64_t ld, const MaskedCopy& d, hipStream_t s);

}  // namespace ocpg
