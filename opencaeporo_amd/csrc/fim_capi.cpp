// C-ABI of the FIM half: connection upload, CalKrPc/CalRock, CalRes and
// AssembleMat, and the hand-off of the assembled system in the reference's
// LinearSystem / FASP dBSRmat layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ocpgpu.h"
#include "device_state.hpp"
#include "fim_host.hpp"
#include "fim_kernels.hpp"
#include "flash_kernel.hpp"

namespace ocpg {
DevBulk& ctxBulk(ocpg_ctx* c);
bool ctxFimn(ocpg_ctx* c);
const double* ctxScm(ocpg_ctx* c);
FimDevice& ctxFim(ocpg_ctx* c);
WellDevice& ctxWells(ocpg_ctx* c);
FimHost& ctxFimHost(ocpg_ctx* c);
const DevEos* ctxEos(ocpg_ctx* c);
hipStream_t ctxStream(ocpg_ctx* c);
int ctxNC(ocpg_ctx* c);
int ctxDevice(ocpg_ctx* c);
const DevSat* ctxSat(ocpg_ctx* c);
const DevRock& ctxRock(ocpg_ctx* c);
void ctxTimerBegin(ocpg_ctx* c, int slot);
void ctxTimerEnd(ocpg_ctx* c, int slot);
int fimReadNorms(ocpg_ctx* c, ocpg_resnorm* norms);
NewtonDevice& ctxNewton(ocpg_ctx* c);
void ctxDropGraph(ocpg_ctx* c);
const DevBulk& ctxLastBulk(ocpg_ctx* c);
bool ctxLastValid(ocpg_ctx* c);
uint8_t* ctxEnsureMask(ocpg_ctx* c);
int ctxDropMask(ocpg_ctx* c);
int ctxRangeCopy(ocpg_ctx* c, int64_t base, ocpg_bulk* h, bool toDevice, bool saved);
int ctxSaveInputsRange(ocpg_ctx* c, int64_t base, int64_t nb, hipStream_t stream);
int ctxPoolTrimLast(ocpg_ctx* c);
hipStream_t ctxCopyStream(ocpg_ctx* c);
int ctxCopyAfterMain(ocpg_ctx* c);
int ctxMainAfterCopy(ocpg_ctx* c);
uint64_t ctxLaunches(ocpg_ctx* c);
uint64_t ctxLaunchesDone(ocpg_ctx* c);
void ctxImageSegments(ocpg_ctx* c, int64_t base, int64_t nb, bool state, bool fixed, std::vector<ImageSeg>& out);
}  // namespace ocpg


using namespace ocpg;

static int hipStatus2(hipError_t e)
{
    if (e == hipSuccess) return OCPG_OK;
    std::fprintf(stderr, "ocpgpu: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
    return OCPG_EHIP;
}
#define HIPCHK2(x)                                    \
    do {                                              \
        hipError_t e_ = (x);                          \
        if (e_ != hipSuccess) return hipStatus2(e_);  \
    } while (0)

template <class T>
static hipError_t devAlloc(T*& p, size_t count, size_t& total)
{
    const size_t b = std::max<size_t>(count, 1) * sizeof(T);
    total += b;
    return hipMalloc(&p, b);
}

static size_t align256w(size_t b) { return (b + 255) & ~size_t(255); }
static int ensureNewton(ocpg_ctx* c);

// A producer's WellOpt::prodPhaseWeight (3 per well) is usable for a rate
// control when given and not all zero: with zero weights the rate equation
// is a constant and the well row of the Jacobian vanishes.
static bool positiveWeights(const double* pw, int w)
{
    if (!pw) return false;
    double s = 0;
    for (int j = 0; j < 3; j++) {
        if (!(pw[(size_t)w * 3 + j] >= 0)) return false;
        s += pw[(size_t)w * 3 + j];
    }
    return s > 0;
}

// The per-step / per-iteration well fields (bhp, control, dG, perforation
// molar densities of injectors).
static void setWellState(WellsHost& WH, const ocpg_wells& w)
{
    const int nw = WH.nw;
    const int nperf = WH.perfStart[nw];
    WH.mode.assign(w.mode, w.mode + nw);
    WH.bhp.assign(w.bhp, w.bhp + nw);
    WH.ctrlBHP.assign(w.ctrlBHP, w.ctrlBHP + nw);
    WH.ctrlRate.assign(w.ctrlRate, w.ctrlRate + nw);
    WH.perfDG.assign(w.perfDG, w.perfDG + nperf);
    WH.perfXi.assign(nperf, 0.0);
    if (w.perfXi) WH.perfXi.assign(w.perfXi, w.perfXi + nperf);
}

static int uploadWellState(ocpg_ctx* c)
{
    WellDevice& W = ctxWells(c);
    const WellsHost& WH = ctxFimHost(c).wells;
    if (W.nw == 0) return OCPG_OK;
    const size_t NW = W.nw, NP = W.nperf;
    HIPCHK2(hipMemcpy(W.mode, WH.mode.data(), NW * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.bhp, WH.bhp.data(), NW * 8, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.ctrlBHP, WH.ctrlBHP.data(), NW * 8, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.ctrlRate, WH.ctrlRate.data(), NW * 8, hipMemcpyHostToDevice));
    if (NP) {
        HIPCHK2(hipMemcpy(W.perfDG, WH.perfDG.data(), NP * 8, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(W.perfXi, WH.perfXi.data(), NP * 8, hipMemcpyHostToDevice));
    }
    return OCPG_OK;
}

// The BSR pattern LinearSystem ends up with after AssembleMatBulksNew and
// AssembleMatWellsNew (include/LinearSystem.hpp:62-125):
//   bulk row n  = [n, its neighbours in the order connections reach them,
//                  then the wells perforating n (open wells in order,
//                  perforations in order)]
//   well row w  = [nb + w, the cells of its perforations in order]
// plus, for the kernels, each slot's connection and mirror slot, the end of
// each bulk row's bulk columns and the slots of every perforation's blocks.
static int buildPattern(ocpg_ctx* c)
{
    FimDevice& F = ctxFim(c);
    WellDevice& W = ctxWells(c);
    const FimHost& H = ctxFimHost(c);
    const WellsHost& WH = H.wells;
    const int64_t nb = F.nb, nconn = F.nconn;
    ctxDropGraph(c);
    const int nw = WH.nw;
    int nwOpen = 0;
    std::vector<int32_t> wrow(nw, -1);
    for (int w = 0; w < nw; w++)
        if (WH.open[w]) wrow[w] = nwOpen++;
    const int nperf = nw > 0 ? WH.perfStart[nw] : 0;
    const int64_t nrow = nb + nwOpen;
    std::vector<int64_t> cnt(nrow, 1);
    for (int64_t k = 0; k < nconn; k++) {
        cnt[H.bId[k]]++;
        cnt[H.eId[k]]++;
    }
    std::vector<int64_t> bulkCnt(cnt.begin(), cnt.begin() + nb);
    for (int w = 0; w < nw; w++) {
        if (!WH.open[w]) continue;
        for (int p = WH.perfStart[w]; p < WH.perfStart[w + 1]; p++) {
            cnt[WH.perfCell[p]]++;
            cnt[nb + wrow[w]]++;
        }
    }
    std::vector<int64_t> IA64(nrow + 1, 0);
    for (int64_t n = 0; n < nrow; n++) IA64[n + 1] = IA64[n] + cnt[n];
    if (IA64[nrow] > (int64_t)INT32_MAX) return OCPG_ETOOBIG;
    const int64_t nnzb = IA64[nrow];
    std::vector<int32_t> IA(nrow + 1), JA(nnzb), slotConn(nnzb), slotMirror(nnzb), bulkEnd(nb);
    std::vector<int64_t> fill(nrow);
    for (int64_t n = 0; n <= nrow; n++) IA[n] = (int32_t)IA64[n];
    for (int64_t n = 0; n < nrow; n++) {
        JA[IA[n]] = (int32_t)n;
        slotConn[IA[n]] = -1;
        slotMirror[IA[n]] = IA[n];
        fill[n] = IA[n] + 1;
    }
    for (int64_t n = 0; n < nb; n++) bulkEnd[n] = (int32_t)(IA[n] + bulkCnt[n]);
    for (int64_t k = 0; k < nconn; k++) {
        const int32_t b = H.bId[k], e = H.eId[k];
        const int32_t se = (int32_t)fill[e]++;  // NewOffDiag(eId, bId)
        const int32_t sb = (int32_t)fill[b]++;  // NewOffDiag(bId, eId)
        JA[se] = b;
        JA[sb] = e;
        slotConn[se] = slotConn[sb] = (int32_t)k;
        slotMirror[se] = sb;
        slotMirror[sb] = se;
    }
    std::vector<int32_t> slotBW(std::max(nperf, 1), -1), slotWB(std::max(nperf, 1), -1);
    std::vector<int32_t> cellPerfPtr(nb + 1, 0), cellPerfIdx, pcell;
    for (int w = 0; w < nw; w++) {
        if (!WH.open[w]) continue;
        const int64_t r = nb + wrow[w];
        for (int p = WH.perfStart[w]; p < WH.perfStart[w + 1]; p++) {
            const int32_t n = WH.perfCell[p];
            const int32_t sBW = (int32_t)fill[n]++;  // NewOffDiag(n, wId)
            const int32_t sWB = (int32_t)fill[r]++;  // NewOffDiag(wId, n)
            JA[sBW] = (int32_t)r;
            JA[sWB] = n;
            slotConn[sBW] = slotConn[sWB] = -2;
            slotMirror[sBW] = sWB;
            slotMirror[sWB] = sBW;
            slotBW[p] = sBW;
            slotWB[p] = sWB;
            cellPerfPtr[n + 1]++;
        }
    }
    for (int64_t n = 0; n < nb; n++) {
        if (cellPerfPtr[n + 1] > 0) pcell.push_back((int32_t)n);
        cellPerfPtr[n + 1] += cellPerfPtr[n];
    }
    cellPerfIdx.assign(std::max<int64_t>(cellPerfPtr[nb], 1), 0);
    {
        std::vector<int32_t> pos(cellPerfPtr.begin(), cellPerfPtr.end() - 1);
        for (int w = 0; w < nw; w++) {
            if (!WH.open[w]) continue;
            for (int p = WH.perfStart[w]; p < WH.perfStart[w + 1]; p++) cellPerfIdx[pos[WH.perfCell[p]]++] = p;
        }
    }
    // device arrays of the pattern (the connection arrays stay)
    const int ncol = F.ncol;
    void* old[] = {F.IA, F.JA, F.slotConn, F.slotMirror, F.bulkEnd, F.res, F.val, F.rhsN};
    for (void* q : old)
        if (q) HIPCHK2(hipFree(q));
    F.IA = F.JA = F.slotConn = F.slotMirror = F.bulkEnd = nullptr;
    F.res = F.val = F.rhsN = nullptr;
    F.fimn = ctxFimn(c) ? 1 : 0;
    size_t tot = F.bytes - F.patternBytes;
    const size_t before = tot;
    HIPCHK2(devAlloc(F.IA, nrow + 1, tot));
    HIPCHK2(devAlloc(F.JA, nnzb, tot));
    HIPCHK2(devAlloc(F.slotConn, nnzb, tot));
    HIPCHK2(devAlloc(F.slotMirror, nnzb, tot));
    HIPCHK2(devAlloc(F.bulkEnd, nb, tot));
    HIPCHK2(devAlloc(F.res, nrow * ncol, tot));
    HIPCHK2(devAlloc(F.val, (size_t)nnzb * ncol * ncol, tot));
    if (F.fimn) HIPCHK2(devAlloc(F.rhsN, nrow * ncol, tot));
    F.patternBytes = tot - before;
    F.bytes = tot;
    F.nnzb = nnzb;
    F.nrow = nrow;
    HIPCHK2(hipMemcpy(F.IA, IA.data(), IA.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.JA, JA.data(), JA.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotConn, slotConn.data(), nnzb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotMirror, slotMirror.data(), nnzb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.bulkEnd, bulkEnd.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemset(F.res, 0, nrow * ncol * sizeof(double)));
    if (F.rhsN) HIPCHK2(hipMemset(F.rhsN, 0, nrow * ncol * sizeof(double)));

    // wells: one allocation for every per-well / per-perforation array
    if (W.mem) HIPCHK2(hipFree(W.mem));
    W = WellDevice{};
    if (nw == 0) return OCPG_OK;
    const int nc = ctxNC(c) + 1;
    const size_t bsize = (size_t)ncol * ncol;
    struct Part {
        void** ptr;
        size_t bytes;
    };
    const size_t NW = nw, NP = std::max(nperf, 1);
    Part parts[] = {
        {(void**)&W.type, NW * 4}, {(void**)&W.open, NW * 4}, {(void**)&W.mode, NW * 4},
        {(void**)&W.unweight, NW * 4}, {(void**)&W.perfStart, (NW + 1) * 4}, {(void**)&W.perfCell, NP * 4},
        {(void**)&W.row, NW * 4}, {(void**)&W.bhp, NW * 8}, {(void**)&W.ctrlBHP, NW * 8},
        {(void**)&W.ctrlRate, NW * 8}, {(void**)&W.injZi, NW * nc * 8}, {(void**)&W.prodPW, NW * 3 * 8},
        {(void**)&W.perfWI, NP * 8}, {(void**)&W.perfMult, NP * 8}, {(void**)&W.perfDG, NP * 8},
        {(void**)&W.perfXi, NP * 8}, {(void**)&W.perfTransj, NP * 3 * 8}, {(void**)&W.perfQi, NP * nc * 8},
        {(void**)&W.wellQi, NW * nc * 8}, {(void**)&W.prodWeight, NW * 8}, {(void**)&W.wellRelRes, NW * 8},
        {(void**)&W.cellPerfPtr, (size_t)(nb + 1) * 4}, {(void**)&W.cellPerfIdx, cellPerfIdx.size() * 4},
        {(void**)&W.pcell, std::max<size_t>(pcell.size(), 1) * 4}, {(void**)&W.perfBlock, NP * bsize * 8},
        {(void**)&W.perfDQW, NP * nc * 8}, {(void**)&W.slotBW, NP * 4}, {(void**)&W.slotWB, NP * 4},
        {(void**)&W.lbhp, NW * 8}, {(void**)&W.initMode, NW * 4}, {(void**)&W.injPhase, NW * 4},
        {(void**)&W.wellDepth, NW * 8}, {(void**)&W.injTemp, NW * 8}, {(void**)&W.perfState, NP * 4},
        {(void**)&W.perfP, NP * 8}, {(void**)&W.status, NW * 4}, {(void**)&W.perfTmp, NP * 8},
        {(void**)&W.perfRhs, NP * ncol * 8},
    };
    size_t total = 0;
    for (const Part& q : parts) total += align256w(q.bytes);
    char* mem = nullptr;
    HIPCHK2(hipMalloc(&mem, total));
    HIPCHK2(hipMemset(mem, 0, total));
    W.mem = mem;
    W.bytes = total;
    size_t off = 0;
    for (const Part& q : parts) {
        *q.ptr = mem + off;
        off += align256w(q.bytes);
    }
    W.nw = nw;
    W.nwOpen = nwOpen;
    W.nperf = nperf;
    W.fimn = F.fimn;
    W.npcell = (int)pcell.size();
    W.Psurf = WH.Psurf;
    W.Tsurf = WH.Tsurf;
    auto up = [](void* d, const void* h, size_t b) { return b ? hipMemcpy(d, h, b, hipMemcpyHostToDevice) : hipSuccess; };
    HIPCHK2(up(W.type, WH.type.data(), NW * 4));
    HIPCHK2(up(W.open, WH.open.data(), NW * 4));
    HIPCHK2(up(W.unweight, WH.unweight.data(), NW * 4));
    HIPCHK2(up(W.perfStart, WH.perfStart.data(), (NW + 1) * 4));
    HIPCHK2(up(W.perfCell, WH.perfCell.data(), (size_t)nperf * 4));
    HIPCHK2(up(W.row, wrow.data(), NW * 4));
    HIPCHK2(up(W.injZi, WH.injZi.data(), NW * nc * 8));
    HIPCHK2(up(W.prodPW, WH.prodPW.data(), NW * 3 * 8));
    HIPCHK2(up(W.perfWI, WH.perfWI.data(), (size_t)nperf * 8));
    HIPCHK2(up(W.perfMult, WH.perfMult.data(), (size_t)nperf * 8));
    HIPCHK2(up(W.cellPerfPtr, cellPerfPtr.data(), (size_t)(nb + 1) * 4));
    HIPCHK2(up(W.cellPerfIdx, cellPerfIdx.data(), cellPerfIdx.size() * 4));
    HIPCHK2(up(W.pcell, pcell.data(), pcell.size() * 4));
    HIPCHK2(up(W.slotBW, slotBW.data(), (size_t)nperf * 4));
    HIPCHK2(up(W.slotWB, slotWB.data(), (size_t)nperf * 4));
    {
        // every perforation starts open (Well::Setup, src/Well.cpp:64-73)
        std::vector<int32_t> st((size_t)nperf, 1);
        HIPCHK2(up(W.perfState, st.data(), (size_t)nperf * 4));
    }
    const int rc = uploadWellState(c);
    if (rc != OCPG_OK) return rc;
    // the last-step BHP starts as the given one (UpdateLastTimeStepBHP of
    // IsoT_FIM::InitReservoir, src/OCPFluidMethod.cpp:836-850)
    HIPCHK2(hipMemcpy(W.lbhp, W.bhp, NW * 8, hipMemcpyDeviceToDevice));
    return OCPG_OK;
}

namespace ocpg {
// The relative-residual maxima of the last CalRes (waits for the stream).
int fimReadNorms(ocpg_ctx* c, ocpg_resnorm* norms)
{
    FimDevice& F = ctxFim(c);
    WellDevice& W = ctxWells(c);
    {
        std::vector<double> v((size_t)F.normBlocks * 2);
        std::vector<int64_t> id((size_t)F.normBlocks * 2);
        HIPCHK2(hipMemcpyAsync(v.data(), F.normPart, v.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctxStream(c)));
        HIPCHK2(hipMemcpyAsync(id.data(), F.normIdx, id.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               ctxStream(c)));
        HIPCHK2(hipStreamSynchronize(ctxStream(c)));
        ocpg_resnorm r = {0, 0, 0, 0, 0};
        if (W.nw > 0) {
            std::vector<double> rel(W.nw);
            HIPCHK2(hipMemcpyAsync(rel.data(), W.wellRelRes, rel.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   ctxStream(c)));
            HIPCHK2(hipStreamSynchronize(ctxStream(c)));
            for (double v : rel) r.maxWellRelRes_mol = std::max(r.maxWellRelRes_mol, v);
        }
        for (int b = 0; b < F.normBlocks; b++) {
            if (r.maxRelRes_V < v[2 * b]) {
                r.maxRelRes_V = v[2 * b];
                r.maxId_V = id[2 * b];
            }
            if (r.maxRelRes_N < v[2 * b + 1]) {
                r.maxRelRes_N = v[2 * b + 1];
                r.maxId_N = id[2 * b + 1];
            }
        }
        *norms = r;
    }
    return OCPG_OK;
}

}  // namespace ocpg

extern "C" {

// BulkConn::Setup/CalAkd output (src/BulkConn.cpp:31-117) and the BSR
// pattern that LinearSystem::AllocateColMem + AssembleMatBulksNew produce.
// Connections must satisfy bId < eId and be sorted by (bId, eId), as
// BulkConn::iteratorConn is.  Removes any wells.
int ocpg_set_conn(ocpg_ctx* c, const ocpg_conn* conn)
{
    if (!c || !conn || (conn->nconn > 0 && (!conn->bId || !conn->eId || !conn->area))) return OCPG_EINVAL;
    DevBulk& B = ctxBulk(c);
    if (B.nb <= 0) return OCPG_ESTATE;
    const int64_t nb = B.nb, nconn = conn->nconn;
    for (int64_t k = 0; k < nconn; k++) {
        const int64_t b = conn->bId[k], e = conn->eId[k];
        if (b < 0 || e >= nb || b >= e) return OCPG_EINVAL;
        if (k > 0 && (conn->bId[k - 1] > b || (conn->bId[k - 1] == b && conn->eId[k - 1] >= e)))
            return OCPG_EINVAL;
    }
    if (nb + 2 * nconn > (int64_t)INT32_MAX) return OCPG_ETOOBIG;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    FimHost& H = ctxFimHost(c);
    H.bId.assign(conn->bId, conn->bId + nconn);
    H.eId.assign(conn->eId, conn->eId + nconn);
    H.wells = WellsHost{};
    H.pool = PoolHost{};
    {
        const int rc = ctxDropMask(c);
        if (rc != OCPG_OK) return rc;
    }
    FimDevice& F = ctxFim(c);
    ctxDropGraph(c);
    fimFree(F);
    F.nb = nb;
    F.nconn = nconn;
    F.ncol = ctxNC(c) + 2;
    F.normBlocks = (int)((nb + 255) / 256);
    size_t tot = 0;
    HIPCHK2(devAlloc(F.bId, nconn, tot));
    HIPCHK2(devAlloc(F.eId, nconn, tot));
    HIPCHK2(devAlloc(F.area, nconn, tot));
    HIPCHK2(devAlloc(F.upblock, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upRho, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upVel, nconn * 3, tot));
    HIPCHK2(devAlloc(F.rec, (size_t)nb * recordDoubles(ctxNC(c)), tot));
    HIPCHK2(devAlloc(F.own, (size_t)nb * ownDoubles(ctxNC(c)), tot));
    HIPCHK2(devAlloc(F.normPart, (size_t)F.normBlocks * 2, tot));
    HIPCHK2(devAlloc(F.normIdx, (size_t)F.normBlocks * 2, tot));
    F.bytes = tot;
    F.patternBytes = 0;
    HIPCHK2(hipMemcpy(F.bId, conn->bId, nconn * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.eId, conn->eId, nconn * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.area, conn->area, nconn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK2(hipMemset(F.upblock, 0, nconn * 3 * sizeof(int32_t)));
    HIPCHK2(hipMemset(F.upRho, 0, nconn * 3 * sizeof(double)));
    HIPCHK2(hipMemset(F.upVel, 0, nconn * 3 * sizeof(double)));
    return buildPattern(c);
}

// ---------------------------------------------------------------------------
// Serving pool: many realisations in one context, each at a reserved range of
// cells, connection slots and BSR blocks, and an activity mask that selects
// the realisations an iteration (the captured one included) works on.  The
// kernels are the ones of an ordinary context; a realisation's rows only
// reach its own cells, so realisations never see each other.

// Capacity cells with at most maxNb neighbours each; every cell starts
// outside any realisation (diagonal-only rows, mask off).
int ocpg_pool_init(ocpg_ctx* c, int64_t capacity, int32_t maxNb)
{
    if (!c || capacity <= 0 || maxNb < 1 || maxNb > 64) return OCPG_EINVAL;
    if (ctxFimn(c)) return OCPG_EUNSUP;  // the serving pool runs IsoT_FIM
    const int64_t bpc = 1 + (int64_t)maxNb, cpc = (maxNb + 1) / 2;
    if (capacity * bpc > (int64_t)INT32_MAX) return OCPG_ETOOBIG;
    int rc = ocpg_set_cells(c, capacity);
    if (rc == OCPG_OK) rc = ctxPoolTrimLast(c);
    if (rc != OCPG_OK) return rc;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    if (!ctxCopyStream(c)) return OCPG_EHIP;
    FimHost& H = ctxFimHost(c);
    H.bId.clear();
    H.eId.clear();
    H.wells = WellsHost{};
    H.pool = PoolHost{};
    H.pool.capacity = capacity;
    H.pool.maxNb = maxNb;
    H.pool.connPerCell = (int)cpc;
    FimDevice& F = ctxFim(c);
    WellDevice& W = ctxWells(c);
    ctxDropGraph(c);
    fimFree(F);
    if (W.mem) HIPCHK2(hipFree(W.mem));
    W = WellDevice{};
    const int NC = ctxNC(c);
    const int64_t nconn = capacity * cpc, nnzb = capacity * bpc;
    F.nb = capacity;
    F.nconn = nconn;
    F.ncol = NC + 2;
    F.normBlocks = (int)((capacity + 255) / 256);
    F.nnzb = nnzb;
    F.nrow = capacity;
    const int ncol = F.ncol;
    size_t tot = 0;
    HIPCHK2(devAlloc(F.bId, nconn, tot));
    HIPCHK2(devAlloc(F.eId, nconn, tot));
    HIPCHK2(devAlloc(F.area, nconn, tot));
    HIPCHK2(devAlloc(F.upblock, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upRho, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upVel, nconn * 3, tot));
    HIPCHK2(devAlloc(F.rec, (size_t)capacity * recordDoubles(NC), tot));
    HIPCHK2(devAlloc(F.own, (size_t)capacity * ownDoubles(NC), tot));
    HIPCHK2(devAlloc(F.normPart, (size_t)F.normBlocks * 2, tot));
    HIPCHK2(devAlloc(F.normIdx, (size_t)F.normBlocks * 2, tot));
    const size_t before = tot;
    HIPCHK2(devAlloc(F.IA, capacity + 1, tot));
    HIPCHK2(devAlloc(F.JA, nnzb, tot));
    HIPCHK2(devAlloc(F.slotConn, nnzb, tot));
    HIPCHK2(devAlloc(F.slotMirror, nnzb, tot));
    HIPCHK2(devAlloc(F.bulkEnd, capacity, tot));
    HIPCHK2(devAlloc(F.res, capacity * ncol, tot));
    HIPCHK2(devAlloc(F.val, (size_t)nnzb * ncol * ncol, tot));
    F.patternBytes = tot - before;
    F.bytes = tot;
    HIPCHK2(hipMemset(F.bId, 0, nconn * sizeof(int32_t)));
    HIPCHK2(hipMemset(F.eId, 0, nconn * sizeof(int32_t)));
    HIPCHK2(hipMemset(F.area, 0, nconn * sizeof(double)));
    HIPCHK2(hipMemset(F.upblock, 0, nconn * 3 * sizeof(int32_t)));
    HIPCHK2(hipMemset(F.upRho, 0, nconn * 3 * sizeof(double)));
    HIPCHK2(hipMemset(F.upVel, 0, nconn * 3 * sizeof(double)));
    HIPCHK2(hipMemset(F.res, 0, capacity * ncol * sizeof(double)));
    {
        std::vector<int32_t> IA(capacity + 1), JA(nnzb, 0), sc(nnzb, -1), sm(nnzb), be(capacity);
        for (int64_t n = 0; n <= capacity; n++) IA[n] = (int32_t)(n * bpc);
        for (int64_t n = 0; n < capacity; n++) {
            for (int64_t k = IA[n]; k < IA[n] + bpc; k++) {
                JA[k] = (int32_t)n;
                sm[k] = (int32_t)k;
            }
            be[n] = IA[n] + 1;
        }
        HIPCHK2(hipMemcpy(F.IA, IA.data(), IA.size() * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.JA, JA.data(), JA.size() * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.slotConn, sc.data(), sc.size() * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.slotMirror, sm.data(), sm.size() * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.bulkEnd, be.data(), be.size() * 4, hipMemcpyHostToDevice));
    }
    if (!ctxEnsureMask(c)) return OCPG_EHIP;
    // the saved iteration inputs, filled per realisation by ocpg_pool_place
    return ocpg_save_flash_inputs(c);
}

// Places a realisation: its cell state (h->nb cells, the layout of
// ocpg_upload_bulk) and its connections (local cell indices, the order and
// rules of ocpg_set_conn) at cells [base, base + h->nb).  The range must not
// overlap a placed realisation.  The state's iteration inputs are also saved
// (the starting point of every iteration over it); its mask stays off until
// ocpg_pool_activate.
int ocpg_pool_place(ocpg_ctx* c, int64_t base, const ocpg_bulk* h, const ocpg_conn* conn)
{
    if (!c || !h || !conn || (conn->nconn > 0 && (!conn->bId || !conn->eId || !conn->area))) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    if (P.capacity <= 0) return OCPG_ESTATE;
    const int64_t nb = h->nb, nconn = conn->nconn, D = P.maxNb, bpc = 1 + D, cpc = P.connPerCell;
    if (nb <= 0 || base < 0 || base + nb > P.capacity || nconn > nb * cpc) return OCPG_EINVAL;
    {
        auto it = P.slots.lower_bound(base);
        if (it != P.slots.end() && it->first < base + nb) return OCPG_EINVAL;
        if (it != P.slots.begin()) {
            --it;
            if (it->first + it->second.nb > base) return OCPG_EINVAL;
        }
    }
    std::vector<int64_t> cnt(nb, 1);
    for (int64_t k = 0; k < nconn; k++) {
        const int64_t b = conn->bId[k], e = conn->eId[k];
        if (b < 0 || e >= nb || b >= e) return OCPG_EINVAL;
        if (k > 0 && (conn->bId[k - 1] > b || (conn->bId[k - 1] == b && conn->eId[k - 1] >= e)))
            return OCPG_EINVAL;
        cnt[b]++;
        cnt[e]++;
    }
    for (int64_t n = 0; n < nb; n++)
        if (cnt[n] > bpc) return OCPG_EINVAL;
    // no wait for the context's stream: a free range has its mask off, so no
    // iteration in flight reads or writes it; copies of a former occupant
    // (and the waits ocpg_pool_remove queued for it) finish first
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxCopyStream(c)));
    // the realisation's rows as buildPattern lays them out, then moved to its
    // reserved ranges
    PoolSlot S;
    S.nb = nb;
    S.nconn = nconn;
    S.IA.assign(nb + 1, 0);
    for (int64_t n = 0; n < nb; n++) S.IA[n + 1] = S.IA[n] + (int32_t)cnt[n];
    S.nnzb = S.IA[nb];
    const int64_t cb = base * cpc, kb = base * bpc;
    std::vector<int32_t> IA(nb), JA(S.nnzb), sc(S.nnzb), sm(S.nnzb), be(nb), bId(nconn), eId(nconn);
    std::vector<int64_t> fill(nb);
    for (int64_t n = 0; n < nb; n++) {
        IA[n] = (int32_t)(kb + S.IA[n]);
        JA[S.IA[n]] = (int32_t)(base + n);
        sc[S.IA[n]] = -1;
        sm[S.IA[n]] = IA[n];
        be[n] = (int32_t)(kb + S.IA[n + 1]);
        fill[n] = S.IA[n] + 1;
    }
    for (int64_t k = 0; k < nconn; k++) {
        const int32_t b = conn->bId[k], e = conn->eId[k];
        const int64_t se = fill[e]++, sb = fill[b]++;
        JA[se] = (int32_t)(base + b);
        JA[sb] = (int32_t)(base + e);
        sc[se] = sc[sb] = (int32_t)(cb + k);
        sm[se] = (int32_t)(kb + sb);
        sm[sb] = (int32_t)(kb + se);
        bId[k] = (int32_t)(base + b);
        eId[k] = (int32_t)(base + e);
    }
    FimDevice& F = ctxFim(c);
    HIPCHK2(hipMemcpy(F.IA + base, IA.data(), nb * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.bulkEnd + base, be.data(), nb * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.JA + kb, JA.data(), S.nnzb * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotConn + kb, sc.data(), S.nnzb * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotMirror + kb, sm.data(), S.nnzb * 4, hipMemcpyHostToDevice));
    if (nconn > 0) {
        HIPCHK2(hipMemcpy(F.bId + cb, bId.data(), nconn * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.eId + cb, eId.data(), nconn * 4, hipMemcpyHostToDevice));
        HIPCHK2(hipMemcpy(F.area + cb, conn->area, nconn * 8, hipMemcpyHostToDevice));
    }
    // the saved inputs on the copy stream: before any export of this
    // realisation, and before the context's later work
    int rc = ctxRangeCopy(c, base, const_cast<ocpg_bulk*>(h), true, false);
    if (rc == OCPG_OK) rc = ctxSaveInputsRange(c, base, nb, ctxCopyStream(c));
    if (rc == OCPG_OK) rc = ctxMainAfterCopy(c);
    if (rc != OCPG_OK) return rc;
    P.slots[base] = std::move(S);
    return OCPG_OK;
}

// Forgets the realisation at base (its mask off); its ranges may be reused.
// Does not wait: what comes next in the range is ordered after the
// iterations that may still include this realisation.
int ocpg_pool_remove(ocpg_ctx* c, int64_t base)
{
    if (!c) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end()) return OCPG_EINVAL;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    PoolSlot& S = it->second;
    if (S.active) {
        uint8_t* m = ctxEnsureMask(c);
        if (!m) return OCPG_EHIP;
        HIPCHK2(launchMaskFill(m, base, S.nb, 0, ctxStream(c)));
        S.active = false;
        S.offAt = ctxLaunches(c);
    }
    // the range's next occupant is written once every iteration that may
    // include this one is done: the copy stream waits for them
    if (S.offAt > ctxLaunchesDone(c)) {
        const int rc = ctxCopyAfterMain(c);
        if (rc != OCPG_OK) return rc;
    }
    P.slots.erase(it);
    return OCPG_OK;
}

// Sets whether the realisation at base takes part in the iterations launched
// after this call on the context's stream (a kernel, in stream order).
int ocpg_pool_activate(ocpg_ctx* c, int64_t base, int on)
{
    if (!c) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end()) return OCPG_EINVAL;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    uint8_t* m = ctxEnsureMask(c);
    if (!m) return OCPG_EHIP;
    HIPCHK2(launchMaskFill(m, base, it->second.nb, on ? 1 : 0, ctxStream(c)));
    PoolSlot& S = it->second;
    if (S.active && !on) S.offAt = ctxLaunches(c);
    S.active = on != 0;
    return OCPG_OK;
}

// The realisation's rows of the last assembly in its own indexing (the
// system ocpg_get_bsr returns for the realisation alone): IA (nb + 1), JA
// and val (nnzb blocks), rhs (nb x ncol).  Waits for the stream.  NULL
// outputs are skipped; *nnzb (if given) receives the block count.
int ocpg_pool_system(ocpg_ctx* c, int64_t base, int64_t* nnzb, int32_t* IA, int32_t* JA, double* val, double* rhs)
{
    if (!c) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end()) return OCPG_EINVAL;
    const PoolSlot& S = it->second;
    if (nnzb) *nnzb = S.nnzb;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    const FimDevice& F = ctxFim(c);
    const int64_t kb = base * (1 + (int64_t)P.maxNb);
    const size_t bsize = (size_t)F.ncol * F.ncol;
    if (IA) std::memcpy(IA, S.IA.data(), S.IA.size() * 4);
    if (JA) {
        HIPCHK2(hipMemcpy(JA, F.JA + kb, S.nnzb * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < S.nnzb; k++) JA[k] -= (int32_t)base;
    }
    if (val) HIPCHK2(hipMemcpy(val, F.val + kb * bsize, S.nnzb * bsize * 8, hipMemcpyDeviceToHost));
    if (rhs) HIPCHK2(hipMemcpy(rhs, F.res + base * F.ncol, S.nb * F.ncol * 8, hipMemcpyDeviceToHost));
    return OCPG_OK;
}

// The realisation's cell state into h (h->nb must be its cell count):
// saved != 0 gives the state its next iteration starts from (the saved
// iteration inputs, with the other fields as they are); otherwise the
// current state.  Waits for the stream.
int ocpg_pool_download(ocpg_ctx* c, int64_t base, ocpg_bulk* h, int saved)
{
    if (!c || !h) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end() || it->second.nb != h->nb) return OCPG_EINVAL;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    HIPCHK2(hipStreamSynchronize(ctxCopyStream(c)));
    return ctxRangeCopy(c, base, h, false, saved != 0);
}

// Realisation images: a placed realisation's device state and pattern as
// one host block (a header, then each piece in the device's own layout, 256
// byte aligned), so that readmitting it is a few straight copies (no
// transposition, no pattern build) on the pool's copy stream, which
// overlap the iterations running on the context's stream.
namespace {
constexpr uint64_t kImageMagic = 0x474D493147504F43ull;  // "OCPG1IMG"
struct ImageHeader {
    uint64_t magic;
    int32_t NC, maxNb;
    int64_t nb, nconn, nnzb, base0;
};
static_assert(sizeof(ImageHeader) <= 256, "image header");

// Every piece of the realisation at base, in image order (the state part
// first: stateSegs of them).
void imagePieces(ocpg_ctx* c, int64_t base, int64_t nb, int64_t nconn, int64_t nnzb, std::vector<ImageSeg>& seg,
                 size_t& stateSegs)
{
    const PoolHost& P = ctxFimHost(c).pool;
    FimDevice& F = ctxFim(c);
    const int64_t kb = base * (1 + (int64_t)P.maxNb), cb = base * P.connPerCell;
    seg.clear();
    ctxImageSegments(c, base, nb, true, false, seg);
    stateSegs = seg.size();
    ctxImageSegments(c, base, nb, false, true, seg);
    auto add = [&](void* p, size_t bytes) { seg.push_back({static_cast<char*>(p), bytes, bytes, 1}); };
    add(F.IA + base, (size_t)nb * 4);
    add(F.bulkEnd + base, (size_t)nb * 4);
    add(F.JA + kb, (size_t)nnzb * 4);
    add(F.slotConn + kb, (size_t)nnzb * 4);
    add(F.slotMirror + kb, (size_t)nnzb * 4);
    add(F.bId + cb, (size_t)nconn * 4);
    add(F.eId + cb, (size_t)nconn * 4);
    add(F.area + cb, (size_t)nconn * 8);
}

size_t imageOffsets(const std::vector<ImageSeg>& seg, std::vector<size_t>& off)
{
    size_t o = 256;
    off.resize(seg.size());
    for (size_t i = 0; i < seg.size(); i++) {
        off[i] = o;
        o += align256w(seg[i].width * seg[i].rows);
    }
    return o;
}
}  // namespace

// Bytes of the image of the realisation at base.
int64_t ocpg_pool_image_bytes(ocpg_ctx* c, int64_t base)
{
    if (!c) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end()) return OCPG_EINVAL;
    std::vector<ImageSeg> seg;
    std::vector<size_t> off;
    size_t ns;
    imagePieces(c, base, it->second.nb, it->second.nconn, it->second.nnzb, seg, ns);
    return (int64_t)imageOffsets(seg, off);
}

// Copies the realisation at base (switched off) into image: everything with
// all != 0, otherwise only its state (saved iteration inputs and last-step
// moles) into an image it filled before.  Asynchronous on the copy stream
// (ocpg_pool_sync_copies waits; page-locked images copy without staging).
int ocpg_pool_export(ocpg_ctx* c, int64_t base, void* image, int all)
{
    if (!c || !image) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    auto it = P.slots.find(base);
    if (P.capacity <= 0 || it == P.slots.end() || it->second.active) return OCPG_EINVAL;
    const PoolSlot& S = it->second;
    ImageHeader* hd = static_cast<ImageHeader*>(image);
    if (!all && (hd->magic != kImageMagic || hd->nb != S.nb || hd->nconn != S.nconn || hd->nnzb != S.nnzb))
        return OCPG_EINVAL;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    hipStream_t cs = ctxCopyStream(c);
    if (!cs) return OCPG_EHIP;
    if (S.offAt > ctxLaunchesDone(c)) {
        const int rc = ctxCopyAfterMain(c);
        if (rc != OCPG_OK) return rc;
    }
    std::vector<ImageSeg> seg;
    std::vector<size_t> off;
    size_t ns;
    imagePieces(c, base, S.nb, S.nconn, S.nnzb, seg, ns);
    imageOffsets(seg, off);
    char* img = static_cast<char*>(image);
    const size_t n = all ? seg.size() : ns;
    for (size_t i = 0; i < n; i++)
        HIPCHK2(hipMemcpy2DAsync(img + off[i], seg[i].width, seg[i].dev, seg[i].pitch, seg[i].width, seg[i].rows,
                                 hipMemcpyDeviceToHost, cs));
    if (all) {
        ImageHeader h{kImageMagic, ctxNC(c), P.maxNb, S.nb, S.nconn, S.nnzb, base};
        std::memcpy(image, &h, sizeof(h));
    }
    return OCPG_OK;
}

// Places a realisation from its image at cells [base, base + nb) (a free
// range): the copies and the pattern's move to the new base run on the copy
// stream, and the context's later work waits for them.  The image must stay
// as it is until ocpg_pool_sync_copies.
int ocpg_pool_import(ocpg_ctx* c, int64_t base, const void* image)
{
    if (!c || !image) return OCPG_EINVAL;
    PoolHost& P = ctxFimHost(c).pool;
    if (P.capacity <= 0) return OCPG_ESTATE;
    ImageHeader h;
    std::memcpy(&h, image, sizeof(h));
    if (h.magic != kImageMagic || h.NC != ctxNC(c) || h.maxNb != P.maxNb || h.nb <= 0 || h.nconn < 0 ||
        h.nconn > h.nb * P.connPerCell || h.nnzb < h.nb || h.nnzb > h.nb * (1 + (int64_t)P.maxNb))
        return OCPG_EINVAL;
    const int64_t nb = h.nb;
    if (base < 0 || base + nb > P.capacity) return OCPG_EINVAL;
    {
        auto it = P.slots.lower_bound(base);
        if (it != P.slots.end() && it->first < base + nb) return OCPG_EINVAL;
        if (it != P.slots.begin()) {
            --it;
            if (it->first + it->second.nb > base) return OCPG_EINVAL;
        }
    }
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    hipStream_t cs = ctxCopyStream(c);
    if (!cs) return OCPG_EHIP;
    std::vector<ImageSeg> seg;
    std::vector<size_t> off;
    size_t ns;
    imagePieces(c, base, nb, h.nconn, h.nnzb, seg, ns);
    imageOffsets(seg, off);
    const char* img = static_cast<const char*>(image);
    // the realisation's own row starts, from the image's (at its old base)
    const int64_t bpc = 1 + (int64_t)P.maxNb;
    PoolSlot S;
    S.nb = nb;
    S.nconn = h.nconn;
    S.nnzb = h.nnzb;
    S.IA.resize(nb + 1);
    const int32_t* IAimg = reinterpret_cast<const int32_t*>(img + off[seg.size() - 8]);
    for (int64_t n = 0; n < nb; n++) S.IA[n] = (int32_t)(IAimg[n] - h.base0 * bpc);
    S.IA[nb] = (int32_t)h.nnzb;
    if (S.IA[0] != 0) return OCPG_EINVAL;
    for (int64_t n = 0; n < nb; n++)
        if (S.IA[n + 1] <= S.IA[n] || S.IA[n + 1] - S.IA[n] > bpc) return OCPG_EINVAL;
    for (size_t i = 0; i < seg.size(); i++)
        HIPCHK2(hipMemcpy2DAsync(seg[i].dev, seg[i].pitch, img + off[i], seg[i].width, seg[i].width, seg[i].rows,
                                 hipMemcpyHostToDevice, cs));
    const int64_t d = base - h.base0;
    if (d != 0) {
        FimDevice& F = ctxFim(c);
        const int64_t kb = base * bpc, cb = base * P.connPerCell;
        HIPCHK2(launchPoolRebase(F.IA + base, F.bulkEnd + base, nb, F.JA + kb, F.slotConn + kb, F.slotMirror + kb,
                                 h.nnzb, F.bId + cb, F.eId + cb, h.nconn, d, d * P.connPerCell, d * bpc, cs));
    }
    const int rc = ctxMainAfterCopy(c);
    if (rc != OCPG_OK) return rc;
    P.slots[base] = std::move(S);
    return OCPG_OK;
}

// Waits for the pool's copies (exports and imports).
int ocpg_pool_sync_copies(ocpg_ctx* c)
{
    if (!c) return OCPG_EINVAL;
    if (ctxFimHost(c).pool.capacity <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxCopyStream(c)));
    return OCPG_OK;
}

int64_t ocpg_bsr_nnzb(ocpg_ctx* c) { return c ? ctxFim(c).nnzb : -1; }
int64_t ocpg_bsr_rows(ocpg_ctx* c) { return c ? ctxFim(c).nrow : -1; }

// AllWells (include/AllWells.hpp) as IsoT_FIM reads it; rebuilds the pattern.
int ocpg_set_wells(ocpg_ctx* c, const ocpg_wells* w)
{
    if (!c || !w || w->nw < 0) return OCPG_EINVAL;
    if (ctxFimHost(c).pool.capacity > 0) return OCPG_ESTATE;  // a pool's realisations carry no wells
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0 || !F.bId) return OCPG_ESTATE;
    if (!flashSupportsNC(ctxNC(c))) return OCPG_EUNSUP;
    const int nw = w->nw;
    const int nc = ctxNC(c) + 1;
    WellsHost WH;
    WH.nw = nw;
    if (nw > 0) {
        if (!w->type || !w->open || !w->mode || !w->bhp || !w->ctrlBHP || !w->ctrlRate || !w->perfStart ||
            !w->perfCell || !w->perfWI || !w->perfDG)
            return OCPG_EINVAL;
        if (w->perfStart[0] != 0) return OCPG_EINVAL;
        for (int i = 0; i < nw; i++) {
            if (w->perfStart[i + 1] < w->perfStart[i]) return OCPG_EINVAL;
            if (w->type[i] != OCPG_WELL_INJ && w->type[i] != OCPG_WELL_PROD) return OCPG_EINVAL;
            if (w->mode[i] != OCPG_WELL_BHP && w->mode[i] != OCPG_WELL_RATE) return OCPG_EINVAL;
            if (w->type[i] == OCPG_WELL_INJ && (!w->injZi || !w->perfXi)) return OCPG_EINVAL;
            if (w->type[i] == OCPG_WELL_PROD && w->mode[i] == OCPG_WELL_RATE && !positiveWeights(w->prodPhaseWeight, i))
                return OCPG_EINVAL;
        }
        const int nperf = w->perfStart[nw];
        for (int p = 0; p < nperf; p++)
            if (w->perfCell[p] < 0 || w->perfCell[p] >= F.nb) return OCPG_EINVAL;
        WH.type.assign(w->type, w->type + nw);
        WH.open.assign(w->open, w->open + nw);
        WH.mode.assign(w->mode, w->mode + nw);
        WH.unweight.assign(nw, 0);
        if (w->unweight) WH.unweight.assign(w->unweight, w->unweight + nw);
        WH.perfStart.assign(w->perfStart, w->perfStart + nw + 1);
        WH.perfCell.assign(w->perfCell, w->perfCell + nperf);
        WH.perfWI.assign(w->perfWI, w->perfWI + nperf);
        WH.perfMult.assign(nperf, 1.0);
        if (w->perfMult) WH.perfMult.assign(w->perfMult, w->perfMult + nperf);
        WH.injZi.assign((size_t)nw * nc, 0.0);
        WH.prodPW.assign((size_t)nw * 3, 0.0);
        WH.hasPW.assign(nw, 0);
        for (int i = 0; i < nw; i++) {
            if (w->type[i] == OCPG_WELL_PROD) WH.hasPW[i] = positiveWeights(w->prodPhaseWeight, i) ? 1 : 0;
            if (w->type[i] == OCPG_WELL_INJ)
                std::copy(w->injZi + (size_t)i * nc, w->injZi + (size_t)(i + 1) * nc, WH.injZi.begin() + (size_t)i * nc);
            else if (w->prodPhaseWeight)
                std::copy(w->prodPhaseWeight + (size_t)i * 3, w->prodPhaseWeight + (size_t)(i + 1) * 3,
                          WH.prodPW.begin() + (size_t)i * 3);
        }
        WH.Psurf = w->Psurf;
        WH.Tsurf = w->Tsurf;
        setWellState(WH, *w);
    }
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    ctxFimHost(c).wells = WH;
    return buildPattern(c);
}

// New values of the per-step / per-iteration well fields.
int ocpg_set_well_state(ocpg_ctx* c, const ocpg_wells* w)
{
    if (!c || !w) return OCPG_EINVAL;
    WellsHost& WH = ctxFimHost(c).wells;
    if (w->nw != WH.nw) return OCPG_EINVAL;
    if (WH.nw == 0) return OCPG_OK;
    if (!w->mode || !w->bhp || !w->ctrlBHP || !w->ctrlRate || !w->perfDG) return OCPG_EINVAL;
    for (int i = 0; i < WH.nw; i++) {
        if (w->mode[i] != OCPG_WELL_BHP && w->mode[i] != OCPG_WELL_RATE) return OCPG_EINVAL;
        if (WH.type[i] == OCPG_WELL_INJ && !w->perfXi) return OCPG_EINVAL;
        if (WH.type[i] == OCPG_WELL_PROD && w->mode[i] == OCPG_WELL_RATE) {
            // new weights may come with the switch to a rate control
            const bool given = w->prodPhaseWeight && positiveWeights(w->prodPhaseWeight, i);
            if (!given && !WH.hasPW[i]) return OCPG_EINVAL;
        }
    }
    if (w->prodPhaseWeight) {
        for (int i = 0; i < WH.nw; i++) {
            if (WH.type[i] != OCPG_WELL_PROD || !positiveWeights(w->prodPhaseWeight, i)) continue;
            std::copy(w->prodPhaseWeight + (size_t)i * 3, w->prodPhaseWeight + (size_t)(i + 1) * 3,
                      WH.prodPW.begin() + (size_t)i * 3);
            WH.hasPW[i] = 1;
        }
    }
    setWellState(WH, *w);
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    const int rc = uploadWellState(c);
    if (rc != OCPG_OK) return rc;
    WellDevice& W = ctxWells(c);
    HIPCHK2(hipMemcpy(W.prodPW, WH.prodPW.data(), (size_t)WH.nw * 3 * 8, hipMemcpyHostToDevice));
    return OCPG_OK;
}

// Well::depth, WellOpt::injTemp / injProdPhase / initOptMode for
// ocpg_prepare_wells and the CaldG of ocpg_reset_to_last_step.
int ocpg_set_well_setup(ocpg_ctx* c, const ocpg_well_setup* su)
{
    if (!c || !su) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    const WellsHost& WH = ctxFimHost(c).wells;
    if (W.nw <= 0) return OCPG_ESTATE;
    if (!su->depth || !su->initMode) return OCPG_EINVAL;
    const int nw = W.nw;
    std::vector<double> temp(nw, 0.0);
    std::vector<int32_t> phase(nw, 1);
    for (int w = 0; w < nw; w++) {
        if (su->initMode[w] != OCPG_WELL_BHP && su->initMode[w] != OCPG_WELL_RATE) return OCPG_EINVAL;
        if (!std::isfinite(su->depth[w])) return OCPG_EINVAL;
        if (WH.type[w] == OCPG_WELL_INJ) {
            if (!su->injTemp || !su->injPhase) return OCPG_EINVAL;
            if (su->injPhase[w] != 1 && su->injPhase[w] != 2) return OCPG_EINVAL;
            temp[w] = su->injTemp[w];
            phase[w] = su->injPhase[w];
        }
        if (WH.type[w] == OCPG_WELL_PROD && su->initMode[w] == OCPG_WELL_RATE && !WH.hasPW[w])
            return OCPG_EINVAL;
    }
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    HIPCHK2(hipMemcpy(W.wellDepth, su->depth, nw * 8, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.initMode, su->initMode, nw * 4, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.injTemp, temp.data(), nw * 8, hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(W.injPhase, phase.data(), nw * 4, hipMemcpyHostToDevice));
    W.hasSetup = true;
    ctxDropGraph(c);
    return OCPG_OK;
}

// AllWells::PrepareWell on the device (asynchronous).
int ocpg_prepare_wells(ocpg_ctx* c)
{
    if (!c) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    if (W.nw <= 0 || !W.hasSetup) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(launchWellPrepare(ctxNC(c), ctxEos(c), ctxBulk(c), W, 0, ctxStream(c)));
    return OCPG_OK;
}

// AllWells::CheckP on the device; waits for the result.
int ocpg_check_wells(ocpg_ctx* c, int32_t* status)
{
    if (!c || !status) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    if (W.nw <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(launchWellCheck(ctxNC(c), ctxBulk(c), W, ctxStream(c)));
    HIPCHK2(hipMemcpyAsync(status, W.status, sizeof(int32_t), hipMemcpyDeviceToHost, ctxStream(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    return OCPG_OK;
}

int ocpg_get_well_state(ocpg_ctx* c, int32_t* mode, double* bhp, double* perfDG, double* perfXi, double* perfMult)
{
    if (!c) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    if (W.nw <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    const size_t NW = W.nw, NP = W.nperf;
    if (mode) HIPCHK2(hipMemcpy(mode, W.mode, NW * 4, hipMemcpyDeviceToHost));
    if (bhp) HIPCHK2(hipMemcpy(bhp, W.bhp, NW * 8, hipMemcpyDeviceToHost));
    if (NP && perfDG) HIPCHK2(hipMemcpy(perfDG, W.perfDG, NP * 8, hipMemcpyDeviceToHost));
    if (NP && perfXi) HIPCHK2(hipMemcpy(perfXi, W.perfXi, NP * 8, hipMemcpyDeviceToHost));
    if (NP && perfMult) HIPCHK2(hipMemcpy(perfMult, W.perfMult, NP * 8, hipMemcpyDeviceToHost));
    return OCPG_OK;
}

// MixtureComp::XiPhase / RhoPhase of one feed.
int ocpg_phase_density(ocpg_ctx* c, double P, double T, const double* zi, int phase, double* xi, double* rho)
{
    if (!c || !zi || !(P > 0) || !std::isfinite(T)) return OCPG_EINVAL;
    if (!flashSupportsNC(ctxNC(c))) return OCPG_EUNSUP;
    const int nc = ctxNC(c) + 1;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    double* d = nullptr;
    HIPCHK2(hipMalloc(&d, (size_t)(nc + 2) * sizeof(double)));
    double out[2] = {0, 0};
    hipError_t e = hipMemcpy(d, zi, nc * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launchPhaseDensity(ctxNC(c), ctxEos(c), P, T, d, phase == 2 ? 2 : 1, d + nc, ctxStream(c));
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + nc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctxStream(c));
    if (e == hipSuccess) e = hipStreamSynchronize(ctxStream(c));
    (void)hipFree(d);
    HIPCHK2(e);
    if (xi) *xi = out[0];
    if (rho) *rho = out[1];
    return OCPG_OK;
}

// Bulk::CheckNi, then Bulk::CheckP; waits for the result.
int ocpg_check_bulk(ocpg_ctx* c, int32_t* status)
{
    if (!c || !status) return OCPG_EINVAL;
    const DevBulk& B = ctxBulk(c);
    if (B.nb <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    int rc = ensureNewton(c);
    if (rc != OCPG_OK) return rc;
    NewtonDevice& N = ctxNewton(c);
    HIPCHK2(hipMemsetAsync(N.chk, 0, sizeof(int32_t), ctxStream(c)));
    HIPCHK2(launchBulkCheck(B, ctxNC(c) + 1, N.chk, ctxStream(c)));
    int32_t f = 0;
    HIPCHK2(hipMemcpyAsync(&f, N.chk, sizeof(int32_t), hipMemcpyDeviceToHost, ctxStream(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    *status = (f & 2) ? OCPG_BULK_NEGATIVE_COMPONENTS_MOLES : (f & 1) ? OCPG_BULK_NEGATIVE_PRESSURE : 0;
    return OCPG_OK;
}

// Bulk::CalMaxChange + AllWells::CalMaxBHPChange; waits for the result.
int ocpg_max_change(ocpg_ctx* c, ocpg_max_change_t* out)
{
    if (!c || !out) return OCPG_EINVAL;
    const DevBulk& B = ctxBulk(c);
    if (B.nb <= 0 || !ctxLastValid(c)) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    int rc = ensureNewton(c);
    if (rc != OCPG_OK) return rc;
    NewtonDevice& N = ctxNewton(c);
    HIPCHK2(launchMaxChange(B, ctxLastBulk(c), ctxNC(c) + 1, N.chg, N.blocks, ctxStream(c)));
    std::vector<double> part((size_t)N.blocks * 4);
    HIPCHK2(hipMemcpyAsync(part.data(), N.chg, part.size() * sizeof(double), hipMemcpyDeviceToHost, ctxStream(c)));
    WellDevice& W = ctxWells(c);
    std::vector<double> bhp(W.nw), lbhp(W.nw);
    std::vector<int32_t> open(W.nw);
    if (W.nw > 0) {
        HIPCHK2(hipMemcpyAsync(bhp.data(), W.bhp, W.nw * 8, hipMemcpyDeviceToHost, ctxStream(c)));
        HIPCHK2(hipMemcpyAsync(lbhp.data(), W.lbhp, W.nw * 8, hipMemcpyDeviceToHost, ctxStream(c)));
        HIPCHK2(hipMemcpyAsync(open.data(), W.open, W.nw * 4, hipMemcpyDeviceToHost, ctxStream(c)));
    }
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    ocpg_max_change_t r = {0, 0, 0, 0, 0};
    for (int b = 0; b < N.blocks; b++) {
        r.dPmax = std::max(r.dPmax, part[4 * b]);
        r.dNmax = std::max(r.dNmax, part[4 * b + 1]);
        r.dSmax = std::max(r.dSmax, part[4 * b + 2]);
        r.eVmax = std::max(r.eVmax, part[4 * b + 3]);
    }
    for (int w = 0; w < W.nw; w++)
        if (open[w]) r.dBHPmax = std::max(r.dBHPmax, std::fabs(bhp[w] - lbhp[w]));
    *out = r;
    return OCPG_OK;
}

// Well::CalFlux / CalProdWeight results of the last ocpg_calres.
int ocpg_get_well_flux(ocpg_ctx* c, double* perfQi, double* wellQi, double* prodWeight)
{
    if (!c) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    if (W.nw == 0) return OCPG_OK;
    const size_t nc = ctxNC(c) + 1;
    if (perfQi && W.nperf) HIPCHK2(hipMemcpy(perfQi, W.perfQi, W.nperf * nc * 8, hipMemcpyDeviceToHost));
    if (wellQi) HIPCHK2(hipMemcpy(wellQi, W.wellQi, W.nw * nc * 8, hipMemcpyDeviceToHost));
    if (prodWeight) HIPCHK2(hipMemcpy(prodWeight, W.prodWeight, W.nw * 8, hipMemcpyDeviceToHost));
    return OCPG_OK;
}

// IsoT_FIM::CalKrPc (src/OCPFluidMethod.cpp:1189-1199).
int ocpg_calkrpc(ocpg_ctx* c)
{
    if (!c || ctxBulk(c).nb <= 0) return OCPG_ESTATE;
    if (!ctxSat(c)) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(launchKrPcRock(ctxBulk(c), ctxSat(c), ctxRock(c), true, false, ctxStream(c)));
    return OCPG_OK;
}

// IsothermalMethod::CalRock (src/OCPFluidMethod.cpp:25-32).
int ocpg_calrock(ocpg_ctx* c)
{
    if (!c || ctxBulk(c).nb <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(launchKrPcRock(ctxBulk(c), ctxSat(c), ctxRock(c), false, true, ctxStream(c)));
    return OCPG_OK;
}

// CalKrPc and CalRock in one pass over the cells.
int ocpg_calkrpc_rock(ocpg_ctx* c)
{
    if (!c || ctxBulk(c).nb <= 0 || !ctxSat(c)) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    ctxTimerBegin(c, T_KRPC);
    HIPCHK2(launchKrPcRock(ctxBulk(c), ctxSat(c), ctxRock(c), true, true, ctxStream(c)));
    ctxTimerEnd(c, T_KRPC);
    return OCPG_OK;
}

// IsoT_FIM::CalRes, bulk part (src/OCPFluidMethod.cpp:1201-1282, 1362-1387).
// The negated residual stays on the device as the right-hand side; with
// norms non-NULL the call waits and returns the maxima.
int ocpg_calres(ocpg_ctx* c, double dt, ocpg_resnorm* norms)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0 || F.nb != ctxBulk(c).nb) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    WellDevice& W = ctxWells(c);
    ctxTimerBegin(c, T_RES);
    if (W.nw > 0) HIPCHK2(launchWellFlux(ctxNC(c), ctxEos(c), ctxBulk(c), F, W, ctxStream(c)));
    HIPCHK2(launchCalRes(ctxNC(c), ctxBulk(c), F, W, dt, ctxStream(c)));
    ctxTimerEnd(c, T_RES);
    if (norms) return fimReadNorms(c, norms);
    return OCPG_OK;
}

// The device arrays of GetSolution for the current cells (allocated on
// first use, kept while the cell count stays).
static int ensureNewton(ocpg_ctx* c)
{
    NewtonDevice& N = ctxNewton(c);
    const DevBulk& B = ctxBulk(c);
    if (N.mem && N.nb == B.nb && N.ld == B.ld) return OCPG_OK;
    if (N.mem) HIPCHK2(hipFree(N.mem));
    N.mem = nullptr;
    N.nb = B.nb;
    N.ld = B.ld;
    N.blocks = (int)std::max<int64_t>((B.nb + 255) / 256, 1);
    N.valid = false;
    const size_t L = (size_t)std::max<int64_t>(B.ld, 64);
    const size_t bD = align256w(3 * L * sizeof(double)), bB = align256w(3 * L);
    const size_t bP = align256w((size_t)N.blocks * 4 * sizeof(double)), bI = align256w((size_t)N.blocks * 2 * 8);
    const size_t bK = 256;
    HIPCHK2(hipMalloc(&N.mem, 2 * bD + bB + 2 * bP + bI + bK));
    char* p = static_cast<char*>(N.mem);
    N.dSNR = reinterpret_cast<double*>(p);
    N.dSNRP = reinterpret_cast<double*>(p + bD);
    N.pexOld = reinterpret_cast<uint8_t*>(p + 2 * bD);
    N.part = reinterpret_cast<double*>(p + 2 * bD + bB);
    N.partIdx = reinterpret_cast<int64_t*>(p + 2 * bD + bB + bP);
    N.chg = reinterpret_cast<double*>(p + 2 * bD + bB + bP + bI);
    N.chk = reinterpret_cast<int32_t*>(p + 2 * bD + bB + 2 * bP + bI);
    return OCPG_OK;
}

// Host reduction of the blocks' first maxima (slot 0 or 1): the value at
// the largest magnitude, lowest index on ties; (0, 0) when none is nonzero.
static int firstMaxOverBlocks(ocpg_ctx* c, int slot, double* val, int64_t* idx)
{
    NewtonDevice& N = ctxNewton(c);
    std::vector<double> part((size_t)N.blocks * 4);
    std::vector<int64_t> pidx((size_t)N.blocks * 2);
    HIPCHK2(hipMemcpyAsync(part.data(), N.part, part.size() * sizeof(double), hipMemcpyDeviceToHost, ctxStream(c)));
    HIPCHK2(hipMemcpyAsync(pidx.data(), N.partIdx, pidx.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                           ctxStream(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    double bv = 0, bk = 0;
    int64_t bi = 0;
    bool any = false;
    for (int b = 0; b < N.blocks; b++) {
        const double v = part[4 * b + 2 * slot], k = part[4 * b + 2 * slot + 1];
        const int64_t i = pidx[2 * b + slot];
        if (k > bk || (any && k == bk && k > 0 && i < bi)) {
            bv = v;
            bk = k;
            bi = i;
            any = true;
        }
    }
    *val = bv;
    if (idx) *idx = any ? bi : 0;
    return OCPG_OK;
}

// IsoT_FIM::GetSolution (src/OCPFluidMethod.cpp:2735-2857) on the resident
// state: u is the linear system's solution, rows x ncol in the LinearSystem
// layout (bulk rows, then open wells), in host memory or (u_on_device)
// device memory.  Updates P, Ni, S, x_ij and the open wells' BHP; keeps the
// saturations before the update and the predicted changes for
// ocpg_nr_change.  out may be NULL (then the call is asynchronous).  In an
// IsoT_FIMn context IsoT_FIMn::GetSolution (:3847-3990): P, Ni, nj, x_ij
// and the BHP (the saturations wait for the next flash).
int ocpg_get_solution(ocpg_ctx* c, const double* u, int u_on_device, double dSmaxlim, ocpg_nr_update* out)
{
    if (!c || !u || !(dSmaxlim > 0)) return OCPG_EINVAL;
    const DevBulk& B = ctxBulk(c);
    FimDevice& F = ctxFim(c);
    if (B.nb <= 0 || F.nb != B.nb) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    int rc = ensureNewton(c);
    if (rc != OCPG_OK) return rc;
    NewtonDevice& N = ctxNewton(c);
    WellDevice& W = ctxWells(c);
    const int64_t ncol = ctxNC(c) + 2;
    const int64_t len = (B.nb + W.nwOpen) * ncol;
    const double* du = u;
    if (!u_on_device) {
        if (N.uLen < len) {
            if (N.u) HIPCHK2(hipFree(N.u));
            N.u = nullptr;
            N.uLen = 0;
            HIPCHK2(hipMalloc(&N.u, (size_t)len * sizeof(double)));
            N.uLen = len;
        }
        HIPCHK2(hipMemcpyAsync(N.u, u, (size_t)len * sizeof(double), hipMemcpyHostToDevice, ctxStream(c)));
        du = N.u;
    }
    if (ctxFimn(c))
        HIPCHK2(launchGetSolutionN(ctxNC(c), B, N, W, B.nb, du, dSmaxlim, ctxScm(c), ctxStream(c)));
    else
        HIPCHK2(launchGetSolution(ctxNC(c), B, N, W, B.nb, du, dSmaxlim, ctxStream(c)));
    N.valid = true;
    if (out) {
        if ((rc = firstMaxOverBlocks(c, 0, &out->NRdPmax, nullptr)) != OCPG_OK) return rc;
        if ((rc = firstMaxOverBlocks(c, 1, &out->NRdNmax, nullptr)) != OCPG_OK) return rc;
    }
    return OCPG_OK;
}

// After the flash that followed ocpg_get_solution: PassFlashValue's dSNR
// and maxNRdSSP (src/OCPFluidMethod.cpp:1135-1146) and Bulk::CalNRdSmax
// (src/Bulk.cpp:1235-1247).  Waits for the stream.
int ocpg_nr_change(ocpg_ctx* c, ocpg_nr_change_t* out)
{
    if (!c || !out) return OCPG_EINVAL;
    NewtonDevice& N = ctxNewton(c);
    const DevBulk& B = ctxBulk(c);
    if (!N.valid || N.nb != B.nb) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(launchNRChange(B, N, ctxStream(c)));
    N.valid = false;
    int rc;
    int64_t k = 0;
    if ((rc = firstMaxOverBlocks(c, 0, &out->NRdSmax, &k)) != OCPG_OK) return rc;
    out->NRdSmax_cell = k / 3;
    if ((rc = firstMaxOverBlocks(c, 1, &out->maxNRdSSP, &out->maxNRdSSP_cell)) != OCPG_OK) return rc;
    return OCPG_OK;
}

// The open and shut wells' bottom-hole pressures held on the device (nw).
int ocpg_get_well_bhp(ocpg_ctx* c, double* bhp)
{
    if (!c || !bhp) return OCPG_EINVAL;
    WellDevice& W = ctxWells(c);
    if (W.nw <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipMemcpyAsync(bhp, W.bhp, (size_t)W.nw * sizeof(double), hipMemcpyDeviceToHost, ctxStream(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    return OCPG_OK;
}

// IsoT_FIM::AssembleMat (src/OCPFluidMethod.cpp:860-875), bulk blocks; the
// right-hand side is the residual CalRes left on the device.  In an
// IsoT_FIMn context IsoT_FIMn::AssembleMat, whose right-hand side adds the
// flash correction's terms to the residual.
int ocpg_assemble(ocpg_ctx* c, double dt)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0 || F.nb != ctxBulk(c).nb) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    WellDevice& W = ctxWells(c);
    ctxTimerBegin(c, T_PACK);
    HIPCHK2(launchPackRecords(ctxNC(c), ctxBulk(c), F, ctxStream(c)));
    ctxTimerEnd(c, T_PACK);
    ctxTimerBegin(c, T_ASM);
    if (F.fimn) {
        // IsoT_FIMn::AssembleMat (src/OCPFluidMethod.cpp:2998-3006): bulks,
        // wells, then AssembleRhsAccumulate(resAbs)
        HIPCHK2(launchAssembleN(ctxNC(c), ctxBulk(c), F, dt, ctxStream(c)));
        if (W.nw > 0) HIPCHK2(launchWellAssemble(ctxNC(c), ctxBulk(c), F, W, dt, ctxStream(c)));
        HIPCHK2(launchRhsAccumulate(F, ctxStream(c)));
    } else {
        HIPCHK2(launchAssemble(ctxNC(c), ctxBulk(c), F, dt, ctxStream(c)));
        if (W.nw > 0) HIPCHK2(launchWellAssemble(ctxNC(c), ctxBulk(c), F, W, dt, ctxStream(c)));
    }
    ctxTimerEnd(c, T_ASM);
    return OCPG_OK;
}

// LinearSystem -> VectorFaspSolver::AssembleMat hand-off (src/FaspSolver.cpp:
// 362-429): IA (rows+1), JA (nnzb), val (nnzb*ncol*ncol, row-major blocks),
// rhs (rows*ncol).  Any pointer may be NULL.
int ocpg_get_bsr(ocpg_ctx* c, int32_t* IA, int32_t* JA, double* val, double* rhs)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    const size_t bs = (size_t)F.ncol * F.ncol;
    if (IA) HIPCHK2(hipMemcpy(IA, F.IA, (F.nrow + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (JA) HIPCHK2(hipMemcpy(JA, F.JA, F.nnzb * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (val) HIPCHK2(hipMemcpy(val, F.val, F.nnzb * bs * sizeof(double), hipMemcpyDeviceToHost));
    if (rhs)
        HIPCHK2(hipMemcpy(rhs, F.fimn ? F.rhsN : F.res, F.nrow * F.ncol * sizeof(double), hipMemcpyDeviceToHost));
    return OCPG_OK;
}

// The same hand-off without waiting: the copies are queued on the
// context's stream behind the assembly and overlap with host work; the
// caller waits with ocpg_synchronize before FASP reads the arrays.  With
// page-locked destinations (ocpg_host_alloc) they run at the link's rate.
int ocpg_get_bsr_async(ocpg_ctx* c, int32_t* IA, int32_t* JA, double* val, double* rhs)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    const size_t bs = (size_t)F.ncol * F.ncol;
    hipStream_t s = ctxStream(c);
    if (IA) HIPCHK2(hipMemcpyAsync(IA, F.IA, (F.nrow + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (JA) HIPCHK2(hipMemcpyAsync(JA, F.JA, F.nnzb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (val) HIPCHK2(hipMemcpyAsync(val, F.val, F.nnzb * bs * sizeof(double), hipMemcpyDeviceToHost, s));
    if (rhs)
        HIPCHK2(hipMemcpyAsync(rhs, F.fimn ? F.rhsN : F.res, F.nrow * F.ncol * sizeof(double), hipMemcpyDeviceToHost,
                               s));
    return OCPG_OK;
}

// Page-locked host memory for the hand-off arrays (NULL on failure).
void* ocpg_host_alloc(int64_t bytes)
{
    void* p = nullptr;
    if (bytes <= 0 || hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

int ocpg_host_free(void* p)
{
    if (!p) return OCPG_OK;
    return hipHostFree(p) == hipSuccess ? OCPG_OK : OCPG_EHIP;
}

// BulkConn upwinding written by CalRes: upblock, upblock_Rho,
// upblock_Velocity (nconn*np each).  Any pointer may be NULL.
int ocpg_get_upwind(ocpg_ctx* c, int32_t* upblock, double* upRho, double* upVel)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    HIPCHK2(hipStreamSynchronize(ctxStream(c)));
    const size_t n = (size_t)F.nconn * 3;
    if (upblock) HIPCHK2(hipMemcpy(upblock, F.upblock, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (upRho) HIPCHK2(hipMemcpy(upRho, F.upRho, n * sizeof(double), hipMemcpyDeviceToHost));
    if (upVel) HIPCHK2(hipMemcpy(upVel, F.upVel, n * sizeof(double), hipMemcpyDeviceToHost));
    return OCPG_OK;
}

}  // extern "C"
