// C-ABI of the FIM half: connection upload, CalKrPc/CalRock, CalRes and
// AssembleMat, and the hand-off of the assembled system in the reference's
// LinearSystem / FASP dBSRmat layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ocpgpu.h"
#include "device_state.hpp"
#include "fim_kernels.hpp"

namespace ocpg {
DevBulk& ctxBulk(ocpg_ctx* c);
FimDevice& ctxFim(ocpg_ctx* c);
hipStream_t ctxStream(ocpg_ctx* c);
int ctxNC(ocpg_ctx* c);
int ctxDevice(ocpg_ctx* c);
const DevSat* ctxSat(ocpg_ctx* c);
const DevRock& ctxRock(ocpg_ctx* c);
void ctxTimerBegin(ocpg_ctx* c, int slot);
void ctxTimerEnd(ocpg_ctx* c, int slot);
}  // namespace ocpg

// timer slots of ocpgpu_capi.cpp
enum { T_FLASH = 0, T_KRPC, T_RES, T_ASM };

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

extern "C" {

// BulkConn::Setup/CalAkd output (src/BulkConn.cpp:31-117) and the BSR
// pattern that LinearSystem::AllocateColMem + AssembleMatBulksNew produce:
// row n = [n, then its neighbours in the order connections reach them].
// Connections must satisfy bId < eId and be sorted by (bId, eId), as
// BulkConn::iteratorConn is.
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
    // pattern on the host
    std::vector<int32_t> cnt(nb, 1);
    for (int64_t k = 0; k < nconn; k++) {
        cnt[conn->bId[k]]++;
        cnt[conn->eId[k]]++;
    }
    std::vector<int32_t> IA(nb + 1, 0);
    for (int64_t n = 0; n < nb; n++) IA[n + 1] = IA[n] + cnt[n];
    const int64_t nnzb = IA[nb];
    std::vector<int32_t> JA(nnzb), slotConn(nnzb), slotMirror(nnzb), fill(nb);
    for (int64_t n = 0; n < nb; n++) {
        JA[IA[n]] = (int32_t)n;
        slotConn[IA[n]] = -1;
        slotMirror[IA[n]] = IA[n];
        fill[n] = IA[n] + 1;
    }
    for (int64_t k = 0; k < nconn; k++) {
        const int32_t b = conn->bId[k], e = conn->eId[k];
        const int32_t se = fill[e]++;  // NewOffDiag(eId, bId)
        const int32_t sb = fill[b]++;  // NewOffDiag(bId, eId)
        JA[se] = b;
        JA[sb] = e;
        slotConn[se] = slotConn[sb] = (int32_t)k;
        slotMirror[se] = sb;
        slotMirror[sb] = se;
    }
    FimDevice& F = ctxFim(c);
    fimFree(F);
    const int ncol = ctxNC(c) + 2;
    F.nb = nb;
    F.nconn = nconn;
    F.nnzb = nnzb;
    F.ncol = ncol;
    F.normBlocks = (int)((nb + 255) / 256);
    size_t tot = 0;
    HIPCHK2(devAlloc(F.bId, nconn, tot));
    HIPCHK2(devAlloc(F.eId, nconn, tot));
    HIPCHK2(devAlloc(F.area, nconn, tot));
    HIPCHK2(devAlloc(F.IA, nb + 1, tot));
    HIPCHK2(devAlloc(F.JA, nnzb, tot));
    HIPCHK2(devAlloc(F.slotConn, nnzb, tot));
    HIPCHK2(devAlloc(F.slotMirror, nnzb, tot));
    HIPCHK2(devAlloc(F.upblock, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upRho, nconn * 3, tot));
    HIPCHK2(devAlloc(F.upVel, nconn * 3, tot));
    HIPCHK2(devAlloc(F.res, nb * ncol, tot));
    HIPCHK2(devAlloc(F.val, (size_t)nnzb * ncol * ncol, tot));
    HIPCHK2(devAlloc(F.normPart, (size_t)F.normBlocks * 2, tot));
    HIPCHK2(devAlloc(F.normIdx, (size_t)F.normBlocks * 2, tot));
    F.bytes = tot;
    HIPCHK2(hipMemcpy(F.bId, conn->bId, nconn * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.eId, conn->eId, nconn * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.area, conn->area, nconn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.IA, IA.data(), IA.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.JA, JA.data(), JA.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotConn, slotConn.data(), nnzb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemcpy(F.slotMirror, slotMirror.data(), nnzb * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK2(hipMemset(F.upblock, 0, nconn * 3 * sizeof(int32_t)));
    HIPCHK2(hipMemset(F.upRho, 0, nconn * 3 * sizeof(double)));
    HIPCHK2(hipMemset(F.upVel, 0, nconn * 3 * sizeof(double)));
    return OCPG_OK;
}

int64_t ocpg_bsr_nnzb(ocpg_ctx* c) { return c ? ctxFim(c).nnzb : -1; }

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
    ctxTimerBegin(c, T_RES);
    HIPCHK2(launchCalRes(ctxNC(c), ctxBulk(c), F, dt, ctxStream(c)));
    ctxTimerEnd(c, T_RES);
    if (norms) {
        std::vector<double> v((size_t)F.normBlocks * 2);
        std::vector<int64_t> id((size_t)F.normBlocks * 2);
        HIPCHK2(hipMemcpyAsync(v.data(), F.normPart, v.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctxStream(c)));
        HIPCHK2(hipMemcpyAsync(id.data(), F.normIdx, id.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               ctxStream(c)));
        HIPCHK2(hipStreamSynchronize(ctxStream(c)));
        ocpg_resnorm r = {0, 0, 0, 0};
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

// IsoT_FIM::AssembleMat (src/OCPFluidMethod.cpp:860-875), bulk blocks; the
// right-hand side is the residual CalRes left on the device.
int ocpg_assemble(ocpg_ctx* c, double dt)
{
    if (!c) return OCPG_EINVAL;
    FimDevice& F = ctxFim(c);
    if (F.nb <= 0 || F.nb != ctxBulk(c).nb) return OCPG_ESTATE;
    HIPCHK2(hipSetDevice(ctxDevice(c)));
    ctxTimerBegin(c, T_ASM);
    HIPCHK2(launchAssemble(ctxNC(c), ctxBulk(c), F, dt, ctxStream(c)));
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
    if (IA) HIPCHK2(hipMemcpy(IA, F.IA, (F.nb + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (JA) HIPCHK2(hipMemcpy(JA, F.JA, F.nnzb * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (val) HIPCHK2(hipMemcpy(val, F.val, F.nnzb * bs * sizeof(double), hipMemcpyDeviceToHost));
    if (rhs) HIPCHK2(hipMemcpy(rhs, F.res, F.nb * F.ncol * sizeof(double), hipMemcpyDeviceToHost));
    return OCPG_OK;
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
