// Host side of the C-ABI declared in include/ocpgpu.h: context, device
// memory layout, transfers in the reference's Bulk layout, and the calls
// that stand in for IsoT_FIM's property pass and assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "../../include/ocpgpu.h"
#include "device_state.hpp"
#include "fim_host.hpp"
#include "fim_kernels.hpp"
#include "flash_kernel.hpp"
#include "mlp_kernels.hpp"

using namespace ocpg;

namespace {

// One SoA field: where its host array lives in ocpg_bulk, its element size
// and how many entries a cell holds.
// C1n, CNPn, CRESn: IsoT_FIMn fields, no entries in an IsoT_FIM context.
enum Cnt { C1, CNC, CNCh, CNP, CNPNC, CNPNP, CLEN, C5, C1n, CNPn, CRESn };

struct FieldDesc {
    size_t hostOff;   // offset of the pointer inside ocpg_bulk
    size_t devOff;    // offset of the pointer inside DevBulk
    int elem;         // bytes per entry
    Cnt cnt;
    bool lastStep;    // part of the last-time-step copy
};

#define HB(f) offsetof(ocpg_bulk, f)
#define DB(f) offsetof(DevBulk, f)

// The fields IsoT_FIM::UpdateLastTimeStep copies (src/OCPFluidMethod.cpp:
// 2915-2957) come first so that they form one contiguous block; SkipStaAnaly
// keeps its own last-step copy (src/AcceleratePVT.cpp:122-138), included.
const FieldDesc kFields[] = {
    {HB(poro), DB(poro), 8, C1, true},          {HB(poroP), DB(poroP), 8, C1, true},
    {HB(rockVp), DB(rockVp), 8, C1, true},      {HB(phaseNum), DB(phaseNum), 4, C1, true},
    {HB(Nt), DB(Nt), 8, C1, true},              {HB(Ni), DB(Ni), 8, CNC, true},
    {HB(vf), DB(vf), 8, C1, true},              {HB(P), DB(P), 8, C1, true},
    {HB(Pj), DB(Pj), 8, CNP, true},             {HB(Pc), DB(Pc), 8, CNP, true},
    {HB(phaseExist), DB(phaseExist), 1, CNP, true}, {HB(S), DB(S), 8, CNP, true},
    {HB(xij), DB(xij), 8, CNPNC, true},         {HB(rho), DB(rho), 8, CNP, true},
    {HB(xi), DB(xi), 8, CNP, true},             {HB(mu), DB(mu), 8, CNP, true},
    {HB(kr), DB(kr), 8, CNP, true},             {HB(vfP), DB(vfP), 8, C1, true},
    {HB(vfi), DB(vfi), 8, CNC, true},           {HB(rhoP), DB(rhoP), 8, CNP, true},
    {HB(rhox), DB(rhox), 8, CNPNC, true},       {HB(xiP), DB(xiP), 8, CNP, true},
    {HB(xix), DB(xix), 8, CNPNC, true},         {HB(muP), DB(muP), 8, CNP, true},
    {HB(mux), DB(mux), 8, CNPNC, true},         {HB(dPcj_dS), DB(dPcj_dS), 8, CNPNP, true},
    {HB(dKr_dS), DB(dKr_dS), 8, CNPNP, true},   {HB(bRowSizedSdP), DB(bRowSizedSdP), 4, C1, true},
    {HB(dSec_dPri), DB(dSec), 8, CLEN, true},   {HB(pSderExist), DB(pSderExist), 1, CNP, true},
    {HB(pVnumCom), DB(pVnumCom), 4, CNP, true}, {HB(skipFlag), DB(skipFlag), 1, C1, true},
    {HB(skipMinEig), DB(skipMinEig), 8, C1, true}, {HB(skipP), DB(skipP), 8, C1, true},
    {HB(skipT), DB(skipT), 8, C1, true},        {HB(skipZi), DB(skipZi), 8, CNCh, true},
    // IsoT_FIMn's additions (src/OCPFluidMethod.cpp:3999-4015: nj, res_n and
    // resPc are part of its last-step copy)
    {HB(nj), DB(nj), 8, CNPn, true},            {HB(res_n), DB(resN), 8, CRESn, true},
    {HB(resPc), DB(resPc), 8, C1n, true},
    // not copied at the end of a time step
    {HB(T), DB(T), 8, C1, false},               {HB(v), DB(v), 8, C1, false},
    {HB(depth), DB(depth), 8, C1, false},       {HB(poroInit), DB(poroInit), 8, C1, false},
    {HB(ftype), DB(ftype), 4, C1, false},       {HB(iters), DB(iters), 4, C5, false},
    {HB(nonconv), DB(nonconv), 4, C1, false},
};
constexpr int kNumFields = sizeof(kFields) / sizeof(kFields[0]);

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Timer slots of ocpg_timing_get.
const char* kTimerNames[kTimerSlots] = {"flash",        "krpc_rock",       "calres",      "assemble",
                                        "init_flash",   "restore",         "flash_stability", "flash_split",
                                        "flash_finish", "flash_tail",      "fixup",       "update_assemble",
                                        "flash_restarts", "pack",          "mlp"};

// Fields the flash reads (and overwrites): what one Newton iteration starts from.
bool isFlashInput(size_t hostOff)
{
    return hostOff == HB(P) || hostOff == HB(Ni) || hostOff == HB(S) || hostOff == HB(phaseNum) ||
           hostOff == HB(xij) || hostOff == HB(skipFlag) || hostOff == HB(skipMinEig) ||
           hostOff == HB(skipP) || hostOff == HB(skipT) || hostOff == HB(skipZi);
}

}  // namespace

struct ocpg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int NC = 0;
    ocpg_eos eosHost{};
    DevEos* dEos = nullptr;
    DevSat satHost{};
    DevSat* dSat = nullptr;
    DevRock rock{};
    DevFlashCounters* dCtr = nullptr;
    // cells, and the last-step copies of the fields IsoT_FIM keeps (lP, lS, ...)
    DevBulk B{};
    DevBulk Blast{};
    char* cur = nullptr;
    char* last = nullptr;
    size_t lastBytes = 0;  // bytes of the last-step block at the front of `cur`
    size_t curBytes = 0;
    size_t fieldOff[kNumFields] = {};
    bool lastValid = false;
    // staged-flash scratch
    FlashWork work{};
    char* workMem = nullptr;
    size_t workBytes = 0;
    // connections and BSR (fim_kernels), wells
    FimDevice fim{};
    WellDevice wells{};
    FimHost fimHost;
    // copy of the flash inputs (ocpg_save_flash_inputs)
    char* snap = nullptr;
    size_t snapBytes = 0;
    // serving pool: per-cell activity mask (B.mask), NULL outside a pool;
    // a pool keeps no last-step copy but the moles (poolLNi), and moves
    // realisation images on a stream of its own
    uint8_t* mask = nullptr;
    bool pool = false;
    double* poolLNi = nullptr;
    hipStream_t copyStream = nullptr;
    hipEvent_t copyEv = nullptr, mainEv = nullptr;
    // replays of the captured iteration launched / known to have finished
    uint64_t launches = 0, launchesDone = 0;
    // learned form (ocpg_set_learned): the network's weights on the device
    // and its per-cell output
    float* mlpMem = nullptr;
    const float* mlpW[8] = {};
    float* guessBuf = nullptr;
    int64_t guessRows = 0;
    // overlapped Newton iteration (ocpg_update_assemble): side streams of the
    // flash and the rows recomputed after they join
    FlashSide side{};
    FixupRows fix{};
    // IsoT_FIM::GetSolution on the device (ocpg_get_solution)
    NewtonDevice newton{};
    // one Newton iteration captured as a graph (ocpg_capture_iteration)
    hipGraphExec_t iterGraph = nullptr;
    hipEvent_t iterDone = nullptr;
    // IsoT_FIMn in place of IsoT_FIM (ocpg_set_method, before ocpg_set_cells)
    // and the critical saturations its GetSolution reads (Bulk::satcm)
    bool fimn = false;
    double scm[3] = {0, 0, 0};
    // per-kernel timers (ocpg_timing_*)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events[kTimerSlots];
    hipEvent_t open[kTimerSlots] = {};
};

static int hipStatus(hipError_t e)
{
    if (e == hipSuccess) return OCPG_OK;
    std::fprintf(stderr, "ocpgpu: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
    return OCPG_EHIP;
}
#define HIPCHK(x)                          \
    do {                                   \
        hipError_t e_ = (x);               \
        if (e_ != hipSuccess) return hipStatus(e_); \
    } while (0)

static int64_t cellEntries(const ocpg_ctx* c, Cnt k)
{
    const int nc = c->NC + 1;
    switch (k) {
        case C1: return 1;
        case CNC: return nc;
        case CNCh: return c->NC;
        case CNP: return OCPG_NP;
        case CNPNC: return OCPG_NP * nc;
        case CNPNP: return OCPG_NP * OCPG_NP;
        case CLEN: return (int64_t)(nc + 1) * (nc + 1) * OCPG_NP;
        case C5: return 5;
        case C1n: return c->fimn ? 1 : 0;
        case CNPn: return c->fimn ? OCPG_NP : 0;
        case CRESn: return c->fimn ? (int64_t)OCPG_NP * (nc + 1) : 0;
    }
    return 0;
}

static void buildDevEos(const ocpg_eos& e, DevEos& d)
{
    std::memset(&d, 0, sizeof(d));
    const int NC = e.nc;
    d.nc = NC;
    d.pvtwRows = e.pvtw_rows;
    for (int i = 0; i < NC; i++) {
        d.Tc[i] = e.Tc[i];
        d.Pc[i] = e.Pc[i];
        d.Vc[i] = e.Vc[i];
        d.MWC[i] = e.MW[i];
        d.Acf[i] = e.Acf[i];
        d.OmegaA[i] = e.OmegaA[i];
        d.OmegaB[i] = e.OmegaB[i];
        d.Vshift[i] = e.Vshift[i];
        d.Vcvis[i] = e.Vcvis[i];
        for (int k = 0; k < NC; k++) d.oneMinusBIC[i * NC + k] = 1 - e.BIC[i * OCPG_MAXNC + k];
        // MixtureComp::AllocateOthers, src/MixtureComp.cpp:2323-2332
        d.sqrtMWi[i] = std::sqrt(e.MW[i]);
        d.muAux1I[i] = 5.4402 * std::pow(e.Tc[i], 1.0 / 6) / std::pow(e.Pc[i], 2.0 / 3);
    }
    for (int q = 0; q < 5; q++) d.LBCcoef[q] = e.LBCcoef[q];
    for (int r = 0; r < e.pvtw_rows; r++)
        for (int q = 0; q < 5; q++) d.pvtw[r][q] = e.pvtw[r][q];
    d.stdRhoW = e.std_rhow;
    d.ssmStaMaxIt = e.ssmsta_maxit;
    d.ssmStaTol2 = e.ssmsta_tol * e.ssmsta_tol;
    d.ssmStaEYt = e.ssmsta_eyt;
    d.ssmStaKtol = 1E-4;   // SSMparamSTA::Ktol default, include/MixtureComp.hpp:33
    d.ssmStaDYtol = 1E-6;  // SSMparamSTA::dYtol default, include/MixtureComp.hpp:34
    d.nrStaMaxIt = e.nrsta_maxit;
    d.nrStaTol = e.nrsta_tol;
    d.ssmSpMaxIt = e.ssmsp_maxit;
    d.ssmSpTol2 = e.ssmsp_tol * e.ssmsp_tol;
    d.nrSpMaxIt = e.nrsp_maxit;
    d.nrSpTol = e.nrsp_tol;
    d.rrMaxIt = e.rr_maxit;
    d.rrTol = e.rr_tol;
}

static void timerBegin(ocpg_ctx* c, int slot)
{
    if (!c->timing) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->open[slot] = e;
}

static void timerEnd(ocpg_ctx* c, int slot)
{
    if (!c->timing || !c->open[slot]) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->events[slot].push_back({c->open[slot], e});
    c->open[slot] = nullptr;
}

static void timerClear(ocpg_ctx* c)
{
    for (int s = 0; s < kTimerSlots; s++) {
        for (auto& p : c->events[s]) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        c->events[s].clear();
        if (c->open[s]) (void)hipEventDestroy(c->open[s]);
        c->open[s] = nullptr;
    }
}

namespace ocpg {
int fimReadNorms(ocpg_ctx* c, ocpg_resnorm* norms);
}

static void destroySide(ocpg_ctx* c)
{
    hipStream_t ss[] = {c->side.sB, c->side.sC, c->side.sD};
    for (hipStream_t st : ss)
        if (st) (void)hipStreamSynchronize(st);
    hipEvent_t evs[] = {c->side.classified, c->side.sliced, c->side.newton, c->side.doneB, c->side.doneC,
                        c->side.doneD};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t st : ss)
        if (st) (void)hipStreamDestroy(st);
    c->side = FlashSide{};
}

// The recomputation lists of the overlapped iteration, for the current cells.
// The captured Newton iteration holds the device addresses it was captured
// with; every call that frees or moves a buffer it reads drops it, so a later
// ocpg_launch_iteration returns OCPG_ESTATE instead of replaying stale
// addresses.
static void dropIterGraph(ocpg_ctx* c)
{
    if (c->iterGraph) (void)hipGraphExecDestroy(c->iterGraph);
    c->iterGraph = nullptr;
}

static int ensureFixup(ocpg_ctx* c)
{
    const int64_t nb = c->B.nb;
    if (c->fix.mem && c->fix.nb == nb) return OCPG_OK;
    if (c->fix.mem) HIPCHK(hipFree(c->fix.mem));
    c->fix = FixupRows{};
    const size_t L = (size_t)std::max<int64_t>(nb, 1);
    const size_t bytes = align256(L * 4) * 3 + 256;
    HIPCHK(hipMalloc(&c->fix.mem, bytes));
    char* p = static_cast<char*>(c->fix.mem);
    c->fix.mark = reinterpret_cast<uint32_t*>(p);
    c->fix.resRows = reinterpret_cast<int32_t*>(p + align256(L * 4));
    c->fix.asmRows = reinterpret_cast<int32_t*>(p + 2 * align256(L * 4));
    c->fix.counts = reinterpret_cast<int32_t*>(p + 3 * align256(L * 4));
    c->fix.nb = nb;
    // the marks start (and stay, between calls) all zero
    HIPCHK(hipMemsetAsync(c->fix.mem, 0, bytes, c->stream));
    return OCPG_OK;
}

static bool tablesValid(int rows, int cols, const double* tab)
{
    if (rows < 1 || rows > OCPG_MAXTAB) return false;
    for (int r = 1; r < rows; r++)
        if (!(tab[r * cols] >= tab[(r - 1) * cols])) return false;
    return true;
}

extern "C" {

const char* ocpg_version(void) { return "ocpgpu 0.1 (gfx950)"; }

// Size of the structures of include/ocpgpu.h, for binding checks.
int64_t ocpg_sizeof(const char* name)
{
    if (!name) return -1;
    if (!std::strcmp(name, "ocpg_eos")) return sizeof(ocpg_eos);
    if (!std::strcmp(name, "ocpg_eos_deck")) return sizeof(ocpg_eos_deck);
    if (!std::strcmp(name, "ocpg_satfunc")) return sizeof(ocpg_satfunc);
    if (!std::strcmp(name, "ocpg_rock")) return sizeof(ocpg_rock);
    if (!std::strcmp(name, "ocpg_bulk")) return sizeof(ocpg_bulk);
    if (!std::strcmp(name, "ocpg_conn")) return sizeof(ocpg_conn);
    if (!std::strcmp(name, "ocpg_resnorm")) return sizeof(ocpg_resnorm);
    if (!std::strcmp(name, "ocpg_wells")) return sizeof(ocpg_wells);
    if (!std::strcmp(name, "ocpg_flash_stats")) return sizeof(ocpg_flash_stats);
    if (!std::strcmp(name, "ocpg_nr_update")) return sizeof(ocpg_nr_update);
    if (!std::strcmp(name, "ocpg_nr_change_t")) return sizeof(ocpg_nr_change_t);
    if (!std::strcmp(name, "ocpg_well_setup")) return sizeof(ocpg_well_setup);
    if (!std::strcmp(name, "ocpg_max_change_t")) return sizeof(ocpg_max_change_t);
    return -1;
}

// MixtureComp::MixtureComp(const ComponentParam&, ...), src/MixtureComp.cpp:
// 28-165: derive Vc (from ZCRIT), default OmegaA/OmegaB, scale SSHIFT, Vcvis,
// LBCcoef x 10, symmetrise BIC, copy controls.
int ocpg_eos_from_deck(const ocpg_eos_deck* d, ocpg_eos* e)
{
    if (!d || !e || d->nc < 1 || d->nc > OCPG_MAXNC) return OCPG_EINVAL;
    if (!d->has_Vc && !d->has_Zc) return OCPG_EINVAL;
    std::memset(e, 0, sizeof(*e));
    const int NC = d->nc;
    const double R = 10.73159;  // GAS_CONSTANT
    e->nc = NC;
    for (int i = 0; i < NC; i++) {
        e->Tc[i] = d->Tc[i];
        e->Pc[i] = d->Pc[i];
        e->MW[i] = d->MW[i];
        e->Acf[i] = d->Acf[i];
        e->Vc[i] = d->has_Vc ? d->Vc[i] : 10.73159 * d->Zc[i] * d->Tc[i] / d->Pc[i];
        e->OmegaA[i] = d->has_OmegaA ? d->OmegaA[i] : 0.457235529;
        e->OmegaB[i] = d->has_OmegaB ? d->OmegaB[i] : 0.077796074;
    }
    for (int i = 0; i < NC; i++) {
        e->Vshift[i] = d->has_Vshift ? d->Vshift[i] * (R * e->OmegaB[i] * e->Tc[i] / e->Pc[i]) : 0.0;
        if (d->has_Vcvis)
            e->Vcvis[i] = d->Vcvis[i];
        else if (d->has_Zcvis)
            e->Vcvis[i] = R * d->Zcvis[i] * e->Tc[i] / e->Pc[i];
        else
            e->Vcvis[i] = e->Vc[i];
    }
    for (int q = 0; q < 5; q++) e->LBCcoef[q] = d->LBCcoef[q] * 10;
    int it = 0;
    for (int i = 1; i < NC; i++)
        for (int j = 0; j < i; j++) {
            e->BIC[i * OCPG_MAXNC + j] = d->BIC_lower[it];
            e->BIC[j * OCPG_MAXNC + i] = d->BIC_lower[it];
            it++;
        }
    e->ssmsta_maxit = d->ssmsta_maxit;
    e->ssmsta_tol = d->ssmsta_tol;
    e->ssmsta_eyt = d->ssmsta_eyt;
    e->nrsta_maxit = d->nrsta_maxit;
    e->nrsta_tol = d->nrsta_tol;
    e->ssmsp_maxit = d->ssmsp_maxit;
    e->ssmsp_tol = d->ssmsp_tol;
    e->nrsp_maxit = d->nrsp_maxit;
    e->nrsp_tol = d->nrsp_tol;
    e->rr_maxit = d->rr_maxit;
    e->rr_tol = d->rr_tol;
    e->pvtw_rows = d->pvtw_rows;
    std::memcpy(e->pvtw, d->pvtw, sizeof(e->pvtw));
    e->std_rhow = d->std_rhow;
    return OCPG_OK;
}

int ocpg_create(ocpg_ctx** out, const ocpg_eos* eos, const ocpg_satfunc* sat, const ocpg_rock* rock,
                int device)
{
    if (!out || !eos) return OCPG_EINVAL;
    *out = nullptr;
    if (eos->nc < 1 || eos->nc > OCPG_MAXNC) return OCPG_EINVAL;
    if (!flashSupportsNC(eos->nc)) return OCPG_EUNSUP;
    if (!tablesValid(eos->pvtw_rows, 5, &eos->pvtw[0][0])) return OCPG_EINVAL;
    if (sat && (!tablesValid(sat->swof_rows, 4, &sat->swof[0][0]) ||
                !tablesValid(sat->sgof_rows, 4, &sat->sgof[0][0])))
        return OCPG_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return OCPG_ENODEV;
    if (device < 0 || device >= ndev) return OCPG_EINVAL;
    ocpg_ctx* c = new (std::nothrow) ocpg_ctx;
    if (!c) return OCPG_EINVAL;
    c->device = device;
    c->NC = eos->nc;
    c->eosHost = *eos;
    int rc = OCPG_OK;
    do {
        if ((rc = hipStatus(hipSetDevice(device)))) break;
        if ((rc = hipStatus(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)))) break;
        if ((rc = hipStatus(hipStreamCreateWithFlags(&c->side.sB, hipStreamNonBlocking)))) break;
        if ((rc = hipStatus(hipStreamCreateWithFlags(&c->side.sC, hipStreamNonBlocking)))) break;
        if ((rc = hipStatus(hipStreamCreateWithFlags(&c->side.sD, hipStreamNonBlocking)))) break;
        {
            hipEvent_t* evs[] = {&c->side.classified, &c->side.sliced, &c->side.newton, &c->side.doneB,
                                 &c->side.doneC, &c->side.doneD};
            for (hipEvent_t* e : evs)
                if ((rc = hipStatus(hipEventCreateWithFlags(e, hipEventDisableTiming)))) break;
            if (rc) break;
        }
        DevEos de;
        buildDevEos(*eos, de);
        if ((rc = hipStatus(hipMalloc(&c->dEos, sizeof(DevEos))))) break;
        if ((rc = hipStatus(hipMemcpy(c->dEos, &de, sizeof(DevEos), hipMemcpyHostToDevice)))) break;
        if ((rc = hipStatus(hipMalloc(&c->dCtr, sizeof(DevFlashCounters))))) break;
        if (sat) {
            std::memset(&c->satHost, 0, sizeof(DevSat));
            c->satHost.swofRows = sat->swof_rows;
            c->satHost.sgofRows = sat->sgof_rows;
            std::memcpy(c->satHost.swof, sat->swof, sizeof(sat->swof));
            std::memcpy(c->satHost.sgof, sat->sgof, sizeof(sat->sgof));
            c->satHost.kroMax = sat->swof[0][2];  // FlowUnit_ODGW01: SWOF col 2, row 0
            // Bulk::satcm of FlowUnit_ODGW01 (src/FlowUnit.cpp:196-202,
            // OCPTable::GetRowZero src/OCPTable.cpp:35-44): oil 0; gas / water
            // the saturation of the last SGOF / SWOF row before kr reaches 1e-8
            auto rowZero = [](const double (*tab)[4], int rows) {
                for (int r = 0; r < rows; r++)
                    if (tab[r][1] >= 1E-8) return r - 1;
                return rows - 1;
            };
            const int rg = rowZero(sat->sgof, sat->sgof_rows), rw = rowZero(sat->swof, sat->swof_rows);
            c->scm[0] = 0;
            c->scm[1] = rg >= 0 ? sat->sgof[rg][0] : 0.0;
            c->scm[2] = rw >= 0 ? sat->swof[rw][0] : 0.0;
            if ((rc = hipStatus(hipMalloc(&c->dSat, sizeof(DevSat))))) break;
            if ((rc = hipStatus(hipMemcpy(c->dSat, &c->satHost, sizeof(DevSat), hipMemcpyHostToDevice)))) break;
        }
        if (rock) {
            c->rock.Pref = rock->Pref;
            c->rock.cp1 = rock->cp1;
            c->rock.cp2 = rock->cp2;
        }
    } while (0);
    if (rc != OCPG_OK) {
        if (c->dEos) (void)hipFree(c->dEos);
        if (c->dCtr) (void)hipFree(c->dCtr);
        if (c->dSat) (void)hipFree(c->dSat);
        destroySide(c);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return OCPG_OK;
}

int ocpg_destroy(ocpg_ctx* c)
{
    if (!c) return OCPG_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    timerClear(c);
    if (c->iterGraph) (void)hipGraphExecDestroy(c->iterGraph);
    if (c->iterDone) (void)hipEventDestroy(c->iterDone);
    fimFree(c->fim);
    if (c->wells.mem) (void)hipFree(c->wells.mem);
    if (c->snap) (void)hipFree(c->snap);
    if (c->mask) (void)hipFree(c->mask);
    if (c->poolLNi) (void)hipFree(c->poolLNi);
    if (c->mlpMem) (void)hipFree(c->mlpMem);
    if (c->guessBuf) (void)hipFree(c->guessBuf);
    if (c->copyStream) {
        (void)hipStreamSynchronize(c->copyStream);
        (void)hipStreamDestroy(c->copyStream);
    }
    if (c->copyEv) (void)hipEventDestroy(c->copyEv);
    if (c->mainEv) (void)hipEventDestroy(c->mainEv);
    if (c->workMem) (void)hipFree(c->workMem);
    if (c->cur) (void)hipFree(c->cur);
    if (c->last) (void)hipFree(c->last);
    if (c->dEos) (void)hipFree(c->dEos);
    if (c->dCtr) (void)hipFree(c->dCtr);
    if (c->dSat) (void)hipFree(c->dSat);
    if (c->fix.mem) (void)hipFree(c->fix.mem);
    if (c->newton.mem) (void)hipFree(c->newton.mem);
    if (c->newton.u) (void)hipFree(c->newton.u);
    destroySide(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OCPG_OK;
}

// IsothermalSolver::SetupMethod's choice (src/IsothermalSolver.cpp:15-34):
// IsoT_FIM (OCPG_METHOD_FIM, the default) or IsoT_FIMn (OCPG_METHOD_FIMN,
// src/OCPFluidMethod.cpp:2962-4017).  Before ocpg_set_cells.
int ocpg_set_method(ocpg_ctx* c, int method)
{
    if (!c || (method != OCPG_METHOD_FIM && method != OCPG_METHOD_FIMN)) return OCPG_EINVAL;
    if (c->cur) return OCPG_ESTATE;
    c->fimn = method == OCPG_METHOD_FIMN;
    return OCPG_OK;
}

int ocpg_get_method(ocpg_ctx* c)
{
    if (!c) return OCPG_EINVAL;
    return c->fimn ? OCPG_METHOD_FIMN : OCPG_METHOD_FIM;
}

// Bulk::satcm (src/Bulk.cpp:350-373): the critical saturation of each phase.
int ocpg_critical_saturations(ocpg_ctx* c, double* scm)
{
    if (!c || !scm) return OCPG_EINVAL;
    if (!c->dSat) return OCPG_ESTATE;
    for (int j = 0; j < 3; j++) scm[j] = c->scm[j];
    return OCPG_OK;
}

// Device memory for nb cells (SoA, leading dimension rounded up to 64 cells);
// zero-filled.  Replaces any earlier allocation.
int ocpg_set_cells(ocpg_ctx* c, int64_t nb)
{
    if (!c || nb < 0 || nb > (int64_t(1) << 31)) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    dropIterGraph(c);
    if (c->cur) HIPCHK(hipFree(c->cur));
    if (c->last) HIPCHK(hipFree(c->last));
    if (c->mask) HIPCHK(hipFree(c->mask));
    if (c->copyStream) HIPCHK(hipStreamSynchronize(c->copyStream));
    if (c->poolLNi) HIPCHK(hipFree(c->poolLNi));
    c->cur = c->last = nullptr;
    c->mask = nullptr;
    c->poolLNi = nullptr;
    c->pool = false;
    c->lastValid = false;
    const int64_t ld = (nb + 63) / 64 * 64;
    size_t off = 0;
    for (int f = 0; f < kNumFields; f++) {
        c->fieldOff[f] = off;
        off += align256((size_t)kFields[f].elem * cellEntries(c, kFields[f].cnt) * (size_t)(ld > 0 ? ld : 64));
    }
    c->lastBytes = 0;
    for (int f = 0; f < kNumFields; f++)
        if (!kFields[f].lastStep) {
            c->lastBytes = c->fieldOff[f];
            break;
        }
    c->curBytes = off;
    HIPCHK(hipMalloc(&c->cur, c->curBytes));
    HIPCHK(hipMalloc(&c->last, c->lastBytes));
    {
        // FlashWork: doubles zi(NC), Nh, Yt, x(2NC), n(2NC), nu(2), Ks(NC),
        // Se; int32 ssmIt, rrIt, nrIt; bytes ftype, lNP, flags; the int32
        // lists; the counters.
        const size_t L = (size_t)(ld > 0 ? ld : 64);
        const int NC = c->NC;
        const size_t dbl[8] = {(size_t)NC, 1, 1, (size_t)2 * NC, (size_t)2 * NC, 2, (size_t)NC, 1};
        size_t w = 0;
        size_t offD[8], offI[3], offB[3], offL[kFlashLists], offC;
        for (int q = 0; q < 8; q++) { offD[q] = w; w += align256(dbl[q] * L * sizeof(double)); }
        for (int q = 0; q < 3; q++) { offI[q] = w; w += align256(L * sizeof(int32_t)); }
        for (int q = 0; q < 3; q++) { offB[q] = w; w += align256(L); }
        for (int q = 0; q < kFlashLists; q++) { offL[q] = w; w += align256(L * sizeof(int32_t)); }
        offC = w;
        w += 256;
        if (c->workMem) HIPCHK(hipFree(c->workMem));
        HIPCHK(hipMalloc(&c->workMem, w));
        c->workBytes = w;
        char* b = c->workMem;
        c->work.zi = reinterpret_cast<double*>(b + offD[0]);
        c->work.Nh = reinterpret_cast<double*>(b + offD[1]);
        c->work.Yt = reinterpret_cast<double*>(b + offD[2]);
        c->work.x = reinterpret_cast<double*>(b + offD[3]);
        c->work.n = reinterpret_cast<double*>(b + offD[4]);
        c->work.nu = reinterpret_cast<double*>(b + offD[5]);
        c->work.Ks = reinterpret_cast<double*>(b + offD[6]);
        c->work.Se = reinterpret_cast<double*>(b + offD[7]);
        c->work.ssmIt = reinterpret_cast<int32_t*>(b + offI[0]);
        c->work.rrIt = reinterpret_cast<int32_t*>(b + offI[1]);
        c->work.nrIt = reinterpret_cast<int32_t*>(b + offI[2]);
        c->work.ftype = reinterpret_cast<int8_t*>(b + offB[0]);
        c->work.lNP = reinterpret_cast<int8_t*>(b + offB[1]);
        c->work.flags = reinterpret_cast<uint8_t*>(b + offB[2]);
        for (int q = 0; q < kFlashLists; q++) c->work.list[q] = reinterpret_cast<int32_t*>(b + offL[q]);
        c->work.count = reinterpret_cast<int32_t*>(b + offC);
        HIPCHK(hipMemsetAsync(c->workMem, 0, w, c->stream));
    }
    HIPCHK(hipMemsetAsync(c->cur, 0, c->curBytes, c->stream));
    HIPCHK(hipMemsetAsync(c->last, 0, c->lastBytes, c->stream));
    std::memset(&c->B, 0, sizeof(DevBulk));
    c->B.nb = nb;
    c->B.ld = ld;
    for (int f = 0; f < kNumFields; f++) {
        void* p = cellEntries(c, kFields[f].cnt) > 0 ? c->cur + c->fieldOff[f] : nullptr;
        std::memcpy(reinterpret_cast<char*>(&c->B) + kFields[f].devOff, &p, sizeof(void*));
    }
    // IsoT_FIM::CalRes reads the last-step moles as lNi.
    for (int f = 0; f < kNumFields; f++)
        if (kFields[f].hostOff == HB(Ni)) c->B.lNi = reinterpret_cast<double*>(c->last + c->fieldOff[f]);
    std::memset(&c->Blast, 0, sizeof(DevBulk));
    c->Blast.nb = nb;
    c->Blast.ld = ld;
    for (int f = 0; f < kNumFields; f++) {
        if (!kFields[f].lastStep) continue;
        void* p = cellEntries(c, kFields[f].cnt) > 0 ? c->last + c->fieldOff[f] : nullptr;
        std::memcpy(reinterpret_cast<char*>(&c->Blast) + kFields[f].devOff, &p, sizeof(void*));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCPG_OK;
}

// Copy every non-NULL field of `h` (reference layout, h->nb cells) into the
// device state.  lNi, when given, goes to the last-step copy.
int ocpg_upload_bulk(ocpg_ctx* c, const ocpg_bulk* h)
{
    if (!c || !h || !c->cur || h->nb != c->B.nb) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const int64_t nb = h->nb, ld = c->B.ld;
    std::vector<char> tmp;
    for (int f = 0; f < kNumFields; f++) {
        const FieldDesc& d = kFields[f];
        const char* src;
        std::memcpy(&src, reinterpret_cast<const char*>(h) + d.hostOff, sizeof(void*));
        const int64_t cnt = cellEntries(c, d.cnt);
        if (!src || cnt == 0) continue;
        tmp.assign((size_t)d.elem * cnt * ld, 0);
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++)
                std::memcpy(&tmp[(size_t)(k * ld + n) * d.elem], src + (size_t)(n * cnt + k) * d.elem, d.elem);
        HIPCHK(hipMemcpy(c->cur + c->fieldOff[f], tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    }
    if (h->lNi) {
        const int64_t cnt = c->NC + 1;
        std::vector<double> t((size_t)cnt * ld, 0.0);
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++) t[k * ld + n] = h->lNi[n * cnt + k];
        HIPCHK(hipMemcpy(c->B.lNi, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return OCPG_OK;
}

int ocpg_download_bulk(ocpg_ctx* c, ocpg_bulk* h)
{
    if (!c || !h || !c->cur || h->nb != c->B.nb) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t nb = h->nb, ld = c->B.ld;
    std::vector<char> tmp;
    for (int f = 0; f < kNumFields; f++) {
        const FieldDesc& d = kFields[f];
        char* dst;
        std::memcpy(&dst, reinterpret_cast<char*>(h) + d.hostOff, sizeof(void*));
        const int64_t cnt = cellEntries(c, d.cnt);
        if (!dst || cnt == 0) continue;
        tmp.resize((size_t)d.elem * cnt * ld);
        HIPCHK(hipMemcpy(tmp.data(), c->cur + c->fieldOff[f], tmp.size(), hipMemcpyDeviceToHost));
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++)
                std::memcpy(dst + (size_t)(n * cnt + k) * d.elem, &tmp[(size_t)(k * ld + n) * d.elem], d.elem);
    }
    if (h->lNi) {
        const int64_t cnt = c->NC + 1;
        std::vector<double> t((size_t)cnt * ld);
        HIPCHK(hipMemcpy(t.data(), c->B.lNi, t.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++) h->lNi[n * cnt + k] = t[k * ld + n];
    }
    return OCPG_OK;
}

// IsoT_FIM::CalFlash (src/OCPFluidMethod.cpp:1108-1120): FlashFIM +
// PassFlashValue for every cell.  Asynchronous unless stats is non-NULL.
int ocpg_flash_fim(ocpg_ctx* c, ocpg_flash_stats* stats)
{
    if (!c || !c->cur) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    hipEvent_t ev[8] = {};
    bool stageEv = c->timing;
    for (int k = 0; k < 6 && stageEv; k++) stageEv = hipEventCreate(&ev[k]) == hipSuccess;
    timerBegin(c, T_FLASH);
    HIPCHK(launchFlash(c->NC, c->dEos, c->B, c->work, stats ? c->dCtr : nullptr, nullptr, c->stream,
                       stageEv ? ev : nullptr, nullptr, !c->fimn));
    if (c->fimn) {
        // IsoT_FIMn::CalFlash: FlashFIMn's own finishing stage
        if (stageEv) HIPCHK(hipEventRecord(ev[4], c->stream));
        HIPCHK(launchFlashFinishN(c->NC, c->dEos, c->B, c->work, 0, c->stream));
        if (stageEv) HIPCHK(hipEventRecord(ev[5], c->stream));
    }
    timerEnd(c, T_FLASH);
    if (stageEv) {
        c->events[T_STAB].push_back({ev[0], ev[1]});
        c->events[T_SPLIT].push_back({ev[2], ev[3]});
        c->events[T_FINISH].push_back({ev[4], ev[5]});
    } else {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    if (stats) {
        DevFlashCounters h;
        HIPCHK(hipMemcpyAsync(&h, c->dCtr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        stats->cells = (int64_t)h.cells;
        for (int q = 0; q < 3; q++) stats->ftype_count[q] = (int64_t)h.ftype[q];
        stats->two_phase = (int64_t)h.twoPhase;
        for (int q = 0; q < 5; q++) stats->iters[q] = (int64_t)h.iters[q];
        stats->nonconverged = (int64_t)h.nonconverged;
    }
    return OCPG_OK;
}

// FlashFIM with a learned initial guess (no reference counterpart): guess is
// device memory, nb rows of ld floats (ln K_1..ln K_NC, stability logit).
// The stability test of a cell whose logit is positive first tries the
// guessed K; the split starts from it where the reference starts from Wilson.
int ocpg_flash_fim_guided(ocpg_ctx* c, const float* guess, int ld, ocpg_flash_stats* stats)
{
    if (!c || !c->cur) return OCPG_ESTATE;
    if (!guess || ld < c->NC + 1) return OCPG_EINVAL;
    c->work.guess = guess;
    c->work.guessLd = ld;
    const int rc = ocpg_flash_fim(c, stats);
    c->work.guess = nullptr;
    c->work.guessLd = 0;
    return rc;
}

// The learned form's network for this context (no reference counterpart):
// host FP32 arrays W1 (nin x 64), b1 (64), W2, W3 (64 x 64), b2, b3 (64),
// W4 (64 x nout), b4 (nout), with nin = NC + 2 and nout = NC + 1; copied to
// the device.  All NULL removes it.
int ocpg_set_learned(ocpg_ctx* c, int nin, int nout, const float* W1, const float* b1, const float* W2,
                     const float* b2, const float* W3, const float* b3, const float* W4, const float* b4)
{
    if (!c) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const float* src[8] = {W1, b1, W2, b2, W3, b3, W4, b4};
    bool none = true, all = true;
    for (const float* p : src) {
        none = none && !p;
        all = all && p;
    }
    if (!none && !all) return OCPG_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->mlpMem) HIPCHK(hipFree(c->mlpMem));
    c->mlpMem = nullptr;
    for (auto& p : c->mlpW) p = nullptr;
    if (none) return OCPG_OK;
    if (nin != c->NC + 2 || nout != c->NC + 1) return OCPG_EINVAL;
    constexpr int H = 64;
    const size_t n[8] = {(size_t)nin * H, H, (size_t)H * H, H, (size_t)H * H, H, (size_t)H * nout, (size_t)nout};
    size_t off[8], tot = 0;
    for (int k = 0; k < 8; k++) {
        off[k] = tot;
        tot += (n[k] * sizeof(float) + 255) & ~size_t(255);
    }
    char* m = nullptr;
    HIPCHK(hipMalloc(&m, tot));
    c->mlpMem = reinterpret_cast<float*>(m);
    for (int k = 0; k < 8; k++) {
        HIPCHK(hipMemcpy(m + off[k], src[k], n[k] * sizeof(float), hipMemcpyHostToDevice));
        c->mlpW[k] = reinterpret_cast<const float*>(m + off[k]);
    }
    return OCPG_OK;
}

// FlashFIM with the learned initial guess computed on the device from the
// cells' own P, T and moles (ocpg_set_learned's network in one fused
// kernel), then ocpg_flash_fim_guided with it.  Asynchronous unless stats
// is non-NULL.
int ocpg_flash_fim_learned(ocpg_ctx* c, ocpg_flash_stats* stats)
{
    if (!c || !c->cur || !c->mlpMem) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    const int64_t nb = c->B.nb;
    const int ldG = c->NC + 1;
    if (c->guessRows != nb) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->guessBuf) HIPCHK(hipFree(c->guessBuf));
        c->guessBuf = nullptr;
        c->guessRows = 0;
        HIPCHK(hipMalloc(&c->guessBuf, (size_t)std::max<int64_t>(nb, 1) * ldG * sizeof(float)));
        c->guessRows = nb;
    }
    timerBegin(c, T_MLP);
    HIPCHK(launchMlpForwardCells(c->stream, nb, c->B.ld, c->NC, c->B.P, c->B.T, c->B.Ni, c->mlpW, c->guessBuf, ldG));
    timerEnd(c, T_MLP);
    return ocpg_flash_fim_guided(c, c->guessBuf, ldG, stats);
}

// IsoT_FIM::UpdateProperty (CalFlash, CalKrPc, CalRock, CalRes;
// src/OCPFluidMethod.cpp:877-897) followed by AssembleMat (:860-875): the
// device work of one Newton iteration, with the flash's slowest cells
// overlapped with the rest.  The stability test and the splits that run far
// longer than the others continue on two side streams while the main stream
// finishes the other cells, runs kr/Pc, rock, CalRes and the assembly on
// every row; after the side streams join, their cells are finished and the
// rows that read them (those cells, their neighbours and, with wells, the
// perforated cells) are computed again.  Every row ends as the serial calls
// leave it.  Asynchronous unless norms or stats is non-NULL.
int ocpg_update_assemble(ocpg_ctx* c, double dt, ocpg_resnorm* norms, ocpg_flash_stats* stats)
{
    if (!c || !c->cur) return OCPG_ESTATE;
    FimDevice& F = c->fim;
    if (F.nb <= 0 || F.nb != c->B.nb) return OCPG_ESTATE;
    if (!c->dSat) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    int rc = ensureFixup(c);
    if (rc != OCPG_OK) return rc;
    WellDevice& W = c->wells;
    hipEvent_t ev[10] = {};
    bool stageEv = c->timing;
    for (int k = 0; k < 10 && stageEv; k++) stageEv = hipEventCreate(&ev[k]) == hipSuccess;
    timerBegin(c, T_UPDASM);
    timerBegin(c, T_FLASH);
    HIPCHK(launchFlash(c->NC, c->dEos, c->B, c->work, stats ? c->dCtr : nullptr, nullptr, c->stream,
                       stageEv ? ev : nullptr, &c->side, !c->fimn));
    if (c->fimn) {
        // IsoT_FIMn: FlashFIMn's finishing stage of the main stream's cells
        if (stageEv) HIPCHK(hipEventRecord(ev[4], c->stream));
        HIPCHK(launchFlashFinishN(c->NC, c->dEos, c->B, c->work, 0, c->stream));
        if (stageEv) HIPCHK(hipEventRecord(ev[5], c->stream));
    }
    timerEnd(c, T_FLASH);
    timerBegin(c, T_KRPC);
    HIPCHK(launchKrPcRock(c->B, c->dSat, c->rock, true, true, c->stream));
    timerEnd(c, T_KRPC);
    // CalRes and the records the assembly reads, side by side in one launch
    // (timed as "calres"; "pack" stays empty here)
    timerBegin(c, T_RES);
    HIPCHK(launchCalResPack(c->NC, c->B, F, W, dt, c->stream));
    timerEnd(c, T_RES);
    timerBegin(c, T_ASM);
    if (c->fimn)
        HIPCHK(launchAssembleN(c->NC, c->B, F, dt, c->stream));
    else
        HIPCHK(launchAssemble(c->NC, c->B, F, dt, c->stream));
    timerEnd(c, T_ASM);
    // join the side streams, finish their cells, recompute what read them
    HIPCHK(hipStreamWaitEvent(c->stream, c->side.doneB, 0));
    HIPCHK(hipStreamWaitEvent(c->stream, c->side.doneC, 0));
    HIPCHK(hipStreamWaitEvent(c->stream, c->side.doneD, 0));
    timerBegin(c, T_FIXUP);
    if (c->fimn)
        HIPCHK(launchFlashSideFinishN(c->NC, c->dEos, c->B, c->work, c->stream));
    else
        HIPCHK(launchFlashSideFinish(c->NC, c->dEos, c->B, c->work, c->stream));
    const CellLists sideCells = flashSideCells(c->work);
    HIPCHK(launchKrPcRockList(c->B, c->dSat, c->rock, sideCells, c->stream));
    HIPCHK(launchFixupRows(F, W, sideCells, c->fix, c->stream));
    if (W.nw > 0) HIPCHK(launchWellFlux(c->NC, c->dEos, c->B, F, W, c->stream));
    HIPCHK(launchCalResRows(c->NC, c->B, F, W, dt, c->fix.resRows, c->fix.counts + 0, c->stream));
    HIPCHK(launchPackRecordsList(c->NC, c->B, F, sideCells, c->stream));
    HIPCHK(launchAssembleRows(c->NC, c->B, F, dt, c->fix.asmRows, c->fix.counts + 1, c->stream));
    if (W.nw > 0) HIPCHK(launchWellAssemble(c->NC, c->B, F, W, dt, c->stream));
    if (c->fimn) HIPCHK(launchRhsAccumulate(F, c->stream));
    HIPCHK(launchResNorm(c->NC, c->B, F, c->stream));
    timerEnd(c, T_FIXUP);
    timerEnd(c, T_UPDASM);
    if (stageEv) {
        c->events[T_STAB].push_back({ev[0], ev[1]});
        c->events[T_SPLIT].push_back({ev[2], ev[3]});
        c->events[T_FINISH].push_back({ev[4], ev[5]});
        c->events[T_TAIL].push_back({ev[6], ev[7]});
        c->events[T_TAILNR].push_back({ev[8], ev[9]});
    } else {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    if (norms && (rc = fimReadNorms(c, norms)) != OCPG_OK) return rc;
    if (stats) {
        DevFlashCounters h;
        HIPCHK(hipMemcpyAsync(&h, c->dCtr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        stats->cells = (int64_t)h.cells;
        for (int q = 0; q < 3; q++) stats->ftype_count[q] = (int64_t)h.ftype[q];
        stats->two_phase = (int64_t)h.twoPhase;
        for (int q = 0; q < 5; q++) stats->iters[q] = (int64_t)h.iters[q];
        stats->nonconverged = (int64_t)h.nonconverged;
    }
    return OCPG_OK;
}

// IsoT_FIM::InitFlash (src/OCPFluidMethod.cpp:1095-1106): InitFlashFIM from
// P, T, Sw (S column 2), rockVp and the feed z (host, nb x NC row-major).
// Ni becomes an output.  IsoT_FIMn::InitFlash (:3129-3150, InitFlashFIMn)
// in an IsoT_FIMn context.
int ocpg_init_flash_fim(ocpg_ctx* c, const double* zi)
{
    if (!c || !c->cur) return OCPG_ESTATE;
    if (!zi) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const int64_t nb = c->B.nb, ld = c->B.ld, NC = c->NC;
    std::vector<double> t((size_t)NC * ld, 0.0);
    for (int64_t n = 0; n < nb; n++)
        for (int64_t i = 0; i < NC; i++) t[i * ld + n] = zi[n * NC + i];
    double* dz = nullptr;
    HIPCHK(hipMalloc(&dz, t.size() * sizeof(double)));
    hipError_t e = hipMemcpy(dz, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice);
    timerBegin(c, T_INIT);
    if (e == hipSuccess) e = launchFlash(c->NC, c->dEos, c->B, c->work, nullptr, dz, c->stream, nullptr, nullptr,
                                         !c->fimn);
    if (e == hipSuccess && c->fimn) e = launchFlashFinishN(c->NC, c->dEos, c->B, c->work, 1, c->stream);
    timerEnd(c, T_INIT);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dz);
    return hipStatus(e);
}

// IsoT_FIM::UpdateLastTimeStep (src/OCPFluidMethod.cpp:2915-2957) plus the
// skip-stability memory: device-to-device copy into the last-step block.
int ocpg_update_last_step(ocpg_ctx* c)
{
    if (!c || !c->cur || c->pool) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->last, c->cur, c->lastBytes, hipMemcpyDeviceToDevice, c->stream));
    // AllWells::UpdateLastTimeStepBHP
    if (c->wells.nw > 0)
        HIPCHK(hipMemcpyAsync(c->wells.lbhp, c->wells.bhp, (size_t)c->wells.nw * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    c->lastValid = true;
    return OCPG_OK;
}

// IsoT_FIM::ResetToLastTimeStep (src/OCPFluidMethod.cpp:2859-2913): the
// Bulk and skip-stability state, then the wells' AllWells::ResetBHP,
// CalTrans and (with ocpg_set_well_setup) CaldG on the restored state; the
// caller's ocpg_calres does CalFlux and CalRes.
int ocpg_reset_to_last_step(ocpg_ctx* c)
{
    if (!c || !c->cur || !c->lastValid || c->pool) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->cur, c->last, c->lastBytes, hipMemcpyDeviceToDevice, c->stream));
    if (c->wells.nw > 0) HIPCHK(launchWellPrepare(c->NC, c->dEos, c->B, c->wells, 1, c->stream));
    return OCPG_OK;
}

// Copy of the fields the flash starts from (P, Ni, S, phaseNum, xij and the
// skip-stability memory), so that one Newton iteration can be re-run from
// the same state.  No reference counterpart.
int ocpg_save_flash_inputs(ocpg_ctx* c)
{
    if (!c || !c->cur) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    size_t need = 0;
    for (int f = 0; f < kNumFields; f++)
        if (isFlashInput(kFields[f].hostOff))
            need += align256((size_t)kFields[f].elem * cellEntries(c, kFields[f].cnt) * c->B.ld);
    if (need != c->snapBytes) {
        HIPCHK(hipStreamSynchronize(c->stream));
        dropIterGraph(c);
        if (c->snap) HIPCHK(hipFree(c->snap));
        c->snap = nullptr;
        HIPCHK(hipMalloc(&c->snap, need));
        c->snapBytes = need;
    }
    size_t off = 0;
    for (int f = 0; f < kNumFields; f++) {
        if (!isFlashInput(kFields[f].hostOff)) continue;
        const size_t b = (size_t)kFields[f].elem * cellEntries(c, kFields[f].cnt) * c->B.ld;
        HIPCHK(hipMemcpyAsync(c->snap + off, c->cur + c->fieldOff[f], b, hipMemcpyDeviceToDevice, c->stream));
        off += align256(b);
    }
    return OCPG_OK;
}

int ocpg_restore_flash_inputs(ocpg_ctx* c)
{
    if (!c || !c->cur || !c->snap) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    timerBegin(c, T_RESTORE);
    if (c->mask) {
        // a serving pool: only the cells of the realisations taking part
        MaskedCopy d{};
        size_t off = 0;
        for (int f = 0; f < kNumFields; f++) {
            if (!isFlashInput(kFields[f].hostOff)) continue;
            const size_t b = (size_t)kFields[f].elem * cellEntries(c, kFields[f].cnt) * c->B.ld;
            d.src[d.nfields] = c->snap + off;
            d.dst[d.nfields] = c->cur + c->fieldOff[f];
            d.elem[d.nfields] = kFields[f].elem;
            d.rows[d.nfields] = (int)cellEntries(c, kFields[f].cnt);
            d.nfields++;
            off += align256(b);
        }
        HIPCHK(launchMaskedCopy(c->mask, c->B.nb, c->B.ld, d, c->stream));
        timerEnd(c, T_RESTORE);
        return OCPG_OK;
    }
    size_t off = 0;
    for (int f = 0; f < kNumFields; f++) {
        if (!isFlashInput(kFields[f].hostOff)) continue;
        const size_t b = (size_t)kFields[f].elem * cellEntries(c, kFields[f].cnt) * c->B.ld;
        HIPCHK(hipMemcpyAsync(c->cur + c->fieldOff[f], c->snap + off, b, hipMemcpyDeviceToDevice, c->stream));
        off += align256(b);
    }
    timerEnd(c, T_RESTORE);
    return OCPG_OK;
}

// Per-kernel timing with HIP events on the context's stream.
int ocpg_timing_enable(ocpg_ctx* c, int on)
{
    if (!c) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    timerClear(c);
    c->timing = on != 0;
    return OCPG_OK;
}

// Total milliseconds and launch count recorded for a timer ("flash",
// "krpc_rock", "calres", "assemble", "init_flash", "restore").
int ocpg_timing_get(ocpg_ctx* c, const char* name, double* ms, int64_t* count)
{
    if (!c || !name) return OCPG_EINVAL;
    int slot = -1;
    for (int s = 0; s < kTimerSlots; s++)
        if (!std::strcmp(name, kTimerNames[s])) slot = s;
    if (slot < 0) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0;
    for (auto& p : c->events[slot]) {
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, p.first, p.second));
        tot += t;
    }
    if (ms) *ms = tot;
    if (count) *count = (int64_t)c->events[slot].size();
    return OCPG_OK;
}

int ocpg_synchronize(ocpg_ctx* c)
{
    if (!c) return OCPG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->launchesDone = c->launches;
    return OCPG_OK;
}

// One Newton iteration on the resident state -- restore the saved inputs,
// FlashFIM + PassFlashValue, CalKrPc + CalRock, the wells' rates and CalRes,
// AssembleMat -- captured once into a HIP graph on the context's stream, so
// a serving loop replays it with one launch.  Needs the connections, the
// saved flash inputs and timing off.
int ocpg_capture_iteration(ocpg_ctx* c, double dt)
{
    if (!c || !c->cur || !c->snap || c->timing) return OCPG_ESTATE;
    if (c->fim.nb != c->B.nb || c->fim.nb <= 0) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    dropIterGraph(c);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int rc = ocpg_restore_flash_inputs(c);
    if (rc == OCPG_OK) rc = ocpg_flash_fim(c, nullptr);
    if (rc == OCPG_OK) rc = ocpg_calkrpc_rock(c);
    if (rc == OCPG_OK) rc = ocpg_calres(c, dt, nullptr);
    if (rc == OCPG_OK) rc = ocpg_assemble(c, dt);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != OCPG_OK) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPCHK(e);
    const hipError_t ei = hipGraphInstantiate(&c->iterGraph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ei);
    if (!c->iterDone) HIPCHK(hipEventCreateWithFlags(&c->iterDone, hipEventDisableTiming));
    return OCPG_OK;
}

// Replays the captured iteration (asynchronous); ocpg_iteration_done polls it.
int ocpg_launch_iteration(ocpg_ctx* c)
{
    if (!c || !c->iterGraph) return OCPG_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipGraphLaunch(c->iterGraph, c->stream));
    HIPCHK(hipEventRecord(c->iterDone, c->stream));
    c->launches++;
    return OCPG_OK;
}

// 1 when the last launched iteration has finished, 0 while it runs.
int ocpg_iteration_done(ocpg_ctx* c)
{
    if (!c || !c->iterDone) return OCPG_ESTATE;
    const hipError_t e = hipEventQuery(c->iterDone);
    if (e == hipSuccess) {
        c->launchesDone = c->launches;
        return 1;
    }
    if (e == hipErrorNotReady) return 0;
    return hipStatus(e);
}

// Internal access for the C++ shim and the benchmark: the context's stream
// and device cell state.
void* ocpg_stream(ocpg_ctx* c) { return c ? (void*)c->stream : nullptr; }
int64_t ocpg_device_bytes(ocpg_ctx* c)
{
    return c ? (int64_t)(c->curBytes + c->lastBytes + c->workBytes + c->snapBytes + fimBytes(c->fim) +
                         c->wells.bytes)
             : -1;
}

}  // extern "C"

// Accessors used by the FIM half of the C-ABI (fim_capi.cpp).
namespace ocpg {
DevBulk& ctxBulk(ocpg_ctx* c) { return c->B; }
bool ctxFimn(ocpg_ctx* c) { return c->fimn; }
const double* ctxScm(ocpg_ctx* c) { return c->scm; }
const DevBulk& ctxLastBulk(ocpg_ctx* c) { return c->Blast; }
bool ctxLastValid(ocpg_ctx* c) { return c->lastValid; }
FimDevice& ctxFim(ocpg_ctx* c) { return c->fim; }
WellDevice& ctxWells(ocpg_ctx* c) { return c->wells; }
FimHost& ctxFimHost(ocpg_ctx* c) { return c->fimHost; }
NewtonDevice& ctxNewton(ocpg_ctx* c) { return c->newton; }
const DevEos* ctxEos(ocpg_ctx* c) { return c->dEos; }
hipStream_t ctxStream(ocpg_ctx* c) { return c->stream; }
int ctxNC(ocpg_ctx* c) { return c->NC; }
int ctxDevice(ocpg_ctx* c) { return c->device; }
const DevSat* ctxSat(ocpg_ctx* c) { return c->dSat; }
const DevRock& ctxRock(ocpg_ctx* c) { return c->rock; }
void ctxTimerBegin(ocpg_ctx* c, int slot) { timerBegin(c, slot); }
void ctxTimerEnd(ocpg_ctx* c, int slot) { timerEnd(c, slot); }
void ctxDropGraph(ocpg_ctx* c) { dropIterGraph(c); }

// The serving pool's activity mask (allocated zero on first use; B.mask then
// points at it, which changes what a captured iteration reads: the graph is
// dropped).
uint8_t* ctxEnsureMask(ocpg_ctx* c)
{
    if (c->mask) return c->mask;
    const size_t n = (size_t)std::max<int64_t>(c->B.ld, 64);
    if (hipMalloc(&c->mask, n) != hipSuccess) return nullptr;
    if (hipMemset(c->mask, 0, n) != hipSuccess) return nullptr;
    c->B.mask = c->mask;
    dropIterGraph(c);
    return c->mask;
}

// A pool's cell state: the last-step copy shrinks to the moles CalRes reads
// (a pool runs single Newton iterations, not time steps).
int ctxPoolTrimLast(ocpg_ctx* c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->last) HIPCHK(hipFree(c->last));
    c->last = nullptr;
    c->lastBytes = 0;
    c->lastValid = false;
    std::memset(&c->Blast, 0, sizeof(DevBulk));
    const size_t n = (size_t)(c->NC + 1) * (size_t)std::max<int64_t>(c->B.ld, 64);
    HIPCHK(hipMalloc(&c->poolLNi, n * sizeof(double)));
    c->B.lNi = c->poolLNi;
    c->pool = true;
    return OCPG_OK;
}

// The pool's copy stream (created on first use) and the orderings between
// it and the context's stream, by events: copies after the work queued so
// far on the context's stream, and the context's later work after the
// copies queued so far.
hipStream_t ctxCopyStream(ocpg_ctx* c)
{
    if (!c->copyStream) {
        if (hipStreamCreateWithFlags(&c->copyStream, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&c->copyEv, hipEventDisableTiming) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&c->mainEv, hipEventDisableTiming) != hipSuccess) return nullptr;
    }
    return c->copyStream;
}
int ctxCopyAfterMain(ocpg_ctx* c)
{
    HIPCHK(hipEventRecord(c->mainEv, c->stream));
    HIPCHK(hipStreamWaitEvent(c->copyStream, c->mainEv, 0));
    return OCPG_OK;
}
int ctxMainAfterCopy(ocpg_ctx* c)
{
    HIPCHK(hipEventRecord(c->copyEv, c->copyStream));
    HIPCHK(hipStreamWaitEvent(c->stream, c->copyEv, 0));
    return OCPG_OK;
}
uint64_t ctxLaunches(ocpg_ctx* c) { return c->launches; }
uint64_t ctxLaunchesDone(ocpg_ctx* c) { return c->launchesDone; }

// Device segments of the cell state of cells [base, base + nb) that a
// realisation image holds, in image order: with `state`, the saved
// iteration inputs and the last-step moles (what an iteration may change
// between admissions); with `fixed`, T, v, depth and poroInit.
void ctxImageSegments(ocpg_ctx* c, int64_t base, int64_t nb, bool state, bool fixed, std::vector<ImageSeg>& out)
{
    const int64_t ld = c->B.ld;
    size_t snapOff = 0;
    for (int f = 0; f < kNumFields; f++) {
        const FieldDesc& d = kFields[f];
        const int64_t cnt = cellEntries(c, d.cnt);
        const size_t fieldBytes = (size_t)d.elem * cnt * ld;
        if (isFlashInput(d.hostOff)) {
            if (state)
                out.push_back({c->snap + snapOff + (size_t)base * d.elem, (size_t)ld * d.elem, (size_t)nb * d.elem,
                               (size_t)cnt});
            snapOff += align256(fieldBytes);
        }
    }
    if (state)
        out.push_back({reinterpret_cast<char*>(c->B.lNi + base), (size_t)ld * 8, (size_t)nb * 8,
                       (size_t)(c->NC + 1)});
    if (fixed)
        for (int f = 0; f < kNumFields; f++) {
            const FieldDesc& d = kFields[f];
            if (d.hostOff == HB(T) || d.hostOff == HB(v) || d.hostOff == HB(depth) || d.hostOff == HB(poroInit))
                out.push_back({c->cur + c->fieldOff[f] + (size_t)base * d.elem, (size_t)ld * d.elem,
                               (size_t)nb * d.elem, (size_t)cellEntries(c, d.cnt)});
        }
}

// Back to an ordinary context: every cell takes part.
int ctxDropMask(ocpg_ctx* c)
{
    if (!c->mask) return OCPG_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(c->mask));
    c->mask = nullptr;
    c->B.mask = nullptr;
    dropIterGraph(c);
    return OCPG_OK;
}

// Copies of the fields of cells [base, base + h->nb) between the host (the
// reference's cell-major layout, h->nb cells) and the device: toDevice
// uploads every non-NULL field; otherwise, with saved, the iteration inputs
// come from the saved copy (ocpg_save_flash_inputs) and the rest from the
// current state, and without it everything from the current state.
// Transposes cnt x nb entries of E bytes between the host's cell-major
// layout (src/dst[n * cnt + k]) and a field-major staging buffer.
template <int E>
static void transposeRange(char* dev, char* host, int64_t nb, int64_t cnt, bool toStage)
{
    struct alignas(E) V { char b[E]; };
    V* d = reinterpret_cast<V*>(dev);
    V* h = reinterpret_cast<V*>(host);
    if (toStage) {
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++) d[k * nb + n] = h[n * cnt + k];
    } else {
        for (int64_t n = 0; n < nb; n++)
            for (int64_t k = 0; k < cnt; k++) h[n * cnt + k] = d[k * nb + n];
    }
}

static void transposeAny(int elem, char* stage, char* host, int64_t nb, int64_t cnt, bool toStage)
{
    if (elem == 8) transposeRange<8>(stage, host, nb, cnt, toStage);
    else if (elem == 4) transposeRange<4>(stage, host, nb, cnt, toStage);
    else transposeRange<1>(stage, host, nb, cnt, toStage);
}

int ctxRangeCopy(ocpg_ctx* c, int64_t base, ocpg_bulk* h, bool toDevice, bool saved)
{
    const int64_t nb = h->nb, ld = c->B.ld;
    if (base < 0 || nb < 0 || base + nb > c->B.nb) return OCPG_EINVAL;
    if (nb == 0) return OCPG_OK;
    std::vector<double> stage;
    size_t snapOff = 0;
    auto move = [&](char* devBase, char* hp, int elem, int64_t cnt) -> hipError_t {
        stage.resize(((size_t)elem * cnt * nb + 7) / 8);
        char* st = reinterpret_cast<char*>(stage.data());
        const size_t w = (size_t)elem * nb, pitch = (size_t)elem * ld;
        char* dp = devBase + (size_t)base * elem;
        if (toDevice) {
            transposeAny(elem, st, hp, nb, cnt, true);
            return hipMemcpy2D(dp, pitch, st, w, w, cnt, hipMemcpyHostToDevice);
        }
        const hipError_t e = hipMemcpy2D(st, w, dp, pitch, w, cnt, hipMemcpyDeviceToHost);
        if (e == hipSuccess) transposeAny(elem, st, hp, nb, cnt, false);
        return e;
    };
    for (int f = 0; f < kNumFields; f++) {
        const FieldDesc& d = kFields[f];
        const int64_t cnt = cellEntries(c, d.cnt);
        const size_t fieldBytes = (size_t)d.elem * cnt * ld;
        const bool input = isFlashInput(d.hostOff);
        char* devBase = c->cur + c->fieldOff[f];
        if (input && !toDevice && saved) devBase = c->snap + snapOff;
        if (input) snapOff += align256(fieldBytes);
        char* hp;
        std::memcpy(&hp, reinterpret_cast<char*>(h) + d.hostOff, sizeof(void*));
        if (hp && cnt > 0) HIPCHK(move(devBase, hp, d.elem, cnt));
    }
    if (h->lNi) HIPCHK(move(reinterpret_cast<char*>(c->B.lNi), reinterpret_cast<char*>(h->lNi), 8, c->NC + 1));
    return OCPG_OK;
}

// Saved iteration inputs of cells [base, base + nb): current -> saved copy,
// on `stream`.
int ctxSaveInputsRange(ocpg_ctx* c, int64_t base, int64_t nb, hipStream_t stream)
{
    if (!c->snap) return OCPG_ESTATE;
    const int64_t ld = c->B.ld;
    size_t off = 0;
    for (int f = 0; f < kNumFields; f++) {
        const FieldDesc& d = kFields[f];
        if (!isFlashInput(d.hostOff)) continue;
        const int64_t cnt = cellEntries(c, d.cnt);
        const size_t o = (size_t)base * d.elem, pitch = (size_t)ld * d.elem;
        HIPCHK(hipMemcpy2DAsync(c->snap + off + o, pitch, c->cur + c->fieldOff[f] + o, pitch, (size_t)nb * d.elem,
                                cnt, hipMemcpyDeviceToDevice, stream));
        off += align256((size_t)d.elem * cnt * ld);
    }
    return OCPG_OK;
}
}  // namespace ocpg
