/*
 * ocpgpu.h — C-ABI of the MI355X (gfx950) EoS flash + FIM assembly library.
 *
 * The library replaces the per-cell compositional flash and the fully implicit
 * residual/Jacobian assembly of OpenCAEPoro 0.5.0.  Every entry point below
 * names the reference interface it stands in for (paths are relative to the
 * OpenCAEPoro source tree).  Plain pointers and sizes only: no C++ or torch
 * types cross this boundary.  Every call returns OCPG_OK (0) or a negative
 * OCPG_E* code; no call aborts the process.
 *
 * Units and meanings follow the reference: P in psia, T in degrees F at the
 * boundary (the flash adds 459.67 internally, src/MixtureComp.cpp:513), moles
 * in lbmol, volumes in ft3.  Hydrocarbon components are 0..NC-1, water is
 * component NC (always appended, src/MixtureComp.cpp:34-35); phases are
 * oil=0, gas=1, water=2 (include/OCPConst.hpp:91-93).  nc = NC+1, np = 3,
 * ncol = nc+1 (block dimension of the Jacobian).
 */
#ifndef OCPGPU_H
#define OCPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCPG_MAXNC 16   /* largest hydrocarbon component count accepted      */
#define OCPG_NP 3       /* oil, gas, water                                   */
#define OCPG_MAXTAB 64  /* rows per saturation / PVTW table                  */

/* Status codes. */
#define OCPG_OK 0
#define OCPG_EINVAL (-1)     /* bad argument (null pointer, size, shape)       */
#define OCPG_EUNSUP (-2)     /* NC without a compiled kernel instantiation     */
#define OCPG_EHIP (-3)       /* HIP runtime error (allocation, launch, copy)   */
#define OCPG_ESTATE (-4)     /* call order: e.g. assemble before set_grid      */
#define OCPG_ENODEV (-5)     /* no gfx950 device visible                       */
#define OCPG_ETOOBIG (-6)    /* request exceeds a documented maximum           */

/* Equation-of-state description, as MixtureComp holds it after its
 * constructor (src/MixtureComp.cpp:28-165): Vc already derived from Zc,
 * Vshift already scaled by R*OmegaB*Tc/Pc, LBCcoef already multiplied by 10,
 * BIC already symmetrised.  ocpg_eos_from_deck() performs those steps from
 * deck-level values.                                                         */
typedef struct ocpg_eos {
    int32_t nc;                         /* NC, hydrocarbon components           */
    int32_t pvtw_rows;                  /* rows used in pvtw                    */
    double Tc[OCPG_MAXNC];              /* degR                                 */
    double Pc[OCPG_MAXNC];              /* psia                                 */
    double Vc[OCPG_MAXNC];              /* ft3/lbmol                            */
    double MW[OCPG_MAXNC];
    double Acf[OCPG_MAXNC];
    double OmegaA[OCPG_MAXNC];
    double OmegaB[OCPG_MAXNC];
    double Vshift[OCPG_MAXNC];          /* ft3/lbmol                            */
    double Vcvis[OCPG_MAXNC];
    double BIC[OCPG_MAXNC * OCPG_MAXNC];/* row-major NC x NC, symmetric         */
    double LBCcoef[5];
    /* EoS solver controls: SSMSTA / NRSTA / SSMSP / NRSP / RR keywords
     * (include/MixtureComp.hpp:28-104, src/MixtureComp.cpp:140-163).          */
    int32_t ssmsta_maxit; double ssmsta_tol; double ssmsta_eyt;
    int32_t nrsta_maxit;  double nrsta_tol;
    int32_t ssmsp_maxit;  double ssmsp_tol;
    int32_t nrsp_maxit;   double nrsp_tol;
    int32_t rr_maxit;     double rr_tol;
    /* PVTW table rows: Pref, Bw, Cw, muw, Cvw (src/MixtureComp.cpp:521-546). */
    double pvtw[OCPG_MAXTAB][5];
    double std_rhow;                    /* lb/ft3 at surface                    */
} ocpg_eos;

/* Deck-level component data, as ParamReservoir's ComponentParam holds it
 * (include/ParamReservoir.hpp:86-193, keywords TCRIT/PCRIT/VCRIT/ZCRIT/MW/
 * ACF/OMEGAA/OMEGAB/SSHIFT/VCRITVIS/ZCRITVIS/BIC/LBCCOEF).  has_* flags say
 * which optional keywords were given.                                        */
typedef struct ocpg_eos_deck {
    int32_t nc;
    int32_t pvtw_rows;
    double Tc[OCPG_MAXNC];
    double Pc[OCPG_MAXNC];
    double MW[OCPG_MAXNC];
    double Acf[OCPG_MAXNC];
    int32_t has_Vc;     double Vc[OCPG_MAXNC];
    int32_t has_Zc;     double Zc[OCPG_MAXNC];
    int32_t has_OmegaA; double OmegaA[OCPG_MAXNC];
    int32_t has_OmegaB; double OmegaB[OCPG_MAXNC];
    int32_t has_Vshift; double Vshift[OCPG_MAXNC]; /* dimensionless SSHIFT    */
    int32_t has_Vcvis;  double Vcvis[OCPG_MAXNC];
    int32_t has_Zcvis;  double Zcvis[OCPG_MAXNC];
    double BIC_lower[OCPG_MAXNC * (OCPG_MAXNC - 1) / 2]; /* strict lower, row by row */
    double LBCcoef[5];  /* deck values (ParamReservoir defaults 0.1023, 0.023364,
                           0.058533, -0.040758, 0.0093324)                    */
    int32_t ssmsta_maxit; double ssmsta_tol; double ssmsta_eyt;
    int32_t nrsta_maxit;  double nrsta_tol;
    int32_t ssmsp_maxit;  double ssmsp_tol;
    int32_t nrsp_maxit;   double nrsp_tol;
    int32_t rr_maxit;     double rr_tol;
    double pvtw[OCPG_MAXTAB][5];
    double std_rhow;
} ocpg_eos_deck;

/* FlowUnit_ODGW01 tables (src/FlowUnit.cpp:178-299): SWOF rows
 * (Sw, krw, krow, Pcow) and SGOF rows (Sg, krg, krog, Pcgo); Stone II.        */
typedef struct ocpg_satfunc {
    int32_t swof_rows;
    int32_t sgof_rows;
    double swof[OCPG_MAXTAB][4];
    double sgof[OCPG_MAXTAB][4];
} ocpg_satfunc;

/* Rock_Linear (src/Rock.cpp:19-27): poro = poroInit*(1+(cp1+cp2/2*dP)*dP). */
typedef struct ocpg_rock {
    double Pref;
    double cp1;
    double cp2;
} ocpg_rock;

/* Cell ("bulk") state in the reference's host layout (include/Bulk.hpp:98-102,
 * src/OCPFluidMethod.cpp:980-1084): cell-major, e.g. Ni[n*nc+i],
 * S[n*np+j], xij[(n*np+j)*nc+i], dSec_dPri[n*maxLendSdP + ...] with
 * maxLendSdP = (nc+1)^2*np.  Any pointer may be NULL where a call does not
 * read or write that field.                                                  */
typedef struct ocpg_bulk {
    int64_t nb;
    double* P;            /* nb                                                 */
    double* T;            /* nb, degF                                           */
    double* Ni;           /* nb*nc                                              */
    double* lNi;          /* nb*nc, moles at the last time step                 */
    double* S;            /* nb*np                                              */
    int32_t* phaseNum;    /* nb: phases present incl. water                     */
    double* xij;          /* nb*np*nc                                           */
    uint8_t* phaseExist;  /* nb*np                                              */
    double* Nt;           /* nb                                                 */
    double* vf;           /* nb                                                 */
    double* vfP;          /* nb                                                 */
    double* vfi;          /* nb*nc                                              */
    double* rho;          /* nb*np                                              */
    double* xi;           /* nb*np                                              */
    double* mu;           /* nb*np                                              */
    double* rhoP;         /* nb*np                                              */
    double* xiP;          /* nb*np                                              */
    double* muP;          /* nb*np                                              */
    double* rhox;         /* nb*np*nc                                           */
    double* xix;          /* nb*np*nc                                           */
    double* mux;          /* nb*np*nc                                           */
    uint8_t* pSderExist;  /* nb*np                                              */
    int32_t* pVnumCom;    /* nb*np                                              */
    int32_t* bRowSizedSdP;/* nb                                                 */
    double* dSec_dPri;    /* nb*maxLendSdP                                      */
    /* SkipStaAnaly state (include/AcceleratePVT.hpp:71-85)                    */
    uint8_t* skipFlag;    /* nb                                                 */
    double* skipMinEig;   /* nb (an FP32 eigenvalue held in FP64)               */
    double* skipP;        /* nb                                                 */
    double* skipT;        /* nb, degR                                           */
    double* skipZi;       /* nb*NC                                              */
    /* saturation functions (IsoT_FIM::CalKrPc)                                */
    double* kr;           /* nb*np                                              */
    double* Pc;           /* nb*np                                              */
    double* Pj;           /* nb*np                                              */
    double* dKr_dS;       /* nb*np*np                                           */
    double* dPcj_dS;      /* nb*np*np                                           */
    /* rock and geometry                                                       */
    double* v;            /* nb, bulk volume                                    */
    double* depth;        /* nb                                                 */
    double* poroInit;     /* nb                                                 */
    double* poro;         /* nb                                                 */
    double* poroP;        /* nb                                                 */
    double* rockVp;       /* nb                                                 */
    /* diagnostics (no reference counterpart; NULL to skip)                    */
    int32_t* ftype;       /* nb: 0 full test, 1 skipped, 2 straight to split    */
    int32_t* iters;       /* nb*5: SSMSTA, NRSTA, SSMSP, NRSP, RR iterations    */
    int32_t* nonconv;     /* nb: 1 where the last flash did not converge (the
                             cell's MixtureComp::ePEC would read 1e8,
                             src/MixtureComp.cpp:1301-1346), 2 where its
                             results are not finite, else 0                    */
    /* IsoT_FIMn state (src/OCPFluidMethod.cpp:3111-3127): phase moles, the
     * flash's secondary-variable correction in dSec_dPri's row order (its
     * first bRowSizedSdP entries are set) and its volume-balance term       */
    double* nj;           /* nb*np                                              */
    double* res_n;        /* nb*np*(nc+1)                                       */
    double* resPc;        /* nb                                                 */
} ocpg_bulk;

/* Connection list in BulkConn order (src/BulkConn.cpp:31-117): bId < eId,
 * sorted by bId then eId; area is the harmonic TPFA transmissibility
 * 1/(1/T1+1/T2) before the CONV1*CONV2 factor.                              */
typedef struct ocpg_conn {
    int64_t nconn;
    const int32_t* bId;
    const int32_t* eId;
    const double* area;
} ocpg_conn;

/* Residual norms of IsoT_FIM::CalRes (src/OCPFluidMethod.cpp:1362-1385). */
typedef struct ocpg_resnorm {
    double maxRelRes_V;
    double maxRelRes_N;
    int64_t maxId_V;
    int64_t maxId_N;
    double maxWellRelRes_mol;   /* over open rate-controlled wells (:1318-1352) */
} ocpg_resnorm;

/* Wells as IsoT_FIM reads them (include/Well.hpp, include/WellOpt.hpp,
 * include/WellPerf.hpp).  Per-step inputs the reference computes once per
 * time step in AllWells::PrepareWell (src/AllWells.cpp:201-214) come in:
 * the perforation pressure offsets dG (Well::CaldG), the injected fluid's
 * molar density at each perforation (MixtureComp::XiPhase in Well::CalFlux
 * with ReCalXi) and the control mode after Well::CheckOptMode.  Arrays are
 * well-major; perforations of well w are perfStart[w] .. perfStart[w+1]-1
 * in the reference's perforation order.                                      */
enum { OCPG_WELL_INJ = 0, OCPG_WELL_PROD = 1 };
enum { OCPG_WELL_BHP = 0, OCPG_WELL_RATE = 1 };  /* RATE: RATE/ORATE/GRATE/WRATE/LRATE */
typedef struct ocpg_wells {
    int32_t nw;                      /* wells                                  */
    const int32_t* type;             /* nw: OCPG_WELL_INJ / OCPG_WELL_PROD     */
    const int32_t* open;             /* nw: nonzero when open (Well::IsOpen)   */
    const int32_t* mode;             /* nw: OCPG_WELL_BHP / OCPG_WELL_RATE     */
    const int32_t* unweight;         /* nw: Well::ifUseUnweight (injectors)    */
    const double* bhp;               /* nw: Well::bhp, the well unknown        */
    const double* ctrlBHP;           /* nw: maxBHP (injector) / minBHP (producer) */
    const double* ctrlRate;          /* nw: WellOpt::maxRate                   */
    const double* injZi;             /* nw*nc: WellOpt::injZi (injectors)      */
    const double* prodPhaseWeight;   /* nw*3: WellOpt::prodPhaseWeight         */
    const int32_t* perfStart;        /* nw+1                                   */
    const int32_t* perfCell;         /* perforated cell (WellPerf::location)   */
    const double* perfWI;            /* WellPerf::WI                           */
    const double* perfMult;          /* WellPerf::multiplier                   */
    const double* perfDG;            /* Well::dG[p]                            */
    const double* perfXi;            /* WellPerf::xi (injectors)               */
    double Psurf, Tsurf;             /* Well::Psurf, Well::Tsurf (F)           */
} ocpg_wells;

/* Totals of MixtureComp's iteration counters (include/MixtureComp.hpp:133-146)
 * over the cells of one flash call.                                          */
typedef struct ocpg_flash_stats {
    int64_t cells;
    int64_t ftype_count[3];
    int64_t two_phase;
    int64_t iters[5];     /* SSMSTA, NRSTA, SSMSP, NRSP, RR                     */
    int64_t nonconverged; /* cells whose ePEC would read 1e8                    */
} ocpg_flash_stats;

/* ------------------------------------------------------------------------ */
/* Entry points.                                                              */
/* ------------------------------------------------------------------------ */

typedef struct ocpg_ctx ocpg_ctx;   /* one device, one stream, one host thread */

const char* ocpg_version(void);
/* sizeof() of a structure above by name ("ocpg_bulk", ...), -1 if unknown.  */
int64_t ocpg_sizeof(const char* name);

/* MixtureComp::MixtureComp (src/MixtureComp.cpp:28-165): deck values to the
 * EoS description the flash uses.  Host only; needs no device.              */
int ocpg_eos_from_deck(const ocpg_eos_deck* deck, ocpg_eos* eos);

/* Context: stands in for the per-region MixtureComp + FlowUnit + Rock
 * objects (Bulk::flashCal/flow/rock) on one GPU.  sat and rock may be NULL
 * when only the flash is used.                                               */
int ocpg_create(ocpg_ctx** ctx, const ocpg_eos* eos, const ocpg_satfunc* sat,
                const ocpg_rock* rock, int device);
int ocpg_destroy(ocpg_ctx* ctx);

/* The FIM method of the context, as IsothermalSolver::SetupMethod chooses it
 * (src/IsothermalSolver.cpp:15-34): IsoT_FIM (secondary variables S and
 * x_ij, the default) or IsoT_FIMn (src/OCPFluidMethod.cpp:2962-4017:
 * secondary variables S and the phase moles n_ij).  Call before
 * ocpg_set_cells (OCPG_ESTATE after).  With IsoT_FIMn the flash, the
 * assembly, the right-hand side, GetSolution and the last-step copy are
 * IsoT_FIMn's (each entry point below says what changes); ocpg_bulk's nj,
 * res_n and resPc carry its extra state.  The serving pool runs IsoT_FIM
 * only (ocpg_pool_init returns OCPG_EUNSUP in an IsoT_FIMn context).        */
#define OCPG_METHOD_FIM 0
#define OCPG_METHOD_FIMN 1
int ocpg_set_method(ocpg_ctx* ctx, int method);
int ocpg_get_method(ocpg_ctx* ctx);
/* Bulk::satcm (src/Bulk.cpp:350-373, FlowUnit_ODGW01's Scm,
 * src/FlowUnit.cpp:196-202): each phase's critical saturation, from the
 * context's tables; IsoT_FIMn::GetSolution's chop reads them.              */
int ocpg_critical_saturations(ocpg_ctx* ctx, double* scm /* 3 */);

/* Device storage for nb cells (IsoT_FIM::AllocateReservoir,
 * src/OCPFluidMethod.cpp:980-1084; IsoT_FIMn::AllocateReservoir, :3111-3127,
 * adds nj, res_n, resPc and their last-step copies), zero-filled.          */
int ocpg_set_cells(ocpg_ctx* ctx, int64_t nb);
/* Host <-> device copies of every non-NULL ocpg_bulk field (reference layout).
 * lNi, when given, is the last-step copy that CalRes reads.                  */
int ocpg_upload_bulk(ocpg_ctx* ctx, const ocpg_bulk* bulk);
int ocpg_download_bulk(ocpg_ctx* ctx, ocpg_bulk* bulk);

/* IsoT_FIM::CalFlash (src/OCPFluidMethod.cpp:1108-1120): MixtureComp::
 * FlashFIM (src/MixtureComp.cpp:494-573) + PassFlashValue (:1122-1187) for
 * every cell, with the skip-stability memory of SkipStaAnaly.  stats may be
 * NULL (then the call is asynchronous).  IsoT_FIMn: IsoT_FIMn::CalFlash
 * (:3152-3177), MixtureComp::FlashFIMn (src/MixtureComp.cpp:575-695) and
 * its PassFlashValue (:3179-3196: also nj, res_n, resPc).                   */
int ocpg_flash_fim(ocpg_ctx* ctx, ocpg_flash_stats* stats);
/* IsoT_FIM::InitFlash (src/OCPFluidMethod.cpp:1095-1106): MixtureComp::
 * InitFlashFIM (src/MixtureComp.cpp:261-353) from P, T, Sw (column 2 of S),
 * rockVp and the feed zi (host array, nb x NC row-major); Ni is an output.
 * IsoT_FIMn: IsoT_FIMn::InitFlash (:3129-3150, InitFlashFIMn :355-442).     */
int ocpg_init_flash_fim(ocpg_ctx* ctx, const double* zi);

/* IsoT_FIM::UpdateLastTimeStep / ResetToLastTimeStep, Bulk and skip-state
 * part (src/OCPFluidMethod.cpp:2859-2957, src/AcceleratePVT.cpp:122-138).
 * IsoT_FIMn's (:3999-4015) also keep nj, res_n and resPc.                  */
int ocpg_update_last_step(ocpg_ctx* ctx);
int ocpg_reset_to_last_step(ocpg_ctx* ctx);

/* BulkConn::iteratorConn (src/BulkConn.cpp:31-117) and the BSR pattern of
 * LinearSystem (include/LinearSystem.hpp:62-125) it implies.                */
int ocpg_set_conn(ocpg_ctx* ctx, const ocpg_conn* conn);
int64_t ocpg_bsr_nnzb(ocpg_ctx* ctx);

/* IsoT_FIM::CalKrPc (src/OCPFluidMethod.cpp:1189-1199) and
 * IsothermalMethod::CalRock (src/OCPFluidMethod.cpp:25-32).                 */
int ocpg_calkrpc(ocpg_ctx* ctx);
int ocpg_calrock(ocpg_ctx* ctx);
int ocpg_calkrpc_rock(ocpg_ctx* ctx);

/* IsoT_FIM::CalRes (src/OCPFluidMethod.cpp:1201-1389), preceded by
 * AllWells::CalTrans and AllWells::CalFlux (src/AllWells.cpp:216-236) when
 * wells are set: bulk fluxes, well sources and the well equations; the
 * negated residual stays on the device as the right-hand side.  norms may be
 * NULL (then the call is asynchronous).                                      */
int ocpg_calres(ocpg_ctx* ctx, double dt, ocpg_resnorm* norms);
/* IsoT_FIM::AssembleMat (src/OCPFluidMethod.cpp:860-875): AssembleMatBulksNew
 * (:1569-1839) and AssembleMatWellsNew (:2417-2733) into the device BSR.
 * Reads the upwinding and well rates of the preceding ocpg_calres.
 * IsoT_FIMn: IsoT_FIMn::AssembleMat (:2998-3006): its AssembleMatBulksNew
 * (:3198-3501) and AssembleMatWellsNew (:3503-3845), whose right-hand side
 * (what ocpg_get_bsr returns as rhs) is -resPc and the -dFdXs res_n terms
 * of the cells with two hydrocarbon phases, plus the residual
 * (AssembleRhsAccumulate).                                                  */
int ocpg_assemble(ocpg_ctx* ctx, double dt);
/* IsoT_FIM::UpdateProperty (CalFlash, CalKrPc, CalRock, the wells'
 * CalTrans/CalFlux, CalRes; src/OCPFluidMethod.cpp:877-897) followed by
 * IsoT_FIM::AssembleMat (:860-875): the results of ocpg_flash_fim,
 * ocpg_calkrpc_rock, ocpg_calres and ocpg_assemble in one call, which lets
 * the flash's slowest cells (long stability tests, split restarts) run on
 * side streams while the other cells' properties, residual and Jacobian are
 * formed, and then recomputes the rows that read the slow cells.  Identical
 * results to the four calls.  norms and stats may be NULL (then the call is
 * asynchronous).  Needs ocpg_set_conn and the kr/Pc tables.  In an
 * IsoT_FIMn context the same with IsoT_FIMn's flash, assembly and
 * right-hand side.                                                          */
int ocpg_update_assemble(ocpg_ctx* ctx, double dt, ocpg_resnorm* norms, ocpg_flash_stats* stats);
/* IsoT_FIM::GetSolution (src/OCPFluidMethod.cpp:2735-2857) on the resident
 * state: the Newton update of P, Ni, S and x_ij with the saturation chop
 * (dSmaxlim: OCPControl's NRdSmax, src/OCPControl.cpp:37) and of the open
 * wells' BHP, from the linear system's solution u (rows x ncol, bulk rows
 * then open wells; host memory, or device memory with u_on_device = 1).
 * out (may be NULL: then asynchronous) gets Bulk::NRdPmax / NRdNmax.       */
typedef struct {
    double NRdPmax;  /* the first cell's dP of largest magnitude             */
    double NRdNmax;  /* the first dNi / Nt of largest magnitude              */
} ocpg_nr_update;
/* IsoT_FIMn: IsoT_FIMn::GetSolution (src/OCPFluidMethod.cpp:3847-3990):
 * P, Ni, the phase moles nj and x_ij (from res_n + dSec_dPri u, with the
 * chop against dSmaxlim and the critical saturations), the BHP; S waits for
 * the next flash.                                                           */
int ocpg_get_solution(ocpg_ctx* ctx, const double* u, int u_on_device, double dSmaxlim, ocpg_nr_update* out);
/* After the flash that follows ocpg_get_solution: IsoT_FIM::PassFlashValue's
 * dSNR / maxNRdSSP (src/OCPFluidMethod.cpp:1135-1146) and Bulk::CalNRdSmax
 * (src/Bulk.cpp:1235-1247), which IsoT_FIM::FinishNR (:937-960) reads.      */
typedef struct {
    double NRdSmax;          /* signed saturation change of largest magnitude */
    int64_t NRdSmax_cell;
    double maxNRdSSP;        /* largest flash-vs-Newton saturation difference */
    int64_t maxNRdSSP_cell;
} ocpg_nr_change_t;
int ocpg_nr_change(ocpg_ctx* ctx, ocpg_nr_change_t* out);
/* Well::BHP of every well as held on the device (nw doubles).              */
int ocpg_get_well_bhp(ocpg_ctx* ctx, double* bhp);
/* The assembled system as VectorFaspSolver::AssembleMat copies it into a
 * dBSRmat (src/FaspSolver.cpp:362-429): IA, JA, val (row-major blocks), rhs.
 * Any pointer may be NULL.                                                    */
int ocpg_get_bsr(ocpg_ctx* ctx, int32_t* IA, int32_t* JA, double* val, double* rhs);
/* The same copies queued on the context's stream without waiting (the
 * caller waits with ocpg_synchronize before the solver reads them), into
 * page-locked arrays from ocpg_host_alloc for the full link rate: the
 * hand-off of the reference's LinearSystem -> VectorFaspSolver::AssembleMat
 * (src/FaspSolver.cpp:362-429) straight into the dBSRmat's arrays.         */
int ocpg_get_bsr_async(ocpg_ctx* ctx, int32_t* IA, int32_t* JA, double* val, double* rhs);
void* ocpg_host_alloc(int64_t bytes);
int ocpg_host_free(void* p);
/* Wells: AllWells as IsoT_FIM::UpdateProperty, CalRes and
 * AssembleMatWellsNew read it (src/OCPFluidMethod.cpp:914-935, 1284-1352,
 * 2417-2733).  ocpg_set_wells fixes the wells and perforations and rebuilds
 * the BSR pattern: after each bulk row's neighbours come its well columns
 * (open wells in order, perforations in order), and one row per open well
 * follows the bulk rows ([well diagonal, perforated cells in perforation
 * order]), as LinearSystem::AddDim/NewDiag/NewOffDiag build them.  Needs
 * ocpg_set_conn first.  nw == 0 removes the wells.                          */
int ocpg_set_wells(ocpg_ctx* ctx, const ocpg_wells* wells);
/* New per-step / per-iteration well values (bhp, mode, ctrlBHP, ctrlRate,
 * perfDG, perfXi) for the wells ocpg_set_wells fixed; other fields ignored. */
int ocpg_set_well_state(ocpg_ctx* ctx, const ocpg_wells* wells);
/* Well::CalFlux and Well::CalProdWeight results of the last ocpg_calres:
 * perforation molar rates perfQi (nperf*nc), well totals wellQi (nw*nc) and
 * producer weights prodWeight (nw).  Any pointer may be NULL.               */
int ocpg_get_well_flux(ocpg_ctx* ctx, double* perfQi, double* wellQi, double* prodWeight);
/* Rows of the assembled system: bulk cells plus open wells.                  */
int64_t ocpg_bsr_rows(ocpg_ctx* ctx);

/* What the wells' per-step preparation reads beyond ocpg_wells (Well::depth,
 * WellOpt::injTemp / injProdPhase / initOptMode; src/Well.cpp:41-92,
 * include/WellOpt.hpp:77-113).  nw entries each; the perforations' depths are
 * their cells' depths (Well::Setup, src/Well.cpp:69).  Set after
 * ocpg_set_wells (which clears it).                                          */
typedef struct ocpg_well_setup {
    const double* depth;      /* Well::depth, the BHP reference depth (ft)    */
    const double* injTemp;    /* injected fluid temperature (F), injectors     */
    const int32_t* injPhase;  /* injectors' phase: 1 gas (solvent), 2 water    */
    const int32_t* initMode;  /* OCPG_WELL_BHP / OCPG_WELL_RATE as scheduled   */
} ocpg_well_setup;
int ocpg_set_well_setup(ocpg_ctx* ctx, const ocpg_well_setup* setup);
/* AllWells::PrepareWell (src/AllWells.cpp:201-214) on the device, once per
 * time step: for each open well Well::CalTrans, Well::CaldG (CalInjdG /
 * CalProddG01, src/Well.cpp:357-555: the perforations' hydrostatic offsets,
 * integrated in 10 ft segments with the injected fluid's density or the
 * flashed perforation stream's), Well::CheckOptMode (:893-944: BHP or rate
 * control from the rate at the BHP limit) and Well::CalFlux's ReCalXi branch
 * (:204-228: the injected fluid's molar density at each perforation).
 * Updates the wells' mode, bhp, dG and perforation densities on the device
 * (ocpg_get_well_state reads them).  Needs ocpg_set_well_setup.             */
int ocpg_prepare_wells(ocpg_ctx* ctx);
/* AllWells::CheckP (src/AllWells.cpp:335-369; Well::CheckP and
 * Well::CheckCrossFlow, src/Well.cpp:946-1066) on the state of the last
 * ocpg_calres: *status = OCPG_WELL_SUCCESS, OCPG_WELL_NEGATIVE_PRESSURE
 * (a BHP or open perforation pressure below zero) or OCPG_WELL_SWITCH (a
 * rate control switched to BHP, or a perforation closed by cross flow; the
 * reference's WELL_SWITCH_TO_BHPMODE / WELL_CROSSFLOW).  Switches and
 * closures are applied to the device state as the reference applies them. */
#define OCPG_WELL_SUCCESS 0
#define OCPG_WELL_NEGATIVE_PRESSURE 1
#define OCPG_WELL_SWITCH 2
int ocpg_check_wells(ocpg_ctx* ctx, int32_t* status);
/* The per-step well state held on the device: control mode (nw), bhp (nw),
 * perforation dG (nperf), injected fluid molar density (nperf) and
 * multiplier (nperf; 0 for a perforation closed by cross flow).  Any pointer
 * may be NULL.                                                              */
int ocpg_get_well_state(ocpg_ctx* ctx, int32_t* mode, double* bhp, double* perfDG, double* perfXi,
                        double* perfMult);
/* MixtureComp::XiPhase and MixtureComp::RhoPhase (src/MixtureComp.cpp:
 * 711-768): molar (lbmol/ft3) and mass (lb/ft3) density of the single phase
 * of feed zi (nc host values) at P (psia), T (F); phase 2 = water (PVTW),
 * otherwise hydrocarbon (one EoS root).  SetupWellOpt's injection factor is
 * 1000 * xi at surface conditions (src/MixtureComp.cpp:790-795).  Blocking. */
int ocpg_phase_density(ocpg_ctx* ctx, double P, double T, const double* zi, int phase, double* xi, double* rho);

/* Bulk::CheckNi then Bulk::CheckP (src/Bulk.cpp:1250-1308), as
 * OCPControl::Check(rs, {"BulkNi", "BulkP"}) runs them in
 * IsoT_FIM::UpdateProperty (src/OCPFluidMethod.cpp:877-885): *status =
 * 0, OCPG_BULK_NEGATIVE_COMPONENTS_MOLES or OCPG_BULK_NEGATIVE_PRESSURE
 * (the reference's BULK_* values, include/Bulk.hpp:39-44).                  */
#define OCPG_BULK_NEGATIVE_PRESSURE (-1)
#define OCPG_BULK_NEGATIVE_COMPONENTS_MOLES (-3)
int ocpg_check_bulk(ocpg_ctx* ctx, int32_t* status);
/* Bulk::CalMaxChange (src/Bulk.cpp:1335-1387) against the last time step
 * and AllWells::CalMaxBHPChange (src/AllWells.cpp:398-406), which
 * OCPControl::CalNextTimeStep reads (src/OCPControl.cpp:292-336).           */
typedef struct {
    double dPmax;    /* max |P - lP|                                         */
    double dNmax;    /* max |Ni - lNi| / |max(Ni, lNi)| over |max| > 1e-8    */
    double dSmax;    /* max |S - lS|                                         */
    double eVmax;    /* max |vf - rockVp| / rockVp                           */
    double dBHPmax;  /* max |bhp - lbhp| over open wells                     */
} ocpg_max_change_t;
int ocpg_max_change(ocpg_ctx* ctx, ocpg_max_change_t* out);

/* The learned form (no reference counterpart).  FlashFIM as ocpg_flash_fim,
 * with a learned initial guess in device memory: nb rows of ld floats
 * (ln K_1 .. ln K_NC, stability logit).  A cell whose logit is positive first
 * tries the guessed K in the stability test; the split starts from it.  The
 * FP64 stability test and split still decide every cell.                   */
int ocpg_flash_fim_guided(ocpg_ctx* ctx, const float* guess, int ld, ocpg_flash_stats* stats);
/* The learned form in the product path: the network (host FP32 weights,
 * nin = NC + 2 inputs, hidden width 64, nout = NC + 1 outputs) is copied to
 * the context; ocpg_flash_fim_learned runs it on every cell's P, T and
 * moles in one fused kernel and then ocpg_flash_fim_guided with its output.
 * All-NULL weights remove the network.  No reference counterpart.         */
int ocpg_set_learned(ocpg_ctx* ctx, int nin, int nout, const float* W1, const float* b1, const float* W2,
                     const float* b2, const float* W3, const float* b3, const float* W4, const float* b4);
int ocpg_flash_fim_learned(ocpg_ctx* ctx, ocpg_flash_stats* stats);
/* Building blocks of the network's training and inference on MI355X (BF16
 * MFMA products, FP32 accumulation): C = A B (+ C) with strided FP32 A, B
 * (rounded to BF16), the K loop split in splitK parts summed in order
 * (partial: splitK*M*N floats when splitK > 1); bias + GELU; GELU';
 * column sums; the loss (masked ln K MSE + weighted stability BCE, dY / B);
 * Adam.  stream is a hipStream_t.                                           */
int ocpg_mlp_gemm(void* stream, int M, int N, int K, const float* A, int64_t sAm, int64_t sAk, const float* B,
                  int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn, int accumulate, int splitK,
                  float* partial);
/* ocpg_mlp_gemm without a workspace argument and without allocating: the
 * split-K partial tiles live in a static buffer of the library, one per GPU,
 * of OCPG_MLP_SCRATCH_FLOATS floats (16 MiB); a call needing more returns
 * OCPG_ETOOBIG.  No reference counterpart: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid allocation).  Calls on one GPU share that buffer, so calls that
 * overlap on one GPU (other streams, other host threads) race; calls
 * ordered on one stream do not.                                            */
#define OCPG_MLP_SCRATCH_FLOATS (1 << 22)
int ocpg_mlp_gemm_noalloc(void* stream, int M, int N, int K, const float* A, int64_t sAm, int64_t sAk,
                          const float* B, int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn,
                          int accumulate, int splitK);
/* The network's whole inference (4 layers of hidden width 64, bias, GELU)
 * in one kernel: Y = MLP(X), X (B x nin, ldX), Y (B x nout, ldY), W_l
 * (in x out row-major FP32, rounded to BF16 in LDS), biases FP32; nin <= 32,
 * nout <= 16.  The same roundings as ocpg_mlp_gemm + ocpg_mlp_bias_act layer
 * by layer.  No reference counterpart (the learned form's own).           */
int ocpg_mlp_forward(void* stream, int B, int nin, int nout, const float* X, int ldX, const float* W1,
                     const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                     const float* W4, const float* b4, float* Y, int ldY);
int ocpg_mlp_bias_act(void* stream, int M, int N, float* Z, const float* b, float* H);
int ocpg_mlp_act_backward(void* stream, int64_t n, const float* Z, float* dH);
int ocpg_mlp_colsum(void* stream, int M, int N, const float* X, float* out);
int ocpg_mlp_loss(void* stream, int B, int nK, int ldY, const float* Y, const float* lnK, const float* label,
                  const float* mask, float wStab, float* dY, float* loss);
int ocpg_mlp_adam(void* stream, int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1,
                  float b2, float eps, int step);

/* One Newton iteration of IsoT_FIM on the resident state (the inputs saved by
 * ocpg_save_flash_inputs restored, then CalFlash, CalKrPc, CalRock, the
 * wells' CalTrans/CalFlux, CalRes and AssembleMat: src/OCPFluidMethod.cpp:
 * 860-935) captured into a HIP graph, for callers that replay iterations
 * many times (the serving loop).  No reference counterpart: IsoT_FIM issues
 * the same calls one by one.  Timing must be off while capturing.            */
int ocpg_capture_iteration(ocpg_ctx* ctx, double dt);
/* Replays it on the context's stream (asynchronous).                        */
int ocpg_launch_iteration(ocpg_ctx* ctx);
/* 1 when the last replay has finished, 0 while it runs, < 0 on error.       */
int ocpg_iteration_done(ocpg_ctx* ctx);
/* BulkConn::upblock / upblock_Rho / upblock_Velocity as CalRes set them.    */
int ocpg_get_upwind(ocpg_ctx* ctx, int32_t* upblock, double* upblock_Rho,
                    double* upblock_Velocity);

/* Serving pool: many realisations (ensemble members of their own sizes) in
 * one context, each at a reserved range of cells, connection slots and BSR
 * blocks, and a per-cell activity mask; the captured iteration then works
 * on the realisations whose mask is on, in one set of launches.  A pool's
 * realisations carry no wells.  No reference counterpart: the reference runs
 * one reservoir per process.
 *   ocpg_pool_init: capacity cells of at most max_neighbours neighbours each
 *     (reallocates like ocpg_set_cells; drops a captured iteration);
 *   ocpg_pool_place: a realisation's cell state (the layout of
 *     ocpg_upload_bulk, state->nb cells) and connections (local indices, the
 *     rules of ocpg_set_conn) at cells [base, base + nb); its iteration
 *     inputs are saved at the same time; the range must be free;
 *   ocpg_pool_activate: whether it takes part in the iterations launched
 *     after the call on the context's stream (stream-ordered);
 *   ocpg_pool_system: its rows of the last assembly in its own indexing (what
 *     ocpg_get_bsr returns for it alone); NULL outputs are skipped;
 *   ocpg_pool_download: its cell state, the saved iteration inputs
 *     (saved != 0) or the current values;
 *   ocpg_pool_remove: frees its range;
 *   ocpg_pool_image_bytes / ocpg_pool_export / ocpg_pool_import: the
 *     realisation as one host block in the device's layout (state, fixed
 *     fields and pattern); export with all == 0 refreshes only the state
 *     (saved iteration inputs, last-step moles) of an image it filled
 *     before; import places it at any free base.  Both run on the pool's
 *     copy stream, overlapping the iterations (page-locked images copy
 *     without staging); ocpg_pool_sync_copies waits for them.  Export and
 *     remove need the realisation switched off.
 * The pool keeps no last-step copy beyond the moles CalRes reads
 * (ocpg_update_last_step / ocpg_reset_to_last_step return OCPG_ESTATE).
 * Placing, importing and removing keep a captured iteration valid.           */
int ocpg_pool_init(ocpg_ctx* ctx, int64_t capacity, int32_t max_neighbours);
int ocpg_pool_place(ocpg_ctx* ctx, int64_t base, const ocpg_bulk* state, const ocpg_conn* conn);
int ocpg_pool_activate(ocpg_ctx* ctx, int64_t base, int on);
int ocpg_pool_system(ocpg_ctx* ctx, int64_t base, int64_t* nnzb, int32_t* IA, int32_t* JA, double* val,
                     double* rhs);
int ocpg_pool_download(ocpg_ctx* ctx, int64_t base, ocpg_bulk* state, int saved);
int ocpg_pool_remove(ocpg_ctx* ctx, int64_t base);
int64_t ocpg_pool_image_bytes(ocpg_ctx* ctx, int64_t base);
int ocpg_pool_export(ocpg_ctx* ctx, int64_t base, void* image, int all);
int ocpg_pool_import(ocpg_ctx* ctx, int64_t base, const void* image);
int ocpg_pool_sync_copies(ocpg_ctx* ctx);

/* Copy / restore of the fields the flash starts from (P, Ni, S, phaseNum,
 * xij, skip-stability memory), to re-run one Newton iteration from the same
 * state (benchmarks, retries).  No reference counterpart.                    */
int ocpg_save_flash_inputs(ocpg_ctx* ctx);
int ocpg_restore_flash_inputs(ocpg_ctx* ctx);

/* Per-kernel HIP-event timers on the context's stream.  Names: "flash",
 * "krpc_rock", "calres", "assemble", "init_flash", "restore", the flash's
 * stages "flash_stability", "flash_split", "flash_finish", and for
 * ocpg_update_assemble "flash_tail" (side stream C), "flash_restarts" (side
 * stream D), "fixup" and "update_assemble" (the whole call); "pack" (the
 * per-cell records the assembly reads; "assemble" is the assembly kernel
 * alone, with the wells' blocks).  No reference
 * counterpart (the reference times whole phases with GetWallTime,
 * include/UtilTiming.hpp:31-62).                                             */
int ocpg_timing_enable(ocpg_ctx* ctx, int on);
int ocpg_timing_get(ocpg_ctx* ctx, const char* name, double* ms, int64_t* count);

int ocpg_synchronize(ocpg_ctx* ctx);
/* The context's HIP stream (as void*), for callers that order their own work
 * after the library's.                                                        */
void* ocpg_stream(ocpg_ctx* ctx);
/* Device bytes held by the context.                                          */
int64_t ocpg_device_bytes(ocpg_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* OCPGPU_H */
