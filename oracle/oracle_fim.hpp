// TEST INFRASTRUCTURE ONLY: CPU restatement of IsoT_FIM's property pass and
// assembly (oracle + CPU baseline).  Not part of the product.
#ifndef ORACLE_FIM_HPP
#define ORACLE_FIM_HPP

#include <vector>

#include "oracle_flash.hpp"

namespace oracle {

// Bulk::maxLendSdP, src/OCPFluidMethod.cpp:1056.
inline int maxLendSdP(int nc) { return (nc + 1) * (nc + 1) * NPH; }

void CalFlash(const ocpg_eos& eos, const ocpg_bulk& bk, int nthreads);
void InitFlash(const ocpg_eos& eos, const ocpg_bulk& bk, const double* zi, int nthreads);
void CalKrPc(const ocpg_satfunc& sat, const ocpg_bulk& bk, int nthreads);
void CalRock(const ocpg_rock& rock, const ocpg_bulk& bk, int nthreads);
// Well::CalTrans / Well::CalFlux / Well::CalProdWeight results, well-major.
struct WellFlux {
    std::vector<double> perfTransj;  // nperf * 3
    std::vector<double> perfQi;      // nperf * nc
    std::vector<double> wellQi;      // nw * nc
    std::vector<double> prodWeight;  // nw (rate-controlled producers)
};
int OpenWells(const ocpg_wells* w);
void CalWellFlux(const ocpg_eos& eos, const ocpg_wells& w, const ocpg_bulk& bk, WellFlux& out);
// Per-step well state the preparation and checks change, besides the
// fields of ocpg_wells they update in place (mode, bhp, perfDG, perfXi,
// perfMult).
struct WellStep {
    double* perfTransj;  // nperf * 3
    double* perfP;       // nperf: WellPerf::P
    int32_t* perfState;  // nperf: 1 open, 0 closed
    const double* lbhp;  // nw, for what = 1
};
// AllWells::PrepareWell (what = 0) or the well part of ResetToLastTimeStep
// (what = 1: ResetBHP, CalTrans, CaldG, the perforation pressures).
void PrepareWells(const ocpg_eos& eos, ocpg_wells& w, const ocpg_well_setup* setup, const ocpg_bulk& bk,
                  WellStep& st, int what);
// AllWells::CheckP: 0 success, 1 negative pressure, 2 switch / cross flow.
int CheckWells(ocpg_wells& w, const ocpg_bulk& bk, WellStep& st);
// With wells, resAbs holds (nb + open wells) * (nc + 1) entries.
// Rows are computed in parallel (nthreads OpenMP threads), each summing its
// connections' terms in the reference's connection order: the result does
// not depend on nthreads.
void CalRes(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt, double* resAbs,
            int32_t* upblock, double* upblock_Rho, double* upblock_Velocity, ocpg_resnorm* norms,
            const ocpg_wells* wells = nullptr, const WellFlux* wf = nullptr, int nthreads = 1);
// Each cell's connections in ascending index order; posB / posE: a
// connection's rank in its bId / eId cell's list.
struct CellConn {
    std::vector<int64_t> ptr, conn;
    std::vector<int32_t> posB, posE;
};
CellConn CellConnections(int64_t nb, const ocpg_conn& conn);
void BsrPattern(int64_t nb, const ocpg_conn& conn, int32_t* IA, int32_t* JA,
                const ocpg_wells* wells = nullptr);
// With fimn, IsoT_FIMn::AssembleMatWellsNew (src/OCPFluidMethod.cpp:
// 3503-3845); rhs (nb * ncol, holding the bulk part) receives each
// perforation's -dQdXs res_n where its cell has two hydrocarbon phases.
void AssembleWells(int nc, const ocpg_conn& conn, const ocpg_wells& w, const WellFlux& wf,
                   const ocpg_bulk& bk, double dt, const int32_t* IA, double* val, bool fimn = false,
                   double* rhs = nullptr);
void AssembleBulks(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt,
                   const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                   const int32_t* JA, double* val, int nthreads = 1, bool fimn = false,
                   double* rhs = nullptr);

// IsoT_FIM::GetSolution (bulk part) and the Newton-change bookkeeping of
// the next PassFlashValue / Bulk::CalNRdSmax.
void GetSolution(int nc, const ocpg_bulk& bk, const double* u, double dSmaxlim, double* dSNR, double* dSNRP,
                 uint8_t* phaseExistOld, double* NRdPmax, double* NRdNmax);
// IsoT_FIMn::GetSolution, bulk part (src/OCPFluidMethod.cpp:3847-3990);
// scm: the critical saturation of each phase (Bulk::satcm).
void GetSolutionN(int nc, const ocpg_bulk& bk, const double* u, double dSmaxlim, const double* scm, double* dSNR,
                  double* dSNRP, uint8_t* phaseExistOld, double* NRdPmax, double* NRdNmax);
// IsoT_FIMn::CalFlash / InitFlash with IsoT_FIMn::PassFlashValue
// (src/OCPFluidMethod.cpp:3129-3196).
void CalFlashN(const ocpg_eos& eos, const ocpg_bulk& bk, int nthreads);
void InitFlashN(const ocpg_eos& eos, const ocpg_bulk& bk, const double* zi, int nthreads);
void NRChange(int64_t nb, const double* S, double* dSNR, const double* dSNRP, const uint8_t* phaseExistOld,
              double* out);

}  // namespace oracle

#endif
