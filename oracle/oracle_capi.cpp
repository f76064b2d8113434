// C entry points of the oracle (loaded by the tests and the benchmark's
// CPU-baseline leg through ctypes).  TEST INFRASTRUCTURE ONLY.
#include <chrono>
#include <cstring>

#include "oracle_fim.hpp"

namespace oracle {
int InitPTSw(const ocpg_eos& eos, const ocpg_satfunc& sat, int64_t nb, const double* depth, const double* dz,
             const double* equil, const double* zmfvd, int nzrows, double rtemp, int tabrow, double* P, double* Sw,
             double* zout);
}

extern "C" {

int orc_flash_fim(const ocpg_eos* eos, ocpg_bulk* bk, int nthreads)
{
    if (!eos || !bk || eos->nc < 1 || eos->nc > OCPG_MAXNC) return OCPG_EINVAL;
    oracle::CalFlash(*eos, *bk, nthreads);
    return OCPG_OK;
}

int orc_init_flash_fim(const ocpg_eos* eos, ocpg_bulk* bk, const double* zi, int nthreads)
{
    if (!eos || !bk || !zi || eos->nc < 1 || eos->nc > OCPG_MAXNC) return OCPG_EINVAL;
    oracle::InitFlash(*eos, *bk, zi, nthreads);
    return OCPG_OK;
}

int orc_calkrpc(const ocpg_satfunc* sat, ocpg_bulk* bk, int nthreads)
{
    if (!sat || !bk) return OCPG_EINVAL;
    oracle::CalKrPc(*sat, *bk, nthreads);
    return OCPG_OK;
}

int orc_calrock(const ocpg_rock* rock, ocpg_bulk* bk, int nthreads)
{
    if (!rock || !bk) return OCPG_EINVAL;
    oracle::CalRock(*rock, *bk, nthreads);
    return OCPG_OK;
}

int orc_calres(int nc, const ocpg_conn* conn, ocpg_bulk* bk, double dt, double* resAbs,
               int32_t* upblock, double* upblock_Rho, double* upblock_Velocity, ocpg_resnorm* norms)
{
    if (!conn || !bk || !resAbs || !upblock || !upblock_Rho || !upblock_Velocity) return OCPG_EINVAL;
    oracle::CalRes(nc, *conn, *bk, dt, resAbs, upblock, upblock_Rho, upblock_Velocity, norms);
    return OCPG_OK;
}

// The same with nthreads OpenMP threads (identical results).
int orc_calres_mt(int nc, const ocpg_conn* conn, ocpg_bulk* bk, double dt, double* resAbs, int32_t* upblock,
                  double* upblock_Rho, double* upblock_Velocity, ocpg_resnorm* norms, int nthreads)
{
    if (!conn || !bk || !resAbs || !upblock || !upblock_Rho || !upblock_Velocity) return OCPG_EINVAL;
    oracle::CalRes(nc, *conn, *bk, dt, resAbs, upblock, upblock_Rho, upblock_Velocity, norms, nullptr, nullptr,
                   nthreads);
    return OCPG_OK;
}

// AllWells::PrepareWell (what 0) / ResetToLastTimeStep's well part (what 1)
// on the wells' arrays in place; setup may be NULL (no CaldG / CheckOptMode).
int orc_prepare_wells(const ocpg_eos* eos, ocpg_wells* w, const ocpg_well_setup* setup, ocpg_bulk* bk,
                      double* perfTransj, double* perfP, int32_t* perfState, const double* lbhp, int what)
{
    if (!eos || !w || !bk || !perfTransj || !perfP || !perfState || (what == 1 && !lbhp)) return OCPG_EINVAL;
    oracle::WellStep st{perfTransj, perfP, perfState, lbhp};
    oracle::PrepareWells(*eos, *w, setup, *bk, st, what);
    return OCPG_OK;
}

int orc_check_wells(ocpg_wells* w, ocpg_bulk* bk, double* perfTransj, double* perfP, int32_t* perfState,
                    int32_t* status)
{
    if (!w || !bk || !perfTransj || !perfP || !perfState || !status) return OCPG_EINVAL;
    oracle::WellStep st{perfTransj, perfP, perfState, nullptr};
    *status = oracle::CheckWells(*w, *bk, st);
    return OCPG_OK;
}

int orc_phase_density(const ocpg_eos* eos, double P, double T, const double* zi, int phase, double* out)
{
    if (!eos || !zi || !out) return OCPG_EINVAL;
    oracle::Mixture mix(*eos);
    mix.PhaseDensity(P, T, zi, phase == 2 ? 2 : 1, &out[0], &out[1]);
    return OCPG_OK;
}

// Bulk::InitPTSw of a compositional model (oracle_equil.cpp).
int orc_init_ptsw(const ocpg_eos* eos, const ocpg_satfunc* sat, int64_t nb, const double* depth, const double* dz,
                  const double* equil, const double* zmfvd, int nzrows, double rtemp, int tabrow, double* P,
                  double* Sw, double* z)
{
    if (!eos || !sat || !depth || !dz || !equil || !zmfvd || nzrows < 1 || tabrow < 2 || !P || !Sw || !z)
        return OCPG_EINVAL;
    return oracle::InitPTSw(*eos, *sat, nb, depth, dz, equil, zmfvd, nzrows, rtemp, tabrow, P, Sw, z);
}

// xi, rho, mu of one hydrocarbon phase of composition x at P, T.
int orc_phase_props(const ocpg_eos* eos, double P, double T, const double* x, double* out)
{
    if (!eos || !x || !out) return OCPG_EINVAL;
    oracle::Mixture mix(*eos);
    mix.PhaseProps(P, T, x, out);
    return OCPG_OK;
}

int orc_flash_phases(const ocpg_eos* eos, double P, double T, const double* Ni, double* out)
{
    // out: ex[3], vj[3], xi[3], rho[3], Nt
    if (!eos || !Ni || !out) return OCPG_EINVAL;
    oracle::Mixture mix(*eos);
    bool ex[oracle::NPH];
    out[12] = mix.FlashPhases(P, T, Ni, ex, out + 3, out + 6, out + 9);
    for (int j = 0; j < oracle::NPH; j++) out[j] = ex[j] ? 1.0 : 0.0;
    return OCPG_OK;
}

int orc_bsr_pattern(int64_t nb, const ocpg_conn* conn, int32_t* IA, int32_t* JA)
{
    if (!conn || !IA || !JA) return OCPG_EINVAL;
    oracle::BsrPattern(nb, *conn, IA, JA);
    return OCPG_OK;
}

int orc_assemble_bulks(int nc, const ocpg_conn* conn, ocpg_bulk* bk, double dt,
                       const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                       const int32_t* JA, double* val)
{
    if (!conn || !bk || !upblock || !upblock_Rho || !IA || !JA || !val) return OCPG_EINVAL;
    oracle::AssembleBulks(nc, *conn, *bk, dt, upblock, upblock_Rho, IA, JA, val);
    return OCPG_OK;
}

int orc_assemble_bulks_mt(int nc, const ocpg_conn* conn, ocpg_bulk* bk, double dt, const int32_t* upblock,
                          const double* upblock_Rho, const int32_t* IA, const int32_t* JA, double* val,
                          int nthreads)
{
    if (!conn || !bk || !upblock || !upblock_Rho || !IA || !JA || !val) return OCPG_EINVAL;
    oracle::AssembleBulks(nc, *conn, *bk, dt, upblock, upblock_Rho, IA, JA, val, nthreads);
    return OCPG_OK;
}

// Well::CalTrans + Well::CalFlux + Well::CalProdWeight of every open well.
int orc_well_flux(const ocpg_eos* eos, const ocpg_wells* wells, ocpg_bulk* bk, double* perfTransj,
                  double* perfQi, double* wellQi, double* prodWeight)
{
    if (!eos || !wells || !bk) return OCPG_EINVAL;
    oracle::WellFlux wf;
    oracle::CalWellFlux(*eos, *wells, *bk, wf);
    const int nc = eos->nc + 1;
    const int nperf = wells->nw > 0 ? wells->perfStart[wells->nw] : 0;
    if (perfTransj) std::memcpy(perfTransj, wf.perfTransj.data(), sizeof(double) * nperf * 3);
    if (perfQi) std::memcpy(perfQi, wf.perfQi.data(), sizeof(double) * nperf * nc);
    if (wellQi) std::memcpy(wellQi, wf.wellQi.data(), sizeof(double) * wells->nw * nc);
    if (prodWeight) std::memcpy(prodWeight, wf.prodWeight.data(), sizeof(double) * wells->nw);
    return OCPG_OK;
}

// MixtureComp::CalProdWeight on the oracle's restatement.
int orc_prod_weight(const ocpg_eos* eos, double Psurf, double Tsurf, const double* Ni, const double* prodPhase,
                    double* weight)
{
    if (!eos || !Ni || !prodPhase || !weight) return OCPG_EINVAL;
    oracle::Mixture mix(*eos);
    *weight = mix.ProdWeight(Psurf, Tsurf, Ni, prodPhase);
    return OCPG_OK;
}

// IsoT_FIM::UpdateProperty's well part + CalRes with wells: resAbs holds
// (nb + open wells) * (nc + 1) entries.
int orc_calres_wells(const ocpg_eos* eos, const ocpg_wells* wells, const ocpg_conn* conn, ocpg_bulk* bk,
                     double dt, double* resAbs, int32_t* upblock, double* upblock_Rho,
                     double* upblock_Velocity, ocpg_resnorm* norms)
{
    if (!eos || !wells || !conn || !bk || !resAbs || !upblock || !upblock_Rho || !upblock_Velocity)
        return OCPG_EINVAL;
    oracle::WellFlux wf;
    oracle::CalWellFlux(*eos, *wells, *bk, wf);
    oracle::CalRes(eos->nc + 1, *conn, *bk, dt, resAbs, upblock, upblock_Rho, upblock_Velocity, norms, wells,
                   &wf);
    return OCPG_OK;
}

int orc_bsr_pattern_wells(int64_t nb, const ocpg_conn* conn, const ocpg_wells* wells, int32_t* IA, int32_t* JA)
{
    if (!conn || !wells || !IA || !JA) return OCPG_EINVAL;
    oracle::BsrPattern(nb, *conn, IA, JA, wells);
    return OCPG_OK;
}

// IsoT_FIM::AssembleMat: AssembleMatBulksNew then AssembleMatWellsNew.
int orc_assemble_wells(const ocpg_eos* eos, const ocpg_wells* wells, const ocpg_conn* conn, ocpg_bulk* bk,
                       double dt, const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                       const int32_t* JA, double* val)
{
    if (!eos || !wells || !conn || !bk || !upblock || !upblock_Rho || !IA || !JA || !val) return OCPG_EINVAL;
    const int nc = eos->nc + 1;
    oracle::AssembleBulks(nc, *conn, *bk, dt, upblock, upblock_Rho, IA, JA, val);
    oracle::WellFlux wf;
    oracle::CalWellFlux(*eos, *wells, *bk, wf);
    oracle::AssembleWells(nc, *conn, *wells, wf, *bk, dt, IA, val);
    return OCPG_OK;
}

// IsoT_FIMn: CalFlash / InitFlash (FlashFIMn / InitFlashFIMn + its
// PassFlashValue; bk->nj, res_n, resPc are outputs too).
int orc_flash_fimn(const ocpg_eos* eos, ocpg_bulk* bk, int nthreads)
{
    if (!eos || !bk || !bk->nj || !bk->res_n || !bk->resPc || eos->nc < 1 || eos->nc > OCPG_MAXNC)
        return OCPG_EINVAL;
    oracle::CalFlashN(*eos, *bk, nthreads);
    return OCPG_OK;
}

int orc_init_flash_fimn(const ocpg_eos* eos, ocpg_bulk* bk, const double* zi, int nthreads)
{
    if (!eos || !bk || !zi || !bk->nj || !bk->res_n || !bk->resPc || eos->nc < 1 || eos->nc > OCPG_MAXNC)
        return OCPG_EINVAL;
    oracle::InitFlashN(*eos, *bk, zi, nthreads);
    return OCPG_OK;
}

// The FlashFIMn of every cell, keeping each cell's CaldXsdXp03 system:
// M (nb x NSYS^2, column-major, dim rows), B (nb x NSYS (NC + 3)), dim (nb;
// 0 for one hydrocarbon phase).  Test helper: the conditioning of that solve.
int orc_fimn_systems(const ocpg_eos* eos, ocpg_bulk* bk, double* M, double* B, int32_t* dim)
{
    if (!eos || !bk || !M || !B || !dim) return OCPG_EINVAL;
    const int NC = eos->nc, nc = NC + 1, np = oracle::NPH;
    constexpr int NS = oracle::Mixture::NSYS;
    oracle::Mixture mix(*eos);
    for (int64_t n = 0; n < bk->nb; n++) {
        oracle::SkipState s;
        s.flag = bk->skipFlag[n] != 0;
        s.minEigen = bk->skipMinEig[n];
        s.P = bk->skipP[n];
        s.T = bk->skipT[n];
        for (int i = 0; i < NC; i++) s.zi[i] = bk->skipZi[n * NC + i];
        mix.FlashFIMn(bk->P[n], bk->T[n], &bk->Ni[n * nc], &bk->S[n * np], bk->phaseNum[n], &bk->xij[n * np * nc],
                      s);
        dim[n] = mix.sysDim;
        for (int k = 0; k < NS * NS; k++) M[n * NS * NS + k] = oracle::todbl(mix.sysM[k]);
        for (int k = 0; k < NS * (OCPG_MAXNC + 3); k++) B[n * NS * (OCPG_MAXNC + 3) + k] = oracle::todbl(mix.sysB[k]);
    }
    return OCPG_OK;
}

// IsoT_FIMn::AssembleMat: AssembleMatBulksNew, AssembleMatWellsNew (wells
// may be NULL), then AssembleRhsAccumulate(resAbs): rhs ((nb + open wells)
// * ncol) = the bulk and perforation terms + resAbs.
int orc_assemble_fimn(const ocpg_eos* eos, const ocpg_wells* wells, const ocpg_conn* conn, ocpg_bulk* bk,
                      double dt, const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                      const int32_t* JA, const double* resAbs, double* val, double* rhs)
{
    if (!eos || !conn || !bk || !upblock || !upblock_Rho || !IA || !JA || !resAbs || !val || !rhs)
        return OCPG_EINVAL;
    const int nc = eos->nc + 1;
    const int ncol = nc + 1;
    const int64_t rows = bk->nb + oracle::OpenWells(wells);
    for (int64_t k = 0; k < rows * ncol; k++) rhs[k] = 0;
    oracle::AssembleBulks(nc, *conn, *bk, dt, upblock, upblock_Rho, IA, JA, val, 1, true, rhs);
    if (wells) {
        oracle::WellFlux wf;
        oracle::CalWellFlux(*eos, *wells, *bk, wf);
        oracle::AssembleWells(nc, *conn, *wells, wf, *bk, dt, IA, val, true, rhs);
    }
    for (int64_t k = 0; k < rows * ncol; k++) rhs[k] += resAbs[k];
    return OCPG_OK;
}

// IsoT_FIMn::GetSolution, bulk part; out[2] = {NRdPmax, NRdNmax}.
int orc_get_solution_n(int nc, ocpg_bulk* bk, const double* u, double dSmaxlim, const double* scm, double* dSNR,
                       double* dSNRP, uint8_t* phaseExistOld, double* out)
{
    if (!bk || !u || !scm || !dSNR || !dSNRP || !phaseExistOld || !out || !bk->nj || !bk->res_n)
        return OCPG_EINVAL;
    oracle::GetSolutionN(nc, *bk, u, dSmaxlim, scm, dSNR, dSNRP, phaseExistOld, &out[0], &out[1]);
    return OCPG_OK;
}

// Wall time of one CalFlash pass (seconds), for the CPU baseline.
double orc_time_flash_fim(const ocpg_eos* eos, ocpg_bulk* bk, int nthreads)
{
    auto t0 = std::chrono::steady_clock::now();
    oracle::CalFlash(*eos, *bk, nthreads);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// IsoT_FIM::GetSolution, bulk part; out[2] = {NRdPmax, NRdNmax}.
int orc_get_solution(int nc, ocpg_bulk* bk, const double* u, double dSmaxlim, double* dSNR, double* dSNRP,
                     uint8_t* phaseExistOld, double* out)
{
    if (!bk || !u || !dSNR || !dSNRP || !phaseExistOld || !out) return OCPG_EINVAL;
    oracle::GetSolution(nc, *bk, u, dSmaxlim, dSNR, dSNRP, phaseExistOld, &out[0], &out[1]);
    return OCPG_OK;
}

// The Newton-change bookkeeping after the flash; out[4] = {NRdSmax, its
// cell, maxNRdSSP, its cell}.
int orc_nr_change(int64_t nb, const double* S, double* dSNR, const double* dSNRP, const uint8_t* phaseExistOld,
                  double* out)
{
    if (!S || !dSNR || !dSNRP || !phaseExistOld || !out) return OCPG_EINVAL;
    oracle::NRChange(nb, S, dSNR, dSNRP, phaseExistOld, out);
    return OCPG_OK;
}

#ifdef OCP_COUNT_FLOPS
// Flop tally of one serial CalFlash pass: out[0] = add/sub/mul/div/sqrt,
// out[1] = transcendental calls, out[2..3] the same within PhaseStable,
// out[4..5] within PhaseSplit.
int orc_count_flash_fim(const ocpg_eos* eos, ocpg_bulk* bk, uint64_t* out)
{
    oracle::g_tally = oracle::FlopTally();
    oracle::CalFlash(*eos, *bk, 1);
    out[0] = oracle::g_tally.flops;
    out[1] = oracle::g_tally.trans;
    out[2] = oracle::g_tally.staFlops;
    out[3] = oracle::g_tally.staTrans;
    out[4] = oracle::g_tally.splitFlops;
    out[5] = oracle::g_tally.splitTrans;
    return OCPG_OK;
}
#endif

#ifdef OCP_ULP_NOISE
// One serial CalFlash pass with every transcendental perturbed by up to one
// ulp from the stream `seed` (nonzero).
int orc_noise_flash_fim(const ocpg_eos* eos, ocpg_bulk* bk, uint64_t seed)
{
    if (!eos || !bk || !seed || eos->nc < 1 || eos->nc > OCPG_MAXNC) return OCPG_EINVAL;
    oracle::g_noise = seed;
    oracle::CalFlash(*eos, *bk, 1);
    return OCPG_OK;
}
#endif

}  // extern "C"
