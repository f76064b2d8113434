// Driver that runs the REFERENCE's own MixtureComp (compiled from
// /root/reference/src by oracle/Makefile into oracle/_ref/) on arrays in the
// reference's Bulk layout.  TEST INFRASTRUCTURE ONLY: it produces golden
// vectors and pins the restatement; nothing in the product loads it.
//
// The reference's per-cell loop (IsoT_FIM::CalFlash, src/OCPFluidMethod.cpp:
// 1108-1120) and PassFlashValue (:1122-1187) live in a file that needs FASP
// headers, so this driver restates those two loops around the reference's
// MixtureComp::FlashFIM / InitFlashFIM and its getters.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>
// The skip-stability memory and the iteration counters are private members
// of the reference classes; the driver reads and seeds them directly.
#define protected public
#define private public
#include "MixtureComp.hpp"
#undef protected
#undef private

#include "../include/ocpgpu.h"

namespace {

struct RefHandle {
    ParamReservoir rp;
    MixtureComp* mix = nullptr;
    OptionalFeatures opt;
    int64_t nb = 0;
    int NC = 0;
};

std::string num(double v)
{
    std::ostringstream os;
    os << std::setprecision(17) << v;
    return os.str();
}

struct Quiet {
    std::streambuf* old;
    std::ostringstream sink;
    Quiet() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~Quiet() { std::cout.rdbuf(old); }
};

}  // namespace

extern "C" {

void* ref_create(const ocpg_eos_deck* d, int64_t nb)
{
    Quiet q;
    RefHandle* h = new RefHandle;
    h->nb = nb;
    h->NC = d->nc;
    const int NC = d->nc;
    ComponentParam& cp = h->rp.comsParam;
    cp.numCom = NC;
    cp.numPhase = 2;
    auto put = [&](Type_A_r<vector<OCP_DBL>>& t, const double* v) {
        t.activity = 1;
        t.data.assign(1, vector<OCP_DBL>(v, v + NC));
    };
    put(cp.Tc, d->Tc);
    put(cp.Pc, d->Pc);
    put(cp.MW, d->MW);
    put(cp.Acf, d->Acf);
    if (d->has_Vc) put(cp.Vc, d->Vc);
    if (d->has_Zc) put(cp.Zc, d->Zc);
    if (d->has_OmegaA) put(cp.OmegaA, d->OmegaA);
    if (d->has_OmegaB) put(cp.OmegaB, d->OmegaB);
    if (d->has_Vshift) put(cp.Vshift, d->Vshift);
    if (d->has_Vcvis) put(cp.Vcvis, d->Vcvis);
    if (d->has_Zcvis) put(cp.Zcvis, d->Zcvis);
    cp.BIC.assign(1, vector<OCP_DBL>(d->BIC_lower, d->BIC_lower + NC * (NC - 1) / 2));
    cp.LBCcoef.assign(d->LBCcoef, d->LBCcoef + 5);
    cp.SSMparamSTA = {std::to_string(d->ssmsta_maxit), num(d->ssmsta_tol), num(d->ssmsta_eyt)};
    cp.NRparamSTA = {std::to_string(d->nrsta_maxit), num(d->nrsta_tol)};
    cp.SSMparamSP = {std::to_string(d->ssmsp_maxit), num(d->ssmsp_tol)};
    cp.NRparamSP = {std::to_string(d->nrsp_maxit), num(d->nrsp_tol)};
    cp.RRparam = {std::to_string(d->rr_maxit), num(d->rr_tol)};
    h->rp.PVTW_T.data.resize(1);
    h->rp.PVTW_T.data[0].assign(5, vector<OCP_DBL>());
    for (int r = 0; r < d->pvtw_rows; r++)
        for (int c = 0; c < 5; c++) h->rp.PVTW_T.data[0][c].push_back(d->pvtw[r][c]);
    h->rp.gravity.activity = 0;
    h->rp.density.activity = 0;
    h->mix = new MixtureComp(h->rp, 0);
    h->mix->std_RhoW = d->std_rhow;
    h->mix->SetupOptionalFeatures(h->opt, (OCP_USI)nb);
    return h;
}

void ref_destroy(void* p)
{
    RefHandle* h = static_cast<RefHandle*>(p);
    delete h->mix;
    delete h;
}

// The EoS description as the reference constructor derived it.
int ref_get_eos(void* p, ocpg_eos* e)
{
    RefHandle* h = static_cast<RefHandle*>(p);
    MixtureComp& m = *h->mix;
    const int NC = h->NC;
    e->nc = NC;
    for (int i = 0; i < NC; i++) {
        e->Tc[i] = m.Tc[i];
        e->Pc[i] = m.Pc[i];
        e->Vc[i] = m.Vc[i];
        e->MW[i] = m.MWC[i];
        e->Acf[i] = m.Acf[i];
        e->OmegaA[i] = m.OmegaA[i];
        e->OmegaB[i] = m.OmegaB[i];
        e->Vshift[i] = m.Vshift[i];
        e->Vcvis[i] = m.Vcvis[i];
        for (int k = 0; k < NC; k++) e->BIC[i * OCPG_MAXNC + k] = m.BIC[i * NC + k];
    }
    for (int q = 0; q < 5; q++) e->LBCcoef[q] = m.LBCcoef[q];
    e->ssmsta_maxit = m.EoSctrl.SSMsta.maxIt;
    e->ssmsta_tol = m.EoSctrl.SSMsta.tol;
    e->ssmsta_eyt = m.EoSctrl.SSMsta.eYt;
    e->nrsta_maxit = m.EoSctrl.NRsta.maxIt;
    e->nrsta_tol = m.EoSctrl.NRsta.tol;
    e->ssmsp_maxit = m.EoSctrl.SSMsp.maxIt;
    e->ssmsp_tol = m.EoSctrl.SSMsp.tol;
    e->nrsp_maxit = m.EoSctrl.NRsp.maxIt;
    e->nrsp_tol = m.EoSctrl.NRsp.tol;
    e->rr_maxit = m.EoSctrl.RR.maxIt;
    e->rr_tol = m.EoSctrl.RR.tol;
    const auto& tab = m.PVTW.data;
    e->pvtw_rows = (int32_t)tab[0].size();
    for (int r = 0; r < e->pvtw_rows; r++)
        for (int c = 0; c < 5; c++) e->pvtw[r][c] = tab[c][r];
    e->std_rhow = m.std_RhoW;
    return 0;
}

static void seed_skip(RefHandle* h, const ocpg_bulk* bk, int64_t n)
{
    SkipStaAnaly& s = h->opt.skipStaAnaly;
    const int NC = h->NC;
    s.flag[n] = bk->skipFlag[n] != 0;
    s.minEigen[n] = bk->skipMinEig[n];
    s.P[n] = bk->skipP[n];
    s.T[n] = bk->skipT[n];
    for (int i = 0; i < NC; i++) s.zi[n * NC + i] = bk->skipZi[n * NC + i];
}

static void read_skip(RefHandle* h, ocpg_bulk* bk, int64_t n)
{
    SkipStaAnaly& s = h->opt.skipStaAnaly;
    const int NC = h->NC;
    bk->skipFlag[n] = s.flag[n] ? 1 : 0;
    bk->skipMinEig[n] = s.minEigen[n];
    bk->skipP[n] = s.P[n];
    bk->skipT[n] = s.T[n];
    for (int i = 0; i < NC; i++) bk->skipZi[n * NC + i] = s.zi[n * NC + i];
}

// Restatement of IsoT_FIM::PassFlashValue around the reference's getters.
static void pass_value(RefHandle* h, ocpg_bulk* bk, int64_t n, const unsigned long long* it0)
{
    MixtureComp& m = *h->mix;
    const int np = 3;
    const int nc = h->NC + 1;
    const int64_t bIdp = n * np;
    int len = 0;
    bk->phaseNum[n] = 0;
    bk->Nt[n] = m.GetNt();
    bk->vf[n] = m.GetVf();
    for (int j = 0; j < np; j++) {
        bk->S[bIdp + j] = m.GetS(j);
        bk->phaseExist[bIdp + j] = m.GetPhaseExist(j);
        if (bk->phaseExist[bIdp + j]) {
            bk->phaseNum[n]++;
            bk->rho[bIdp + j] = m.GetRho(j);
            bk->xi[bIdp + j] = m.GetXi(j);
            bk->mu[bIdp + j] = m.GetMu(j);
            bk->rhoP[bIdp + j] = m.GetRhoP(j);
            bk->xiP[bIdp + j] = m.GetXiP(j);
            bk->muP[bIdp + j] = m.GetMuP(j);
            for (int i = 0; i < nc; i++) {
                bk->xij[bIdp * nc + j * nc + i] = m.GetXij(j, i);
                bk->rhox[bIdp * nc + j * nc + i] = m.GetRhoX(j, i);
                bk->xix[bIdp * nc + j * nc + i] = m.GetXiX(j, i);
                bk->mux[bIdp * nc + j * nc + i] = m.GetMuX(j, i);
            }
        }
        bk->pSderExist[bIdp + j] = m.GetPSderExist(j);
        bk->pVnumCom[bIdp + j] = m.GetPVnumCom(j);
        if (bk->pSderExist[bIdp + j]) len++;
        len += bk->pVnumCom[bIdp + j];
    }
    bk->vfP[n] = m.GetVfP();
    for (int i = 0; i < nc; i++) bk->vfi[n * nc + i] = m.GetVfi(i);
    bk->bRowSizedSdP[n] = len;
    const int L = (nc + 1) * (nc + 1) * np;
    const vector<OCP_DBL>& d = m.GetDXsDXp();
    for (int k = 0; k < len * (nc + 1); k++) bk->dSec_dPri[n * L + k] = d[k];
    if (bk->ftype) bk->ftype[n] = m.ftype;
    if (bk->nonconv) bk->nonconv[n] = m.GetErrorPEC() >= 1e8 ? 1 : 0;
    if (bk->iters) {
        bk->iters[n * 5 + 0] = (int32_t)(m.itersSSMSTA - it0[0]);
        bk->iters[n * 5 + 1] = (int32_t)(m.itersNRSTA - it0[1]);
        bk->iters[n * 5 + 2] = (int32_t)(m.itersSSMSP - it0[2]);
        bk->iters[n * 5 + 3] = (int32_t)(m.itersNRSP - it0[3]);
        bk->iters[n * 5 + 4] = (int32_t)(m.itersRR - it0[4]);
    }
}

int ref_flash_fim(void* p, ocpg_bulk* bk)
{
    Quiet q;
    RefHandle* h = static_cast<RefHandle*>(p);
    if (bk->nb > h->nb) return -1;
    MixtureComp& m = *h->mix;
    const int np = 3;
    const int nc = h->NC + 1;
    for (int64_t n = 0; n < bk->nb; n++) {
        seed_skip(h, bk, n);
        unsigned long long it0[5] = {m.itersSSMSTA, m.itersNRSTA, m.itersSSMSP, m.itersNRSP,
                                     m.itersRR};
        m.FlashFIM(bk->P[n], bk->T[n], &bk->Ni[n * nc], &bk->S[n * np], (USI)bk->phaseNum[n],
                   &bk->xij[n * np * nc], (OCP_USI)n);
        read_skip(h, bk, n);
        pass_value(h, bk, n, it0);
    }
    return 0;
}

int ref_init_flash_fim(void* p, ocpg_bulk* bk, const double* zi)
{
    Quiet q;
    RefHandle* h = static_cast<RefHandle*>(p);
    if (bk->nb > h->nb) return -1;
    MixtureComp& m = *h->mix;
    const int np = 3;
    const int NC = h->NC;
    const int nc = NC + 1;
    for (int64_t n = 0; n < bk->nb; n++) {
        seed_skip(h, bk, n);
        unsigned long long it0[5] = {m.itersSSMSTA, m.itersNRSTA, m.itersSSMSP, m.itersNRSP,
                                     m.itersRR};
        m.InitFlashFIM(bk->P[n], 0.0, bk->T[n], &bk->S[n * np], bk->rockVp[n], &zi[n * NC],
                       (OCP_USI)n);
        for (int i = 0; i < nc; i++) bk->Ni[n * nc + i] = m.GetNi(i);
        read_skip(h, bk, n);
        pass_value(h, bk, n, it0);
    }
    return 0;
}

// IsoT_FIMn::PassFlashValue (src/OCPFluidMethod.cpp:3179-3196): IsoT_FIM's,
// then nj, res and resPc.
static void pass_value_n(RefHandle* h, ocpg_bulk* bk, int64_t n, const unsigned long long* it0)
{
    MixtureComp& m = *h->mix;
    pass_value(h, bk, n, it0);
    const int np = 3;
    const int nc = h->NC + 1;
    for (int j = 0; j < np; j++)
        if (bk->phaseExist[n * np + j]) bk->nj[n * np + j] = m.GetNj(j);
    const vector<OCP_DBL>& r = m.GetRes();
    for (int k = 0; k < bk->bRowSizedSdP[n]; k++) bk->res_n[n * (np * (nc + 1)) + k] = r[k];
    bk->resPc[n] = m.GetResPc();
}

// IsoT_FIMn::CalFlash (src/OCPFluidMethod.cpp:3152-3177) around the
// reference's MixtureComp::FlashFIMn.
int ref_flash_fimn(void* p, ocpg_bulk* bk)
{
    Quiet q;
    RefHandle* h = static_cast<RefHandle*>(p);
    if (bk->nb > h->nb || !bk->nj || !bk->res_n || !bk->resPc) return -1;
    MixtureComp& m = *h->mix;
    const int np = 3;
    const int nc = h->NC + 1;
    std::vector<USI> flagB(np, 0);
    for (int64_t n = 0; n < bk->nb; n++) {
        seed_skip(h, bk, n);
        for (int j = 0; j < np; j++) flagB[j] = bk->phaseExist[n * np + j];
        unsigned long long it0[5] = {m.itersSSMSTA, m.itersNRSTA, m.itersSSMSP, m.itersNRSP,
                                     m.itersRR};
        m.FlashFIMn(bk->P[n], bk->T[n], &bk->Ni[n * nc], &bk->S[n * np], &bk->xij[n * np * nc],
                    &bk->nj[n * np], flagB.data(), (USI)bk->phaseNum[n], (OCP_USI)n);
        read_skip(h, bk, n);
        pass_value_n(h, bk, n, it0);
    }
    return 0;
}

// IsoT_FIMn::InitFlash (src/OCPFluidMethod.cpp:3129-3150).
int ref_init_flash_fimn(void* p, ocpg_bulk* bk, const double* zi)
{
    Quiet q;
    RefHandle* h = static_cast<RefHandle*>(p);
    if (bk->nb > h->nb || !bk->nj || !bk->res_n || !bk->resPc) return -1;
    MixtureComp& m = *h->mix;
    const int np = 3;
    const int NC = h->NC;
    const int nc = NC + 1;
    for (int64_t n = 0; n < bk->nb; n++) {
        seed_skip(h, bk, n);
        unsigned long long it0[5] = {m.itersSSMSTA, m.itersNRSTA, m.itersSSMSP, m.itersNRSP,
                                     m.itersRR};
        m.InitFlashFIMn(bk->P[n], 0.0, bk->T[n], &bk->S[n * np], bk->rockVp[n], &zi[n * NC],
                        (OCP_USI)n);
        for (int i = 0; i < nc; i++) bk->Ni[n * nc + i] = m.GetNi(i);
        read_skip(h, bk, n);
        pass_value_n(h, bk, n, it0);
    }
    return 0;
}

// MixtureComp::CalProdWeight (src/MixtureComp.cpp:830-853) on the
// reference's own object.  The phase volumes vj start from zero, as in a
// freshly allocated object, so a phase absent at surface conditions
// contributes nothing.
int ref_prod_weight(void* p, double Psurf, double Tsurf, const double* Ni, const double* prodPhase,
                    double* weight)
{
    RefHandle* h = static_cast<RefHandle*>(p);
    if (!h || !Ni || !prodPhase || !weight) return OCPG_EINVAL;
    Quiet q;
    std::fill(h->mix->vj.begin(), h->mix->vj.end(), 0.0);
    std::vector<OCP_DBL> pp(prodPhase, prodPhase + 3), w(h->NC + 1, 0.0);
    h->mix->CalProdWeight(Psurf, Tsurf, Ni, pp, w);
    *weight = w[0];
    return OCPG_OK;
}

// MixtureComp::XiPhase and MixtureComp::RhoPhase of one feed: out[0] xi,
// out[1] rho (phase 1 gas, 2 water).
int ref_phase_density(void* p, double P, double T, const double* zi, int phase, double* out)
{
    RefHandle* h = static_cast<RefHandle*>(p);
    if (!h || !zi || !out) return OCPG_EINVAL;
    Quiet q;
    const USI ph = phase == 2 ? WATER : GAS;
    out[0] = h->mix->XiPhase(P, T, zi, ph);
    out[1] = h->mix->RhoPhase(P, 0, T, zi, ph);
    return OCPG_OK;
}

// MixtureComp::Flash of the moles Ni (nc): out = phaseExist[3], vj[3],
// xi[3], rho[3] (zeros where a phase does not exist), Nt.
int ref_flash_phases(void* p, double P, double T, const double* Ni, double* out)
{
    RefHandle* h = static_cast<RefHandle*>(p);
    if (!h || !Ni || !out) return OCPG_EINVAL;
    Quiet q;
    h->mix->Flash(P, T, Ni);
    for (int j = 0; j < 3; j++) {
        const bool ex = h->mix->phaseExist[j];
        out[j] = ex ? 1.0 : 0.0;
        out[3 + j] = ex ? h->mix->vj[j] : 0.0;
        out[6 + j] = ex ? h->mix->xi[j] : 0.0;
        out[9 + j] = ex ? h->mix->rho[j] : 0.0;
    }
    out[12] = h->mix->Nt;
    return OCPG_OK;
}

}  // extern "C"
