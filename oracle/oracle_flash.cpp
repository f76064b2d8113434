// CPU restatement of MixtureComp's FIM flash.  TEST INFRASTRUCTURE ONLY:
// the oracle for the HIP kernels and the CPU baseline of the benchmark.
// Each function cites the reference function it restates; expressions keep
// the reference's operand order so that a build without FP contraction
// reproduces the reference's rounding.
#include "oracle_flash.hpp"

#include <algorithm>
#include <cstring>

namespace oracle {

#ifdef OCP_COUNT_FLOPS
thread_local FlopTally g_tally;
#ifdef OCP_ULP_NOISE
thread_local uint64_t g_noise = 88172645463325252ull;
#endif
#endif

static inline real deltaij(int i, int j) { return i == j ? 1.0 : 0.0; }

static inline real signD(const real& d)
{
    if (d > 0) return 1.0;
    if (d < 0) return -1.0;
    return 0.0;
}

// Sum of |x| (dasum), Euclidean norm (dnrm2).
static inline real Dnorm1(int n, const real* x)
{
    real s = 0;
    for (int i = 0; i < n; i++) s += fabs(x[i]);
    return s;
}
static inline real Dnorm2(int n, const real* x)
{
    real s = 0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

// NTcubicroot, src/MixtureComp.cpp:5078-5102.
static void NTcubicroot(real& root, const real& a, const real& b, const real& c)
{
    real e = root * (root * (root + a) + b) + c;
    real df;
    double iter = 0;
    real optroot = root;
    real opte = fabs(e);
    while (fabs(e) > 1E-8) {
        df = root * (3 * root + 2 * a) + b;
        root = root - e / df;
        iter++;
        if (iter > 10) break;
        e = root * (root * (root + a) + b) + c;
        if (fabs(e) <= opte) {
            opte = fabs(e);
            optroot = root;
        }
    }
    root = optroot;
}

Mixture::Mixture(const ocpg_eos& eos)
{
    // MixtureComp::MixtureComp, src/MixtureComp.cpp:28-165 (values arrive
    // already derived), AllocateEoS :875-887, AllocateOthers :2321-2350.
    NC = eos.nc;
    numCom = NC + 1;
    numPhase = NPH;
    for (int i = 0; i < NC; i++) {
        Tc[i] = eos.Tc[i];
        Pc[i] = eos.Pc[i];
        Vc[i] = eos.Vc[i];
        MWC[i] = eos.MW[i];
        Acf[i] = eos.Acf[i];
        OmegaA[i] = eos.OmegaA[i];
        OmegaB[i] = eos.OmegaB[i];
        Vshift[i] = eos.Vshift[i];
        Vcvis[i] = eos.Vcvis[i];
    }
    for (int i = 0; i < NC * NC; i++) BIC[i] = eos.BIC[(i / NC) * OCPG_MAXNC + (i % NC)];
    for (int i = 0; i < 5; i++) LBCcoef[i] = eos.LBCcoef[i];
    delta1 = 2.41421356237;
    delta2 = -0.41421356237;
    delta1P2 = delta1 + delta2;
    delta1M2 = delta1 - delta2;
    delta1T2 = delta1 * delta2;
    for (int i = 0; i < NC; i++) sqrtMWi[i] = sqrt(MWC[i]);
    for (int i = 0; i < NC; i++)
        muAux1I[i] = 5.4402 * pow(Tc[i], 1.0 / 6) / pow(Pc[i], 2.0 / 3);
    pvtwRows = eos.pvtw_rows;
    for (int r = 0; r < pvtwRows; r++)
        for (int c = 0; c < 5; c++) pvtw[r][c] = eos.pvtw[r][c];
    std_RhoW = eos.std_rhow;

    ssmStaMaxIt = eos.ssmsta_maxit;
    ssmStaTol = eos.ssmsta_tol;
    ssmStaEYt = eos.ssmsta_eyt;
    ssmStaTol2 = ssmStaTol * ssmStaTol;
    ssmStaKtol = 1E-4;
    ssmStaDYtol = 1E-6;
    nrStaMaxIt = eos.nrsta_maxit;
    nrStaTol = eos.nrsta_tol;
    ssmSpMaxIt = eos.ssmsp_maxit;
    ssmSpTol = eos.ssmsp_tol;
    ssmSpTol2 = ssmSpTol * ssmSpTol;
    nrSpMaxIt = eos.nrsp_maxit;
    nrSpTol = eos.nrsp_tol;
    rrMaxIt = eos.rr_maxit;
    rrTol = eos.rr_tol;

    lNP = 0;
    NP = 1;
    P = T = 0;
    vf = Nt = vfP = 0;
    for (int i = 0; i <= MAXNC; i++) Ni[i] = vfi[i] = 0;
    for (int j = 0; j < NPH; j++) {
        phaseExist[j] = false;
        S[j] = vj[j] = nj[j] = rho[j] = xi[j] = mu[j] = 0;
        rhoP[j] = xiP[j] = muP[j] = 0;
        pSderExist[j] = false;
        pVnumCom[j] = 0;
        vjp[j] = 0;
        for (int i = 0; i <= MAXNC; i++) vji[j][i] = 0;
    }
    for (int k = 0; k < NPH * (MAXNC + 1); k++) xij[k] = rhox[k] = xix[k] = mux[k] = 0;
    for (auto& v : dXsdXp) v = 0;
    ftype = 0;
    ePEC = 0;
    for (auto& it : iters) it = 0;
}

// MixtureComp::InitPTN / InitPTZ, include/MixtureComp.hpp:170-184.
void Mixture::InitPTN(double Pin, double Tin, const double* Niin)
{
    P = Pin;
    T = Tin;
    for (int i = 0; i < numCom; i++) Ni[i] = Niin[i];
    Nh = Dnorm1(NC, Ni);
    for (int i = 0; i < NC; i++) zi[i] = Ni[i] / Nh;
}

void Mixture::InitPTZ(double Pin, double Tin, const double* Ziin)
{
    P = Pin;
    T = Tin;
    for (int i = 0; i < NC; i++) zi[i] = Ziin[i];
}

// SkipStaAnaly::IfSkip + CalFtypeFIM, src/AcceleratePVT.cpp:54-120.
void Mixture::CalFtypeFIM(const double* Sjin, const SkipState& skip)
{
    bool skipIt = false;
    if (skip.flag) {
        skipIt = true;
        if (fabs(1 - skip.P / P) >= skip.minEigen / 10) {
            skipIt = false;
        } else if (fabs(skip.T - T) >= skip.minEigen * 10) {
            skipIt = false;
        } else {
            for (int i = 0; i < NC; i++) {
                if (fabs(Ni[i] / Nh - skip.zi[i]) >= skip.minEigen / 10) {
                    skipIt = false;
                    break;
                }
            }
        }
    }
    if (skipIt) {
        ftype = 1;
    } else if (lNP >= 2) {
        ftype = 2;
        for (int j = 0; j < NPMAX; j++) {
            if (Sjin[j] < 1E-4) {
                ftype = 0;
                break;
            }
        }
    } else {
        ftype = 0;
    }
}

// MixtureComp::FlashFIM, src/MixtureComp.cpp:494-573.
void Mixture::FlashFIM(double Pin, double Tin, const double* Niin, const double* Sjin,
                       int lastNP, const double* xijin, SkipState& skip)
{
    for (auto& it : iters) it = 0;
    lNP = lastNP > 0 ? lastNP - 1 : 0;
    if (lNP == 2) {
        for (int i = 0; i < NC; i++) lKs[i] = xijin[i] / xijin[i + numCom];
    }
    InitPTN(Pin, Tin + CONV5, Niin);
    CalFtypeFIM(Sjin, skip);
    CalFlash();
    CalVfiVfp_full02();
    WaterProperties(false, Sjin, 0);
    CalSaturation();
    CaldXsdXpAPI02p();
    CalXiPNX_partial();
    CalRhoPX_partial();
    CalMuPXLBC_partial();
    SetDerivFlags();
    CalSkipForNextStep(skip);
}

// MixtureComp::InitFlashFIM, src/MixtureComp.cpp:261-353.
void Mixture::InitFlashFIM(double Pin, double Tin, const double* Sjin, double Vpore,
                           const double* Ziin, SkipState& skip)
{
    for (auto& it : iters) it = 0;
    ftype = 0;
    lNP = 0;
    InitPTZ(Pin, Tin + CONV5, Ziin);
    PhaseEquilibrium();
    CalMW();
    CalVfXiRho();
    CalViscoLBC();
    IdentifyPhase();
    CopyPhase();
    real Sw = Sjin[numPhase - 1];
    Nh = Vpore * (1 - Sw) / vf;
    for (int j = 0; j < NP; j++) nu[j] = Nh * nu[j];
    for (int j = 0; j < NPMAX; j++) vj[j] = Nh * vj[j];
    vf *= Nh;
    CalVfiVfp_full02();
    for (int i = 0; i < NC; i++) Ni[i] = zi[i] * Nh;
    WaterProperties(true, Sjin, Vpore);
    CalSaturation();
    CaldXsdXpAPI02p();
    CalXiPNX_partial();
    CalRhoPX_partial();
    CalMuPXLBC_partial();
    SetDerivFlags();
    CalSkipForNextStep(skip);
}

// PVTW table evaluation, OCPTable::Eval_All(0, ...) src/OCPTable.cpp:47-86.
// The search cursor does not change the result, so none is kept.
void Mixture::PVTWEval(real Pv, real* data, real* cdata) const
{
    const int nRow = pvtwRows;
    int seg = -1;
    for (int i = 0; i + 1 < nRow; i++) {
        if (Pv >= pvtw[i][0] && Pv < pvtw[i + 1][0]) {
            seg = i;
            break;
        }
    }
    if (seg < 0) {
        const int r = (Pv < pvtw[0][0]) ? 0 : nRow - 1;
        for (int k = 0; k < 5; k++) {
            cdata[k] = 0;
            data[k] = pvtw[r][k];
        }
        return;
    }
    for (int k = 0; k < 5; k++) {
        cdata[k] = (pvtw[seg + 1][k] - pvtw[seg][k]) / (pvtw[seg + 1][0] - pvtw[seg][0]);
        data[k] = pvtw[seg][k] + cdata[k] * (Pv - pvtw[seg][0]);
    }
}

// Water block of FlashFIM (src/MixtureComp.cpp:521-546) and of InitFlashFIM
// (:321-345, where Ni[water] comes from Vpore*Sw*xi).
void Mixture::WaterProperties(bool init, const double* Sjin, double Vpore)
{
    const int Wpid = numPhase - 1;
    const int Wcid = numCom - 1;
    phaseExist[Wpid] = true;
    xij[Wpid * numCom + Wcid] = 1.0;
    if (!init) {
        nj[Wpid] = Ni[Wcid];
        Nt = Nh + Ni[Wcid];
    }
    real data[5], cdata[5];
    PVTWEval(P, data, cdata);
    real Pw0 = data[0];
    real bw0 = data[1];
    real cbw = data[2];
    real bw = bw0 * (1 - cbw * (P - Pw0));
    real bwp = -cbw * bw0;
    mu[Wpid] = data[3];
    xi[Wpid] = 1 / (CONV1 * bw);
    rho[Wpid] = std_RhoW / bw;
    muP[Wpid] = cdata[3];
    xiP[Wpid] = -bwp / (bw * bw * CONV1);
    rhoP[Wpid] = CONV1 * xiP[Wpid] * std_RhoW;
    if (init) {
        real Sw = Sjin[numPhase - 1];
        Ni[Wcid] = Vpore * Sw * xi[Wpid];
        nj[Wpid] = Ni[Wcid];
        Nt = Nh + Ni[Wcid];
    }
    vj[Wpid] = CONV1 * Ni[Wcid] * bw;
    vfi[Wcid] = CONV1 * bw;
    const real vwp = CONV1 * Ni[Wcid] * bwp;
    vf += vj[Wpid];
    vfP += vwp;
    vji[numPhase - 1][numCom - 1] = vfi[Wcid];
    vjp[numPhase - 1] = vwp;
}

// pSderExist / pVnumCom, src/MixtureComp.cpp:559-570.
void Mixture::SetDerivFlags()
{
    for (int j = 0; j < NPH; j++) {
        pSderExist[j] = false;
        pVnumCom[j] = 0;
    }
    if (phaseExist[0]) {
        pSderExist[0] = true;
        pVnumCom[0] = NC;
    }
    if (phaseExist[1]) {
        pSderExist[1] = true;
        pVnumCom[1] = NC;
    }
    pSderExist[2] = true;
}

// MixtureComp::CalFlash, src/MixtureComp.cpp:697-708 (no miscibility).
void Mixture::CalFlash()
{
    PhaseEquilibrium();
    for (int j = 0; j < NP; j++) nu[j] = Nh * nu[j];
    CalMW();
    CalVfXiRho();
    CalViscoLBC();
    IdentifyPhase();
    CopyPhase();
}

// MixtureComp::CalProdWeight (src/MixtureComp.cpp:830-853) after
// MixtureComp::Flash (:167-193): a cold flash of the well stream at surface
// conditions, water from PVTW, then the stream's volume per mole in the
// producer's rate units.
double Mixture::ProdWeight(double Pin, double Tin, const double* Niin, const double* prodPhase)
{
    ftype = 0;
    lNP = 0;
    InitPTN(Pin, Tin + CONV5, Niin);
    for (int j = 0; j < NPH; j++) vj[j] = 0;
    CalFlash();
    const int Wpid = numPhase - 1;
    const int Wcid = numCom - 1;
    Nt = Nh + Ni[Wcid];
    real data[5], cdata[5];
    PVTWEval(P, data, cdata);
    const real Pw0 = data[0];
    const real bw0 = data[1];
    const real cbw = data[2];
    const real bw = bw0 * (1 - cbw * (P - Pw0));
    xi[Wpid] = 1 / (CONV1 * bw);
    vj[Wpid] = CONV1 * Ni[Wcid] * bw;
    const real qt = Nt;
    real factor[NPH];
    factor[0] = vj[0] / qt / CONV1;
    factor[1] = vj[1] / qt / 1000;
    factor[2] = xi[2] * vj[2] / qt;
    real tmp = 0;
    for (int i = 0; i < 3; i++) tmp += factor[i] * prodPhase[i];
    return static_cast<double>(tmp);
}

double Mixture::FlashPhases(double Pin, double Tin, const double* Niin, bool* ex, double* vjo, double* xio,
                            double* rhoo)
{
    ftype = 0;
    lNP = 0;
    InitPTN(Pin, Tin + CONV5, Niin);
    CalFlash();
    for (int j = 0; j < NPH - 1; j++) {
        ex[j] = phaseExist[j];
        vjo[j] = ex[j] ? static_cast<double>(vj[j]) : 0.0;
        xio[j] = ex[j] ? static_cast<double>(xi[j]) : 0.0;
        rhoo[j] = ex[j] ? static_cast<double>(rho[j]) : 0.0;
    }
    const int Wpid = numPhase - 1;
    const int Wcid = numCom - 1;
    real data[5], cdata[5];
    PVTWEval(P, data, cdata);
    const real bw = data[1] * (1 - data[2] * (P - data[0]));
    ex[Wpid] = true;
    xio[Wpid] = static_cast<double>(1 / (CONV1 * bw));
    rhoo[Wpid] = static_cast<double>(std_RhoW / bw);
    vjo[Wpid] = static_cast<double>(CONV1 * Ni[Wcid] * bw);
    return static_cast<double>(Nh + Ni[Wcid]);
}

void Mixture::ProdRate(double Pin, double Tin, const double* Niin, double* rate)
{
    bool ex[NPH];
    double v[NPH], x_[NPH], r[NPH];
    FlashPhases(Pin, Tin, Niin, ex, v, x_, r);
    rate[0] = v[0] / CONV1;
    rate[1] = v[1] / 1000;
    rate[2] = v[2] * x_[2];
}

void Mixture::PhaseDensity(double Pin, double Tin, const double* Ziin, int phase, double* xio, double* rhoo)
{
    if (phase == NPH - 1) {
        real data[5], cdata[5];
        PVTWEval(Pin, data, cdata);
        const real bw = data[1] * (1 - data[2] * (Pin - data[0]));
        *xio = static_cast<double>(1 / (CONV1 * bw));
        *rhoo = static_cast<double>(std_RhoW / bw);
        return;
    }
    InitPTZ(Pin, Tin + CONV5, Ziin);
    NP = 1;
    CalAiBi();
    CalAjBj(Aj[0], Bj[0], zi);
    SolEoS(Zj[0], Aj[0], Bj[0]);
    real vtmp = Zj[0] * GAS_CONSTANT * T / P;
    for (int i = 0; i < NC; i++) vtmp -= zi[i] * Vshift[i];
    const real xitmp = 1 / vtmp;
    real mw = 0;
    for (int i = 0; i < NC; i++) mw += zi[i] * MWC[i];
    *xio = static_cast<double>(xitmp);
    *rhoo = static_cast<double>(mw * xitmp);
}

// MixtureComp::SolEoS, src/MixtureComp.cpp:889-914.
void Mixture::SolEoS(real& ZjT, const real& AjT, const real& BjT)
{
    const real aj = AjT;
    const real bj = BjT;
    const real a = (delta1 + delta2 - 1) * bj - 1;
    const real b = (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1));
    const real c = -(aj * bj + delta1 * delta2 * bj * bj * (bj + 1));
    int flag = CubicRoot(a, b, c, true);
    if (flag == 1) {
        ZjT = Ztmp[0];
    } else {
        real zj1 = Ztmp[0];
        real zj2 = Ztmp[2];
        real dG = (zj2 - zj1) + log((zj1 - bj) / (zj2 - bj)) -
                  aj / (bj * (delta2 - delta1)) *
                      log((zj1 + delta1 * bj) * (zj2 + delta2 * bj) /
                          ((zj1 + delta2 * bj) * (zj2 + delta1 * bj)));
        if (dG > 0)
            ZjT = zj1;
        else
            ZjT = zj2;
    }
}

// MixtureComp::CubicRoot, src/MixtureComp.cpp:4999-5056.
int Mixture::CubicRoot(const real& a, const real& b, const real& c, bool NTflag)
{
    real Q = (a * a - 3 * b) / 9;
    real R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
    real Q3 = Q * Q * Q;
    real M = R * R - Q3;
    if (M <= 0) {
        real theta = acos(R / sqrt(Q3));
        Ztmp[0] = -2 * sqrt(Q) * cos(theta / 3) - a / 3;
        Ztmp[1] = -2 * sqrt(Q) * cos((theta + 2 * PI_TRUNC) / 3) - a / 3;
        Ztmp[2] = -2 * sqrt(Q) * cos((theta - 2 * PI_TRUNC) / 3) - a / 3;
        if (NTflag) {
            NTcubicroot(Ztmp[0], a, b, c);
            NTcubicroot(Ztmp[1], a, b, c);
            NTcubicroot(Ztmp[2], a, b, c);
        }
        std::sort(Ztmp, Ztmp + 3);
        return 3;
    } else {
        real tmp1 = -R + sqrt(M);
        real tmp2 = R + sqrt(M);
        real S_ = signD(tmp1) * pow(fabs(tmp1), 1.0 / 3);
        real T_ = -signD(tmp2) * pow(fabs(tmp2), 1.0 / 3);
        Ztmp[0] = S_ + T_ - a / 3;
        if (NTflag) NTcubicroot(Ztmp[0], a, b, c);
        return 1;
    }
}

// MixtureComp::CalAiBi, src/MixtureComp.cpp:916-937.
void Mixture::CalAiBi()
{
    real acf, mwi, Pri, Tri;
    for (int i = 0; i < NC; i++) {
        acf = Acf[i];
        if (acf <= 0.49) {
            mwi = 0.37464 + 1.54226 * acf - 0.26992 * pow(acf, 2);
        } else {
            mwi = 0.379642 + 1.48503 * acf - 0.164423 * pow(acf, 2) +
                  0.016667 * pow(acf, 3);
        }
        Pri = P / Pc[i];
        Tri = T / Tc[i];
        Ai[i] = OmegaA[i] * Pri / pow(Tri, 2) * pow((1 + mwi * (1 - sqrt(Tri))), 2);
        Bi[i] = OmegaB[i] * Pri / Tri;
    }
}

// MixtureComp::CalAjBj, src/MixtureComp.cpp:955-969.
void Mixture::CalAjBj(real& AjT, real& BjT, const real* xj) const
{
    AjT = 0;
    BjT = 0;
    for (int i1 = 0; i1 < NC; i1++) {
        BjT += Bi[i1] * xj[i1];
        AjT += xj[i1] * xj[i1] * Ai[i1] * (1 - BIC[i1 * NC + i1]);
        for (int i2 = 0; i2 < i1; i2++) {
            AjT += 2 * xj[i1] * xj[i2] * sqrt(Ai[i1] * Ai[i2]) * (1 - BIC[i1 * NC + i2]);
        }
    }
}

// MixtureComp::CalFugPhi (phi and fug), src/MixtureComp.cpp:1035-1071.
void Mixture::CalFugPhi(real* phiT, real* fugT, const real* xj)
{
    real aj, bj, zj;
    CalAjBj(aj, bj, xj);
    SolEoS(zj, aj, bj);
    const real m1 = delta1;
    const real m2 = delta2;
    const real m1Mm2 = delta1M2;
    real tmp;
    for (int i = 0; i < NC; i++) {
        tmp = 0;
        for (int k = 0; k < NC; k++) {
            tmp += 2 * (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]) * xj[k];
        }
        phiT[i] = exp(Bi[i] / bj * (zj - 1) - log(zj - bj) -
                      aj / (m1Mm2) / bj * (tmp / aj - Bi[i] / bj) *
                          log((zj + m1 * bj) / (zj + m2 * bj)));
        fugT[i] = phiT[i] * xj[i] * P;
    }
    Asta = aj;
    Bsta = bj;
    Zsta = zj;
}

// MixtureComp::CalFugPhi (fug only), src/MixtureComp.cpp:1073-1101.
void Mixture::CalFugPhiFugOnly(real* fugT, const real* xj)
{
    real aj, bj, zj;
    CalAjBj(aj, bj, xj);
    SolEoS(zj, aj, bj);
    const real m1 = delta1;
    const real m2 = delta2;
    const real m1Mm2 = delta1M2;
    real tmp;
    for (int i = 0; i < NC; i++) {
        tmp = 0;
        for (int k = 0; k < NC; k++) {
            tmp += 2 * (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]) * xj[k];
        }
        tmp = exp(Bi[i] / bj * (zj - 1) - log(zj - bj) -
                  aj / (m1Mm2) / bj * (tmp / aj - Bi[i] / bj) *
                      log((zj + m1 * bj) / (zj + m2 * bj)));
        fugT[i] = tmp * xj[i] * P;
    }
    Asta = aj;
    Bsta = bj;
    Zsta = zj;
}

// MixtureComp::CalFugPhiAll, src/MixtureComp.cpp:1103-1145.
void Mixture::CalFugPhiAll()
{
    real tmp;
    const real m1 = delta1;
    const real m2 = delta2;
    const real m1Mm2 = delta1M2;
    for (int j = 0; j < NP; j++) {
        const real* xj = x[j];
        real* phiT = phi[j];
        real* fugT = fug[j];
        real& aj = Aj[j];
        real& bj = Bj[j];
        real& zj = Zj[j];
        CalAjBj(aj, bj, xj);
        SolEoS(zj, aj, bj);
        for (int i = 0; i < NC; i++) {
            tmp = 0;
            for (int k = 0; k < NC; k++) {
                tmp += 2 * (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]) * xj[k];
            }
            phiT[i] = exp(Bi[i] / bj * (zj - 1) - log(zj - bj) -
                          aj / (m1Mm2) / bj * (tmp / aj - Bi[i] / bj) *
                              log((zj + m1 * bj) / (zj + m2 * bj)));
            fugT[i] = phiT[i] * xj[i] * P;
        }
    }
}

// MixtureComp::CalMW, CalVfXiRho, CalSaturation, FindMWmax, x2n
// (src/MixtureComp.cpp:1147-1212).
void Mixture::CalMW()
{
    for (int j = 0; j < NP; j++) {
        MW[j] = 0;
        for (int i = 0; i < NC; i++) MW[j] += x[j][i] * MWC[i];
    }
}

void Mixture::CalVfXiRho()
{
    vf = 0;
    real tmp;
    for (int j = 0; j < NP; j++) {
        const real* xj = x[j];
        tmp = Zj[j] * GAS_CONSTANT * T / P;
        for (int i = 0; i < NC; i++) tmp -= xj[i] * Vshift[i];
        vC[j] = tmp * nu[j];
        vf += vC[j];
        xiC[j] = 1 / tmp;
        rhoC[j] = MW[j] * xiC[j];
    }
}

void Mixture::CalSaturation()
{
    for (int j = 0; j < numPhase; j++) {
        S[j] = 0;
        if (phaseExist[j]) S[j] = vj[j] / vf;
    }
}

int Mixture::FindMWmax()
{
    int tmpId = 0;
    real tmpMW = MW[0];
    for (int j = 1; j < NP; j++) {
        if (tmpMW < MW[j]) {
            tmpMW = MW[j];
            tmpId = j;
        }
    }
    return tmpId;
}

void Mixture::x2n()
{
    for (int j = 0; j < NP; j++)
        for (int i = 0; i < NC; i++) n[j][i] = nu[j] * x[j][i];
}

// MixtureComp::CalKwilson, src/MixtureComp.cpp:1269-1277.
void Mixture::CalKwilson()
{
    for (int i = 0; i < NC; i++) {
        Kw[0][i] = (Pc[i] / P) * exp(5.373 * (1 + Acf[i]) * (1 - Tc[i] / T));
        Kw[1][i] = 1 / Kw[0][i];
        Kw[2][i] = pow(Kw[0][i], 1.0 / 3);
        Kw[3][i] = pow(Kw[1][i], 1.0 / 3);
    }
}

// MixtureComp::PhaseEquilibrium, src/MixtureComp.cpp:1279-1353.
// PhaseStable / PhaseSplit, with their share of the flop tally in the
// counting build.
bool Mixture::CountedStable()
{
#ifdef OCP_COUNT_FLOPS
    const uint64_t f0 = g_tally.flops, t0 = g_tally.trans;
    const bool r = PhaseStable();
    g_tally.staFlops += g_tally.flops - f0;
    g_tally.staTrans += g_tally.trans - t0;
    return r;
#else
    return PhaseStable();
#endif
}

void Mixture::CountedSplit()
{
#ifdef OCP_COUNT_FLOPS
    const uint64_t f0 = g_tally.flops, t0 = g_tally.trans;
    PhaseSplit();
    g_tally.splitFlops += g_tally.flops - f0;
    g_tally.splitTrans += g_tally.trans - t0;
#else
    PhaseSplit();
#endif
}

void Mixture::PhaseEquilibrium()
{
    switch (ftype) {
        case 0:
            flagSkip = true;
            NP = 1;
            nu[0] = 1;
            for (int i = 0; i < NC; i++) x[0][i] = zi[i];
            CalAiBi();
            CalAjBj(Aj[0], Bj[0], x[0]);
            SolEoS(Zj[0], Aj[0], Bj[0]);
            CalKwilson();
            while (!CountedStable()) {
                NP++;
                CountedSplit();
                if (NP == NPMAX || NP == 1) break;
            }
            if (NP > 1) flagSkip = false;
            if (NP == 1) {
                if (nrStaConflag)
                    ePEC = nrStaRealTol;
                else if (ssmStaConflag)
                    ePEC = ssmStaRealTol;
                else
                    ePEC = 1E8;
            } else {
                if (nrSpConflag)
                    ePEC = nrSpRealTol;
                else if (ssmSpConflag)
                    ePEC = ssmSpRealTol;
                else
                    ePEC = 1E8;
            }
            break;
        case 1:
            flagSkip = true;
            NP = 1;
            nu[0] = 1;
            for (int i = 0; i < NC; i++) x[0][i] = zi[i];
            CalAiBi();
            CalAjBj(Aj[0], Bj[0], x[0]);
            SolEoS(Zj[0], Aj[0], Bj[0]);
            ePEC = 0.0;
            break;
        case 2:
            flagSkip = false;
            NP = 2;
            Yt = 1.01;
            CalAiBi();
            CalKwilson();
            CountedSplit();
            if (nrSpConflag)
                ePEC = nrSpRealTol;
            else if (ssmSpConflag)
                ePEC = ssmSpRealTol;
            else
                ePEC = 1E8;
            break;
    }
}

// MixtureComp::PhaseStable, src/MixtureComp.cpp:1355-1390.
bool Mixture::PhaseStable()
{
    int testPId;
    if (NP == 1) {
        testPId = 0;
    } else {
        CalMW();
        testPId = FindMWmax();
    }
    ssmStaConflag = false;
    ssmStaCurIt = 0;
    nrStaConflag = false;
    nrStaCurIt = 0;

    bool flag;
    int tmpNP = NP;
    if (lNP == 0) {
        flag = StableSSM(testPId);
    } else {
        flag = StableSSM01(testPId);
        if (!flag) tmpNP++;
        if (tmpNP != lNP) {
            flag = StableSSM(testPId);
            flagSkip = false;
        }
    }
    iters[0] += ssmStaCurIt;
    iters[1] += nrStaCurIt;
    return flag;
}

// MixtureComp::StableSSM01 (relaxed), src/MixtureComp.cpp:1392-1490.
bool Mixture::StableSSM01(int Id)
{
    real Stol = ssmStaTol2;
    int maxIt = ssmStaMaxIt;
    real eYt = ssmStaEYt;
    real Ktol = ssmStaKtol;
    real dYtol = ssmStaDYtol;
    real Se = 0, Sk = 0, dY;
    bool flag, Tsol;
    int iter, k;

    const real* xj = x[Id];
    CalFugPhi(phi[Id], fug[Id], xj);
    const real* fugId = fug[Id];
    real* ks = Ks[0];
    for (k = 0; k < 2; k++) {
        for (int i = 0; i < NC; i++) ks[i] = Kw[k][i];
        iter = 0;
        flag = false;
        Tsol = false;
        while (true) {
            Yt = 0;
            for (int i = 0; i < NC; i++) {
                Y[i] = xj[i] * ks[i];
                Yt += Y[i];
            }
            {
                const real s = 1 / Yt;
                for (int i = 0; i < NC; i++) Y[i] = s * Y[i];
            }
            if (iter > 0) {
                dY = 0;
                for (int i = 0; i < NC; i++) dY += pow((Y[i] - di[i]), 2);
                if (dY < dYtol) {
                    flag = true;
                    break;
                }
            }
            CalFugPhiFugOnly(fugSta, Y);
            Se = 0;
            Sk = 0;
            for (int i = 0; i < NC; i++) {
                di[i] = fugId[i] / (fugSta[i] * Yt);
                ks[i] *= di[i];
                Se += pow(di[i] - 1, 2);
                Sk += pow(log(ks[i]), 2);
            }
            iter++;
            if (Se < Stol) {
                flag = true;
                break;
            }
            if (Sk < Ktol) {
                flag = true;
                Tsol = true;
                break;
            }
            if (iter > maxIt) break;
            for (int i = 0; i < NC; i++) di[i] = Y[i];
        }
        ssmStaCurIt += iter;
        if (flag && Yt > 1 - 0.1 && Sk > 1) flagSkip = false;
        if (flag && Yt > 1 + eYt) {
            ssmStaConflag = true;
            ssmStaRealTol = sqrt(Se);
            return false;
        }
    }
    ssmStaRealTol = sqrt(Se);
    return true;
}

// MixtureComp::StableSSM (strict), src/MixtureComp.cpp:1492-1588.
bool Mixture::StableSSM(int Id)
{
    const real* xj = x[Id];
    CalFugPhi(phi[Id], fug[Id], xj);
    const real* fugId = fug[Id];
    for (int i = 0; i < NC; i++) di[i] = phi[Id][i] * xj[i];

    real Stol = ssmStaTol2;
    int maxIt = ssmStaMaxIt;
    real eYt = ssmStaEYt;
    real Se;
    bool flag;
    int iter, k;
    for (k = 0; k < 4; k++) {
        Yt = 0;
        for (int i = 0; i < NC; i++) {
            Y[i] = xj[i] / Kw[k][i];
            Yt += Y[i];
        }
        {
            const real s = 1 / Yt;
            for (int i = 0; i < NC; i++) Y[i] = s * Y[i];
        }
        CalFugPhi(phiSta, fugSta, Y);
        Se = 0;
        for (int i = 0; i < NC; i++) Se += pow(log(fugSta[i] / fugId[i] * Yt), 2);
        flag = true;
        iter = 0;
        while (Se > Stol) {
            Yt = 0;
            for (int i = 0; i < NC; i++) {
                Y[i] = di[i] / phiSta[i];
                Yt += Y[i];
            }
            {
                const real s = 1 / Yt;
                for (int i = 0; i < NC; i++) Y[i] = s * Y[i];
            }
            CalFugPhi(phiSta, fugSta, Y);
            Se = 0;
            for (int i = 0; i < NC; i++) Se += pow(log(fugSta[i] / fugId[i] * Yt), 2);
            iter++;
            if (iter > maxIt) {
                flag = false;
                break;
            }
        }
        (void)flag;
        ssmStaCurIt += iter;
        flag = StableNR(Id);
        if (flag && Yt > 1 + eYt) {
            ssmStaConflag = true;
            ssmStaRealTol = sqrt(Se);
            return false;
        }
    }
    ssmStaRealTol = sqrt(Se);
    return true;
}

// MixtureComp::StableNR, src/MixtureComp.cpp:1590-1636.
bool Mixture::StableNR(int Id)
{
    for (int i = 0; i < NC; i++) resSTA[i] = log(fug[Id][i] / (fugSta[i] * Yt));
    int maxIt = nrStaMaxIt;
    real Stol = nrStaTol;
    real Se = Dnorm2(NC, resSTA);
    real alpha = 1;
    int iter = 0;
    while (Se > Stol) {
        CalFugXSTA();
        AssembleJmatSTA();
        SymIndefSolve(NC, JmatSTA, resSTA);
        for (int i = 0; i < NC; i++) Y[i] = Yt * Y[i];
        for (int i = 0; i < NC; i++) Y[i] = Y[i] + alpha * resSTA[i];
        Yt = 0;
        for (int i = 0; i < NC; i++) Yt += Y[i];
        {
            const real s = 1 / Yt;
            for (int i = 0; i < NC; i++) Y[i] = s * Y[i];
        }
        CalFugPhiFugOnly(fugSta, Y);
        for (int i = 0; i < NC; i++) resSTA[i] = log(fug[Id][i] / (fugSta[i] * Yt));
        Se = Dnorm2(NC, resSTA);
        iter++;
        if (iter > maxIt) {
            nrStaCurIt += iter;
            nrStaConflag = false;
            nrStaRealTol = Se;
            return false;
        }
    }
    nrStaCurIt += iter;
    nrStaConflag = true;
    nrStaRealTol = Se;
    return true;
}

// MixtureComp::CalFugXSTA, src/MixtureComp.cpp:1638-1714.
void Mixture::CalFugXSTA()
{
    real* fugx = fugX;
    real aj = Asta;
    real bj = Bsta;
    real zj = Zsta;
    real tmp = 0;
    const real m1 = delta1;
    const real m2 = delta2;
    const real m1Pm2 = delta1P2;
    const real m1Mm2 = delta1M2;
    const real m1Tm2 = delta1T2;

    for (int i = 0; i < NC; i++) Bx[i] = Bi[i];
    for (int i = 0; i < NC; i++) {
        tmp = 0;
        for (int k = 0; k < NC; k++) tmp += Y[k] * (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]);
        Ax[i] = 2 * tmp;
        Zx[i] = ((bj - zj) * Ax[i] + ((aj + m1Tm2 * (3 * bj * bj + 2 * bj)) +
                                      ((m1Pm2) * (2 * bj + 1) - 2 * m1Tm2 * bj) * zj -
                                      (m1Pm2 - 1) * zj * zj) *
                                         Bx[i]) /
                (3 * zj * zj + 2 * ((m1Pm2 - 1) * bj - 1) * zj +
                 (aj + m1Tm2 * bj * bj - (m1Pm2)*bj * (bj + 1)));
    }
    real C, E, G;
    real Cxk, Dxk, Exk, Gxk;
    real aik;
    G = (zj + m1 * bj) / (zj + m2 * bj);
    for (int i = 0; i < NC; i++) {
        C = Y[i] * P / (zj - bj);
        E = -aj / ((m1Mm2)*bj) * (Ax[i] / aj - Bx[i] / bj);
        for (int k = 0; k < NC; k++) {
            aik = (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]);
            Cxk = ((zj - bj) * deltaij(i, k) - Y[i] * (Zx[k] - Bx[k])) * P /
                  ((zj - bj) * (zj - bj));
            Dxk = Bx[i] / bj * (Zx[k] - Bx[k] * (zj - 1) / bj);
            Exk = (2 * (aj / bj * Bx[k] * Bx[i] + bj * aik) - Ax[i] * Bx[k] -
                   Ax[k] * Bi[i]) /
                  (bj * bj);
            Exk /= -(m1Mm2);
            Gxk = (m1Mm2) / (zj + m2 * bj) / (zj + m2 * bj) * (zj * Bx[k] - Zx[k] * bj);
            fugx[i * NC + k] = 1 / C * Cxk + Dxk + Exk * log(G) + E / G * Gxk;
        }
    }
}

// MixtureComp::AssembleJmatSTA, src/MixtureComp.cpp:1716-1741.
void Mixture::AssembleJmatSTA()
{
    const real* fugx = fugX;
    real tmp;
    for (int i = 0; i < NC; i++) {
        tmp = 0;
        for (int k = 0; k < NC; k++) tmp += Y[k] * fugx[i * NC + k];
        for (int j = 0; j < NC; j++) JmatSTA[i * NC + j] = (fugx[i * NC + j] - tmp + 1) / Yt;
    }
}

// MixtureComp::PhaseSplit, src/MixtureComp.cpp:1743-1782.
void Mixture::PhaseSplit()
{
    ssmSpConflag = false;
    ssmSpCurIt = 0;
    nrSpConflag = false;
    nrSpCurIt = 0;
    rrCurIt = 0;
    SplitSSM2(false);
    SplitNR();
    // Bounded at 100 restarts (the reference has no bound; the HIP kernel has
    // the same one so that no cell can spin forever).
    int restarts = 0;
    while (!nrSpConflag) {
        SplitSSM2(true);
        SplitNR();
        if (!CheckSplit()) break;
        if (ssmSpConflag) break;
        if (++restarts >= 100) break;
    }
    CheckSplit();
    iters[2] += ssmSpCurIt;
    iters[3] += nrSpCurIt;
    iters[4] += rrCurIt;
}

// MixtureComp::CheckSplit, src/MixtureComp.cpp:1784-1836.  The Gibbs-energy
// comparison only prints in the reference and is not restated.
bool Mixture::CheckSplit()
{
    if (NP == 2) {
        real eX = 0;
        real nuMax = std::max(nu[0], nu[1]);
        for (int i = 0; i < NC; i++) eX += (x[0][i] - x[1][i]) * (x[0][i] - x[1][i]);
        if (nuMax < 1 && nrSpConflag && isfinite(eX)) {
            // accept
        } else {
            if (!isfinite(eX) || 1 - nuMax < 1E-3) {
                NP = 1;
                for (int i = 0; i < NC; i++) x[0][i] = zi[i];
                nu[0] = 1;
                CalAjBj(Aj[0], Bj[0], x[0]);
                SolEoS(Zj[0], Aj[0], Bj[0]);
                ssmStaConflag = false;
                nrStaConflag = false;
                return false;
            }
        }
    }
    return true;
}

// MixtureComp::SplitSSM2, src/MixtureComp.cpp:1847-1903.
void Mixture::SplitSSM2(bool flag)
{
    ssmSpConflag = true;
    real Se = 1;
    real Stol = ssmSpTol2;
    int maxIt = ssmSpMaxIt;
    if (!flag) {
        if (lNP == 2) {
            for (int i = 0; i < NC; i++) Ks[NP - 2][i] = lKs[i];
        } else {
            for (int i = 0; i < NC; i++) Ks[NP - 2][i] = Kw[0][i];
        }
    } else {
        Stol *= 1E-1;
        maxIt *= 2;
    }
    if (Yt < 1.1) {
        Stol *= 1E-1;
        maxIt *= 2;
    }
    int iter = 0;
    while (Se > Stol) {
        RachfordRice2();
        UpdateXRR();
        CalFugPhiAll();
        Se = 0;
        for (int i = 0; i < NC; i++) {
            Se += pow(fug[1][i] / fug[0][i] - 1, 2);
            Ks[0][i] = phi[1][i] / phi[0][i];
        }
        iter++;
        if (iter > maxIt) {
            ssmSpConflag = false;
            break;
        }
    }
    ssmSpRealTol = sqrt(Se);
    ssmSpCurIt += iter;
}

// MixtureComp::RachfordRice2, src/MixtureComp.cpp:1907-1959.
void Mixture::RachfordRice2()
{
    const real* Ktmp = Ks[0];
    real Kmin = Ktmp[0];
    real Kmax = Ktmp[0];
    for (int i = 1; i < NC; i++) {
        if (Ktmp[i] < Kmin) Kmin = Ktmp[i];
        if (Ktmp[i] > Kmax) Kmax = Ktmp[i];
    }
    const real numin = 1 / (1 - Kmax);
    const real numax = 1 / (1 - Kmin);
    nu[0] = 0.5 * (numin + numax);
    real tmp, rj, J, dnuj, tmpnu;
    int iter = 0;
    const real tol = rrTol;
    const double maxIt = rrMaxIt;
    while (true) {
        rj = 0;
        J = 0;
        for (int i = 0; i < NC; i++) {
            tmp = 1 + nu[0] * (Ktmp[i] - 1);
            rj += zi[i] * (Ktmp[i] - 1) / tmp;
            J -= zi[i] * (Ktmp[i] - 1) * (Ktmp[i] - 1) / (tmp * tmp);
        }
        if (fabs(rj) < tol || iter > maxIt) break;
        dnuj = -rj / J;
        tmpnu = nu[0] + dnuj;
        if (tmpnu < numax && tmpnu > numin) {
            nu[0] = tmpnu;
        } else {
            if (dnuj > 0)
                nu[0] = (nu[0] + numax) / 2;
            else
                nu[0] = (nu[0] + numin) / 2;
        }
        iter++;
    }
    rrCurIt += iter;
    nu[1] = 1 - nu[0];
}

// MixtureComp::UpdateXRR, src/MixtureComp.cpp:2030-2043 (NP = 2).
void Mixture::UpdateXRR()
{
    real tmp;
    for (int i = 0; i < NC; i++) {
        tmp = 1;
        tmp += nu[0] * (Ks[0][i] - 1);
        x[1][i] = zi[i] / tmp;
        x[0][i] = Ks[0][i] * x[1][i];
    }
}

// MixtureComp::SplitNR, src/MixtureComp.cpp:2057-2156 (NP = 2).
void Mixture::SplitNR()
{
    nrSpConflag = false;
    const int len = NC;
    x2n();
    CalResSP();
    real eNR0;
    real eNR = Dnorm2(len, resSP);
    real NRtol = nrSpTol;
    real alpha;
    real en;
    int iter = 0;
    eNR0 = eNR;
    while (eNR > NRtol) {
        for (int j = 0; j < NP; j++)
            for (int i = 0; i < NC; i++) ln[j][i] = n[j][i];
        CalFugNAll(true);
        AssembleJmatSP();
        SymIndefSolve(len, JmatSP, resSP);
        alpha = CalStepNRsp();
        for (int i = 0; i < NC; i++) n[1][i] = zi[i];
        for (int i = 0; i < NC; i++) n[0][i] = n[0][i] + alpha * resSP[i];
        for (int i = 0; i < NC; i++) n[1][i] = n[1][i] - n[0][i];
        nu[0] = 0;
        for (int i = 0; i < NC; i++) nu[0] += n[0][i];
        for (int i = 0; i < NC; i++) x[0][i] = fabs(n[0][i] / nu[0]);
        nu[1] = 0;
        for (int i = 0; i < NC; i++) nu[1] += n[1][i];
        for (int i = 0; i < NC; i++) x[1][i] = fabs(n[1][i] / nu[1]);
        CalFugPhiAll();
        CalResSP();
        eNR = Dnorm2(len, resSP);
        iter++;
        if (eNR > eNR0 || iter > nrSpMaxIt) break;
        en = 0;
        for (int j = 0; j < NP; j++) {
            for (int i = 0; i < NC; i++) ln[j][i] = ln[j][i] - n[j][i];
            en += Dnorm2(NC, ln[j]);
        }
        if (en / (NP * NC) < 1E-8) {
            nrSpConflag = true;
            break;
        }
    }
    nrSpRealTol = eNR;
    if (eNR < NRtol) nrSpConflag = true;
    nrSpCurIt += iter;
}

// MixtureComp::CalResSP, src/MixtureComp.cpp:2158-2166.
void Mixture::CalResSP()
{
    for (int i = 0; i < NC; i++) resSP[i] = log(fug[NP - 1][i] / fug[0][i]);
}

// MixtureComp::CalFugNAll, src/MixtureComp.cpp:2168-2250.
void Mixture::CalFugNAll(bool Znflag)
{
    real C, E, G;
    real Cnk, Dnk, Enk, Gnk;
    real tmp, aik;
    for (int j = 0; j < NP; j++) {
        real* fugn = fugN[j];
        const real& aj = Aj[j];
        const real& bj = Bj[j];
        const real& zj = Zj[j];
        const real* xj = x[j];
        real* Znj = Zn[j];
        for (int i = 0; i < NC; i++) {
            tmp = 0;
            for (int m = 0; m < NC; m++) tmp += (1 - BIC[i * NC + m]) * sqrt(Ai[i] * Ai[m]) * xj[m];
            An[i] = 2 / nu[j] * (tmp - aj);
            Bn[i] = 1 / nu[j] * (Bi[i] - bj);
            if (Znflag) {
                Znj[i] = ((bj - zj) * An[i] +
                          ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                           ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                           (delta1 + delta2 - 1) * zj * zj) *
                              Bn[i]) /
                         (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                          (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
            }
        }
        G = (zj + delta1 * bj) / (zj + delta2 * bj);
        for (int i = 0; i < NC; i++) {
            C = xj[i] * P / (zj - bj);
            tmp = 0;
            for (int k = 0; k < NC; k++) tmp += (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]) * xj[k];
            E = -aj / ((delta1 - delta2) * bj) * (2 * tmp / aj - Bi[i] / bj);
            for (int k = 0; k <= i; k++) {
                aik = (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]);
                Cnk = P / (zj - bj) / (zj - bj) *
                      ((zj - bj) / nu[j] * (deltaij(i, k) - xj[i]) - xj[i] * (Znj[k] - Bn[k]));
                Dnk = Bi[i] / bj * (Znj[k] - (Bi[k] - bj) * (zj - 1) / (nu[j] * bj));
                Gnk = (delta1 - delta2) / ((zj + delta2 * bj) * (zj + delta2 * bj)) *
                      (Bn[k] * zj - Znj[k] * bj);
                Enk = -1 / (delta1 - delta2) / (bj * bj) *
                      (2 * (bj * aik / nu[j] + Bn[k] * (Bi[i] * aj / bj - tmp)) - An[k] * Bi[i] -
                       aj * Bi[i] / nu[j]);
                Enk -= E / nu[j];
                fugn[i * NC + k] = 1 / C * Cnk + Dnk + Enk * log(G) + E / G * Gnk;
                fugn[k * NC + i] = fugn[i * NC + k];
            }
        }
    }
}

// MixtureComp::AssembleJmatSP, src/MixtureComp.cpp:2265-2299 (NP = 2).
void Mixture::AssembleJmatSP()
{
    for (int i = 0; i < NC; i++)
        for (int k = 0; k < NC; k++) JmatSP[i * NC + k] = fugN[0][i * NC + k] + fugN[1][i * NC + k];
}

// MixtureComp::CalStepNRsp, src/MixtureComp.cpp:2301-2319 (NP = 2).
real Mixture::CalStepNRsp()
{
    real alpha = 1;
    real tmp;
    const real* nj0 = n[0];
    const real* r = resSP;
    for (int i = 0; i < NC; i++) {
        tmp = nj0[i] + alpha * r[i];
        if (tmp < 0) alpha *= 0.9 * fabs(nj0[i] / r[i]);
    }
    return alpha;
}

// MixtureComp::IdentifyPhase, src/MixtureComp.cpp:2352-2384.
void Mixture::IdentifyPhase()
{
    phaseExist[0] = false;
    phaseExist[1] = false;
    if (NP == 1) {
        real A = 0;
        real B = 0;
        for (int i = 0; i < NC; i++) {
            A += x[0][i] * Vc[i] * Tc[i];
            B += x[0][i] * Vc[i];
        }
        real Tcm = A / B;
        if (T > Tcm) {
            phaseLabel[0] = 1;
            phaseExist[1] = true;
        } else {
            phaseLabel[0] = 0;
            phaseExist[0] = true;
        }
    } else {
        if (MW[0] > MW[1]) {
            phaseLabel[0] = 0;
            phaseLabel[1] = 1;
        } else {
            phaseLabel[0] = 1;
            phaseLabel[1] = 0;
        }
        phaseExist[0] = true;
        phaseExist[1] = true;
    }
}

// MixtureComp::CopyPhase, src/MixtureComp.cpp:2386-2398.
void Mixture::CopyPhase()
{
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        vj[j1] = vC[j];
        for (int i = 0; i < NC; i++) xij[j1 * numCom + i] = x[j][i];
        mu[j1] = muC[j];
        xi[j1] = xiC[j];
        rho[j1] = rhoC[j];
        nj[j1] = nu[j];
    }
}

// MixtureComp::CalViscoLBC, src/MixtureComp.cpp:2402-2446 (polynomial branch).
void Mixture::CalViscoLBC()
{
    real tmp, Tri, xijT, xijP, xijV;
    for (int j = 0; j < NP; j++) {
        const real* xj = x[j];
        real* muA = muAux[j];
        for (int q = 0; q < 5; q++) muA[q] = 0;
        xijT = 0;
        xijP = 0;
        xijV = 0;
        for (int i = 0; i < NC; i++) {
            tmp = 5.4402 * pow(Tc[i], 1.0 / 6) / sqrt(MW[j]) / pow(Pc[i], 2.0 / 3);
            Tri = T / Tc[i];
            if (Tri <= 1.5) {
                tmp = 34 * 1E-5 * pow(Tri, 0.94) / tmp;
            } else {
                tmp = 17.78 * 1E-5 * pow((4.58 * Tri - 1.67), 0.625) / tmp;
            }
            muA[0] += xj[i] * sqrt(MWC[i]) * tmp;
            muA[1] += xj[i] * sqrt(MWC[i]);
            xijT += xj[i] * Tc[i];
            xijP += xj[i] * Pc[i];
            xijV += xj[i] * Vcvis[i];
        }
        muA[2] = 5.4402 * pow(xijT, 1.0 / 6) / sqrt(MW[j]) / pow(xijP, 2.0 / 3);
        muA[3] = xiC[j] * xijV;
        muA[4] = muA[3] * (muA[3] * (muA[3] * (LBCcoef[4] * muA[3] + LBCcoef[3]) + LBCcoef[2]) +
                           LBCcoef[1]) +
                 LBCcoef[0];
        muC[j] = muA[0] / muA[1] + 1E-4 * (pow(muA[4], 4) - 1) / muA[2];
    }
}

// MixtureComp::CalFugPAll, src/MixtureComp.cpp:2519-2579.
void Mixture::CalFugPAll(bool Zpflag)
{
    real C, E, G;
    real Cp, Dp, Gp;
    real tmp;
    for (int j = 0; j < NP; j++) {
        real* fugp = fugP[j];
        const real* xj = x[j];
        real& aj = Aj[j];
        real& bj = Bj[j];
        real& zj = Zj[j];
        real Ap = aj / P;
        real Bp = bj / P;
        if (Zpflag) {
            Zp[j] = ((bj - zj) * Ap +
                     ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                      ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                      (delta1 + delta2 - 1) * zj * zj) *
                         Bp) /
                    (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                     (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
        }
        G = (zj + delta1 * bj) / (zj + delta2 * bj);
        Gp = (delta1 - delta2) / ((zj + delta2 * bj) * (zj + delta2 * bj)) * (Bp * zj - Zp[j] * bj);
        for (int i = 0; i < NC; i++) {
            C = P / (zj - bj);
            tmp = 0;
            for (int m = 0; m < NC; m++) tmp += (1 - BIC[i * NC + m]) * sqrt(Ai[i] * Ai[m]) * xj[m];
            E = -aj / ((delta1 - delta2) * bj) * (2 * tmp / aj - Bi[i] / bj);
            Cp = ((zj - bj) - P * (Zp[j] - Bp)) / ((zj - bj) * (zj - bj));
            Dp = Bi[i] / bj * Zp[j];
            fugp[i] = Cp / C + Dp + E / G * Gp;
        }
    }
}

// MixtureComp::CalVfiVfp_full02 (+ AssembleMat/RhsVfiVfp_full02),
// src/MixtureComp.cpp:3911-4025.
void Mixture::CalVfiVfp_full02()
{
    real CgTP = GAS_CONSTANT * T / P;
    if (NP == 1) {
        const real& aj = Aj[0];
        const real& bj = Bj[0];
        const real& zj = Zj[0];
        const real* xj = x[0];
        real* Znij = Zn[0];
        real tmp;
        for (int i = 0; i < NC; i++) {
            tmp = 0;
            for (int m = 0; m < NC; m++) tmp += (1 - BIC[i * NC + m]) * sqrt(Ai[i] * Ai[m]) * xj[m];
            An[i] = 2 / nu[0] * (tmp - aj);
            Bn[i] = 1 / nu[0] * (Bi[i] - bj);
            Znij[i] = ((bj - zj) * An[i] +
                       ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                        ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                        (delta1 + delta2 - 1) * zj * zj) *
                           Bn[i]) /
                      (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                       (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
            vfi[i] = CgTP * (zj + nu[0] * Znij[i]) - Vshift[i];
        }
        Zp[0] = ((bj - zj) * aj +
                 ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                  ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                  (delta1 + delta2 - 1) * zj * zj) *
                     bj) /
                P /
                (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                 (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
        vfP = CgTP * nu[0] * (Zp[0] - zj / P);
    } else {
        CalFugNAll(true);
        CalFugPAll(true);
        // JmatDer column-sorted (column-major), rhsDer NC x (NC+1) column-major.
        for (int i = 0; i < NC; i++)
            for (int m = 0; m < NC; m++) JmatDer[i * NC + m] = fugN[0][i * NC + m] + fugN[1][i * NC + m];
        for (int i = 0; i < NC; i++) rhsDer[i] = fugP[1][i] - fugP[0][i];
        for (int k = 0; k < NC; k++)
            for (int i = 0; i < NC; i++) rhsDer[(k + 1) * NC + i] = fugN[1][k * NC + i];
        LUSolveColMajor(NC, NC + 1, JmatDer, rhsDer);
        const real* dnkjdNP = rhsDer;
        const int j0 = phaseLabel[0];
        const int j1 = phaseLabel[1];
        vjp[j0] = CgTP * nu[0] * (Zp[0] - Zj[0] / P);
        vjp[j1] = CgTP * nu[1] * (Zp[1] - Zj[1] / P);
        for (int k = 0; k < NC; k++) {
            vjp[j0] += (CgTP * (Zj[0] + nu[0] * Zn[0][k]) - Vshift[k]) * dnkjdNP[k];
            vjp[j1] -= (CgTP * (Zj[1] + nu[1] * Zn[1][k]) - Vshift[k]) * dnkjdNP[k];
        }
        vfP = vjp[j0] + vjp[j1];
        dnkjdNP += NC;
        for (int i = 0; i < NC; i++) {
            vfi[i] = 0;
            vji[j0][i] = 0;
            vji[j1][i] = 0;
            for (int k = 0; k < NC; k++) {
                vji[j0][i] += (CgTP * (Zj[0] + nu[0] * Zn[0][k]) - Vshift[k]) * dnkjdNP[k];
                vji[j1][i] += (CgTP * (Zj[1] + nu[1] * Zn[1][k]) - Vshift[k]) *
                              (deltaij(i, k) - dnkjdNP[k]);
            }
            vfi[i] = vji[j0][i] + vji[j1][i];
            dnkjdNP += NC;
        }
    }
}

// MixtureComp::CaldXsdXpAPI02p, src/MixtureComp.cpp:4386-4493.
void Mixture::CaldXsdXpAPI02p()
{
    const int nrowMax = NPH + NPH * numCom;
    const int ncol = numCom + 1;
    for (int k = 0; k < nrowMax * ncol; k++) dXsdXp[k] = 0;
    real vf2 = vf * vf;
    real vwp = vjp[numPhase - 1];
    real vw = vj[numPhase - 1];
    if (NP == 1) {
        const int j1 = phaseLabel[0];
        real* bId = dXsdXp;
        bId[0] = ((vfP - vwp) * vf - vfP * vj[j1]) / vf2;
        bId++;
        for (int m = 0; m < NC; m++) bId[m] = vfi[m] * vw / vf2;
        bId += NC;
        bId[0] = -vfi[numCom - 1] * vj[j1] / vf2;
        bId++;
        for (int m = 0; m < ncol; m++) bId[m] = -dXsdXp[m];
        bId += ncol + 1;
        for (int i = 0; i < NC; i++) {
            for (int m = 0; m < NC; m++) bId[m] = (deltaij(i, m) * Nh - Ni[i]) / (Nh * Nh);
            bId += ncol;
        }
    } else {
        real* bId;
        for (int j = 0; j < 2; j++) {
            const int j1 = phaseLabel[j];
            bId = &dXsdXp[j1 * ncol];
            bId[0] = (vjp[j1] * vf - vfP * vj[j1]) / vf2;
            bId++;
            for (int m = 0; m < numCom; m++) bId[m] = (vji[j1][m] * vf - vfi[m] * vj[j1]) / vf2;
        }
        bId = &dXsdXp[(numPhase - 1) * ncol];
        bId[0] = (vwp * vf - vfP * vw) / vf2;
        bId++;
        for (int m = 0; m < numCom; m++) bId[m] = (vji[(numPhase - 1)][m] * vf - vfi[m] * vw) / vf2;
        bId += numCom;
        const int j0 = phaseLabel[0];
        const int j1 = phaseLabel[1];
        const real* dnkjdNP = rhsDer;
        real* bId1 = bId + j0 * NC * ncol;
        real* bId2 = bId + j1 * NC * ncol;
        real njDerSum = 0;
        for (int i = 0; i < NC; i++) njDerSum += dnkjdNP[i];
        for (int i = 0; i < NC; i++) {
            bId1[0] = (dnkjdNP[i] * nu[0] - njDerSum * Nh * n[0][i]) / (nu[0] * nu[0]);
            bId2[0] = (-dnkjdNP[i] * nu[1] + njDerSum * Nh * n[1][i]) / (nu[1] * nu[1]);
            bId1 += ncol;
            bId2 += ncol;
        }
        dnkjdNP += NC;
        for (int m = 0; m < NC; m++) {
            njDerSum = 0;
            for (int i = 0; i < NC; i++) njDerSum += dnkjdNP[i];
            bId1 = bId + j0 * NC * ncol + m + 1;
            bId2 = bId + j1 * NC * ncol + m + 1;
            for (int i = 0; i < NC; i++) {
                bId1[0] = (dnkjdNP[i] * nu[0] - njDerSum * Nh * n[0][i]) / (nu[0] * nu[0]);
                bId2[0] = ((deltaij(i, m) - dnkjdNP[i]) * nu[1] - (1 - njDerSum) * Nh * n[1][i]) /
                          (nu[1] * nu[1]);
                bId1 += ncol;
                bId2 += ncol;
            }
            dnkjdNP += NC;
        }
    }
}

// MixtureComp::CalXiPNX_partial, src/MixtureComp.cpp:3037-3092.
void Mixture::CalXiPNX_partial()
{
    real CgTP = GAS_CONSTANT * T / P;
    real tmp;
    for (int j = 0; j < NP; j++) {
        const real* xj = x[j];
        real aj = Aj[j];
        real bj = Bj[j];
        real zj = Zj[j];
        for (int i = 0; i < NC; i++) Bx[i] = Bi[i];
        for (int i = 0; i < NC; i++) {
            tmp = 0;
            for (int k = 0; k < NC; k++) tmp += xj[k] * (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]);
            Ax[i] = 2 * tmp;
            Zx[i] = ((bj - zj) * Ax[i] +
                     ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                      ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                      (delta1 + delta2 - 1) * zj * zj) *
                         Bx[i]) /
                    (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                     (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
        }
        tmp = -xiC[j] * xiC[j];
        for (int i = 0; i < NC; i++) xixC[j * NC + i] = tmp * (CgTP * Zx[i] - Vshift[i]);
        xiPC[j] = tmp * CgTP * (Zp[j] - Zj[j] / P);
    }
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        xiP[j1] = xiPC[j];
        for (int i = 0; i < NC; i++) xix[j1 * numCom + i] = xixC[j * NC + i];
    }
}

// MixtureComp::CalRhoPX_partial, src/MixtureComp.cpp:3094-3114.
void Mixture::CalRhoPX_partial()
{
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        rhoP[j1] = xiP[j1] * MW[j];
        for (int i = 0; i < NC; i++)
            rhox[j1 * numCom + i] = xix[j1 * numCom + i] * MW[j] + xi[j1] * MWC[i];
    }
}

// MixtureComp::CalMuPXLBC_partial, src/MixtureComp.cpp:3125-3213.  The
// reference's der1IJ = -(1 / 2) * ... is an integer zero, kept as such.
void Mixture::CalMuPXLBC_partial()
{
    real val1IJ, val2IJ;
    real der1IJ, der2IJ, der3J, der4J, der6J, der7J, der8J;
    real Tri, tmp;
    real xTj, xPj, xVj;
    real derxTj, derxPj, derMWj;
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        const real* xj = x[j];
        const real* muAuxj = muAux[j];
        xTj = xPj = xVj = 0;
        for (int i = 0; i < NC; i++) {
            xTj += xj[i] * Tc[i];
            xPj += xj[i] * Pc[i];
            xVj += xj[i] * Vcvis[i];
        }
        der7J = xVj * xiP[j1];
        der8J = der7J * (LBCcoef[1] +
                         muAuxj[3] * (2 * LBCcoef[2] +
                                      muAuxj[3] * (3 * LBCcoef[3] + muAuxj[3] * 4 * LBCcoef[4])));
        muP[j1] = (4 * 1E-4) * pow(muAuxj[4], 3) * der8J / muAuxj[2];

        const int bId = numCom * j1;
        for (int k = 0; k < NC; k++) {
            derxTj = Tc[k];
            derxPj = Pc[k];
            derMWj = MWC[k];
            der3J = 0;
            for (int i = 0; i < NC; i++) {
                val1IJ = muAux1I[i] / sqrt(MW[j]);
#ifdef OCP_ORACLE_EXACT_LBC
                // test-only variant: the derivative the reference meant to take
                der1IJ = -(1.0 / 2) * muAux1I[i] * pow(MW[j], -1.5) * derMWj;
#else
                der1IJ = -(1 / 2) * muAux1I[i] * pow(MW[j], -1.5) * derMWj;
#endif
                Tri = T / Tc[i];
                if (Tri <= 1.5) {
                    tmp = 34 * 1E-5 * pow(Tri, 0.94);
                } else {
                    tmp = 17.78 * 1E-5 * pow(4.58 * Tri - 1.67, 0.625);
                }
                val2IJ = tmp / val1IJ;
                der2IJ = -tmp * der1IJ / (val1IJ * val1IJ);
                der3J += xj[i] * sqrtMWi[i] * der2IJ + deltaij(i, k) * sqrtMWi[k] * val2IJ;
            }
            der4J = sqrtMWi[k];
            der6J = 5.4402 *
                    (1.0 / 6 * pow(xTj, -5.0 / 6) * derxTj -
                     pow(xTj, 1.0 / 6) * (0.5 / MW[j] * derMWj + 2.0 / 3 / xPj * derxPj)) /
                    (sqrt(MW[j]) * pow(xPj, 2.0 / 3));
            der7J = xix[j1 * numCom + k] * xVj + xi[j1] * Vcvis[k];
            der8J = der7J * (LBCcoef[1] +
                             muAuxj[3] * (2 * LBCcoef[2] +
                                          muAuxj[3] * (3 * LBCcoef[3] + muAuxj[3] * 4 * LBCcoef[4])));
            mux[bId + k] =
                (der3J * muAuxj[1] - muAuxj[0] * der4J) / (muAuxj[1] * muAuxj[1]) +
                1E-4 * (4 * pow(muAuxj[4], 3) * der8J * muAuxj[2] - (pow(muAuxj[4], 4) - 1) * der6J) /
                    (muAuxj[2] * muAuxj[2]);
        }
    }
}

// MixtureComp::CalPhiNSkip, src/MixtureComp.cpp:5155-5217.
void Mixture::CalPhiNSkip()
{
    real C, E, G;
    real Cnk, Dnk, Enk, Gnk;
    real tmp, aik;
    const real& aj = Aj[0];
    const real& bj = Bj[0];
    const real& zj = Zj[0];
    const real* xj = x[0];
    real* Znj = Zn[0];
    for (int i = 0; i < NC; i++) {
        tmp = 0;
        for (int m = 0; m < NC; m++) tmp += (1 - BIC[i * NC + m]) * sqrt(Ai[i] * Ai[m]) * xj[m];
        An[i] = 2 / nu[0] * (tmp - aj);
        Bn[i] = 1 / nu[0] * (Bi[i] - bj);
        Znj[i] = ((bj - zj) * An[i] +
                  ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                   ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                   (delta1 + delta2 - 1) * zj * zj) *
                      Bn[i]) /
                 (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                  (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
    }
    G = (zj + delta1 * bj) / (zj + delta2 * bj);
    for (int i = 0; i < NC; i++) {
        C = 1 / (zj - bj);
        tmp = 0;
        for (int k = 0; k < NC; k++) tmp += (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]) * xj[k];
        E = -aj / ((delta1 - delta2) * bj) * (2 * tmp / aj - Bi[i] / bj);
        for (int k = 0; k <= i; k++) {
            aik = (1 - BIC[i * NC + k]) * sqrt(Ai[i] * Ai[k]);
            Cnk = (Bn[k] - Znj[k]) / ((zj - bj) * (zj - bj));
            Dnk = Bi[i] / bj * (Znj[k] - (Bi[k] - bj) * (zj - 1) / (nu[0] * bj));
            Gnk = (delta1 - delta2) / ((zj + delta2 * bj) * (zj + delta2 * bj)) * (Bn[k] * zj - Znj[k] * bj);
            Enk = -1 / (delta1 - delta2) / (bj * bj) *
                  (2 * (bj * aik / nu[0] + Bn[k] * (Bi[i] * aj / bj - tmp)) - An[k] * Bi[i] -
                   aj * Bi[i] / nu[0]);
            Enk -= E / nu[0];
            phiN[i * NC + k] = 1 / C * Cnk + Dnk + Enk * log(G) + E / G * Gnk;
            phiN[k * NC + i] = phiN[i * NC + k];
        }
    }
}

// MixtureComp::CalSkipForNextStep (+ AssembleSkipMatSTA), src/MixtureComp.cpp:5219-5254.
void Mixture::CalSkipForNextStep(SkipState& skip)
{
    if (flagSkip && ftype == 0) {
        CalPhiNSkip();
        float mat[MAXNC * MAXNC];
        const real* xj = x[0];
        for (int i = 0; i < NC; i++) {
            for (int j = 0; j <= i; j++) {
                mat[i * NC + j] = tofloat(deltaij(i, j) + nu[0] * sqrt(xj[i] * xj[j]) * phiN[i * NC + j]);
                mat[j * NC + i] = mat[i * NC + j];
            }
        }
        const float lam = MinEigenSym(NC, mat);
        skip.minEigen = lam;
        skip.P = todbl(P);
        skip.T = todbl(T);
        for (int i = 0; i < NC; i++) skip.zi[i] = todbl(zi[i]);
    }
    skip.flag = flagSkip;
}

// ---------------------------------------------------------------------------
// LAPACK restatements.
// ---------------------------------------------------------------------------

static inline int iamax(int n, const real* x, int incx)
{
    int best = 0;
    real bmax = fabs(x[0]);
    for (int i = 1; i < n; i++) {
        real v = fabs(x[i * incx]);
        if (v > bmax) {
            bmax = v;
            best = i;
        }
    }
    return best;
}

#ifdef OCP_ORACLE_MKL
// Test-only variant: the LAPACK calls go to the library the reference links
// (MKL), to tell the restated algorithms' rounding from the rest.
extern "C" void dsysv_(const char*, const int*, const int*, double*, const int*, int*, double*, const int*,
                       double*, const int*, int*);
extern "C" void dgesv_(const int*, const int*, double*, const int*, int*, double*, const int*, int*);

int SymIndefSolve(int n, real* a, real* b)
{
    const char uplo = 'U';
    const int one = 1, lwork = MAXNC * (NPMAX - 1);
    int ipiv[MAXNC], info = 0;
    double work[MAXNC * (NPMAX - 1)];
    dsysv_(&uplo, &n, &one, a, &n, ipiv, b, &n, work, &lwork, &info);
    return info;
}

int LUSolveColMajor(int n, int nrhs, real* a, real* b)
{
    int ipiv[NPH + NPMAX * MAXNC], info = 0;
    dgesv_(&n, &nrhs, a, &n, ipiv, b, &n, &info);
    return info;
}
#else
// dsysv with UPLO='U': dsytf2 (Bunch-Kaufman, reference LAPACK) then dsytrs.
// a is column-major: A(r,c) = a[r + c*n]; 0-based indices throughout.
int SymIndefSolve(int n, real* a, real* b)
{
#define A_(r, c) a[(r) + (c)*n]
    int ipiv[MAXNC];
    const real alpha = (1 + sqrt(real(17.0))) / 8;
    int info = 0;
    int k = n - 1;
    while (k >= 0) {
        int kstep = 1;
        real absakk = fabs(A_(k, k));
        int imax = 0;
        real colmax = 0;
        if (k > 0) {
            imax = iamax(k, &A_(0, k), 1);
            colmax = fabs(A_(imax, k));
        }
        int kp;
        if (std::max(absakk, colmax) == 0) {
            if (info == 0) info = k + 1;
            kp = k;
        } else {
            if (absakk >= alpha * colmax) {
                kp = k;
            } else {
                // row imax: columns imax+1..k
                int jmax = imax + 1 + iamax(k - imax, &A_(imax, imax + 1), n);
                real rowmax = fabs(A_(imax, jmax));
                if (imax > 0) {
                    jmax = iamax(imax, &A_(0, imax), 1);
                    rowmax = std::max(rowmax, fabs(A_(jmax, imax)));
                }
                if (absakk >= alpha * colmax * (colmax / rowmax)) {
                    kp = k;
                } else if (fabs(A_(imax, imax)) >= alpha * rowmax) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            const int kk = k - kstep + 1;
            if (kp != kk) {
                for (int i = 0; i < kp; i++) std::swap(A_(i, kk), A_(i, kp));
                for (int j = kp + 1; j < kk; j++) std::swap(A_(j, kk), A_(kp, j));
                std::swap(A_(kk, kk), A_(kp, kp));
                if (kstep == 2) std::swap(A_(k - 1, k), A_(kp, k));
            }
            if (kstep == 1) {
                // dsyr upper on A(0:k-1, 0:k-1) with x = A(0:k-1, k), alpha = -1/A(k,k)
                real r1 = 1 / A_(k, k);
                const real malpha = -r1;
                for (int j = 0; j < k; j++) {
                    if (A_(j, k) != 0) {
                        real temp = malpha * A_(j, k);
                        for (int i = 0; i <= j; i++) A_(i, j) = A_(i, j) + A_(i, k) * temp;
                    }
                }
                for (int i = 0; i < k; i++) A_(i, k) = r1 * A_(i, k);
            } else {
                if (k > 1) {
                    real d12 = A_(k - 1, k);
                    real d22 = A_(k - 1, k - 1) / d12;
                    real d11 = A_(k, k) / d12;
                    real t = 1 / (d11 * d22 - 1);
                    d12 = t / d12;
                    for (int j = k - 2; j >= 0; j--) {
                        real wkm1 = d12 * (d11 * A_(j, k - 1) - A_(j, k));
                        real wk = d12 * (d22 * A_(j, k) - A_(j, k - 1));
                        for (int i = j; i >= 0; i--)
                            A_(i, j) = A_(i, j) - A_(i, k) * wk - A_(i, k - 1) * wkm1;
                        A_(j, k) = wk;
                        A_(j, k - 1) = wkm1;
                    }
                }
            }
        }
        if (kstep == 1) {
            ipiv[k] = kp + 1;
        } else {
            ipiv[k] = -(kp + 1);
            ipiv[k - 1] = -(kp + 1);
        }
        k -= kstep;
    }
    if (info != 0) return info;
    // dsytrs, UPLO='U', one right-hand side.
    k = n - 1;
    while (k >= 0) {
        if (ipiv[k] > 0) {
            int kp = ipiv[k] - 1;
            if (kp != k) std::swap(b[k], b[kp]);
            const real bk = -b[k];
            if (bk != 0)
                for (int i = 0; i < k; i++) b[i] = b[i] + A_(i, k) * bk;
            b[k] = (1 / A_(k, k)) * b[k];
            k -= 1;
        } else {
            int kp = -ipiv[k] - 1;
            if (kp != k - 1) std::swap(b[k - 1], b[kp]);
            real t = -b[k];
            if (t != 0)
                for (int i = 0; i < k - 1; i++) b[i] = b[i] + A_(i, k) * t;
            t = -b[k - 1];
            if (t != 0)
                for (int i = 0; i < k - 1; i++) b[i] = b[i] + A_(i, k - 1) * t;
            real akm1k = A_(k - 1, k);
            real akm1 = A_(k - 1, k - 1) / akm1k;
            real ak = A_(k, k) / akm1k;
            real denom = akm1 * ak - 1;
            real bkm1 = b[k - 1] / akm1k;
            real bk = b[k] / akm1k;
            b[k - 1] = (ak * bkm1 - bk) / denom;
            b[k] = (akm1 * bk - bkm1) / denom;
            k -= 2;
        }
    }
    k = 0;
    while (k < n) {
        if (ipiv[k] > 0) {
            real s = 0;
            for (int i = 0; i < k; i++) s += A_(i, k) * b[i];
            b[k] = b[k] - s;
            int kp = ipiv[k] - 1;
            if (kp != k) std::swap(b[k], b[kp]);
            k += 1;
        } else {
            real s = 0;
            for (int i = 0; i < k; i++) s += A_(i, k) * b[i];
            b[k] = b[k] - s;
            s = 0;
            for (int i = 0; i < k; i++) s += A_(i, k + 1) * b[i];
            b[k + 1] = b[k + 1] - s;
            int kp = -ipiv[k] - 1;
            if (kp != k) std::swap(b[k], b[kp]);
            k += 2;
        }
    }
    return 0;
#undef A_
}

// dgesv: dgetf2 (partial pivoting) then dgetrs, column-major, nrhs columns.
int LUSolveColMajor(int n, int nrhs, real* a, real* b)
{
#define A_(r, c) a[(r) + (c)*n]
    int ipiv[NPH + NPMAX * MAXNC];
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j + iamax(n - j, &A_(j, j), 1);
        ipiv[j] = p;
        if (A_(p, j) != 0) {
            if (p != j)
                for (int c = 0; c < n; c++) std::swap(A_(j, c), A_(p, c));
            const real r = 1 / A_(j, j);
            for (int i = j + 1; i < n; i++) A_(i, j) = r * A_(i, j);
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; c++) {
            const real t = A_(j, c);
            if (t != 0)
                for (int i = j + 1; i < n; i++) A_(i, c) = A_(i, c) - A_(i, j) * t;
        }
    }
    if (info != 0) return info;
    for (int c = 0; c < nrhs; c++) {
        real* bc = b + c * n;
        for (int j = 0; j < n; j++)
            if (ipiv[j] != j) std::swap(bc[j], bc[ipiv[j]]);
        for (int j = 0; j < n; j++) {
            const real t = bc[j];
            if (t != 0)
                for (int i = j + 1; i < n; i++) bc[i] = bc[i] - t * A_(i, j);
        }
        for (int j = n - 1; j >= 0; j--) {
            if (bc[j] != 0) {
                bc[j] = bc[j] / A_(j, j);
                const real t = bc[j];
                for (int i = 0; i < j; i++) bc[i] = bc[i] - t * A_(i, j);
            }
        }
    }
    return 0;
#undef A_
}
#endif  // OCP_ORACLE_MKL

float MinEigenSym(int n, const float* af)
{
    double a[MAXNC * MAXNC];
    for (int i = 0; i < n * n; i++) a[i] = af[i];
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                const double apq = a[p * n + q];
                if (apq == 0) continue;
                const double app = a[p * n + p];
                const double aqq = a[q * n + q];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1);
                const double s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = a[k * n + p];
                    const double akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = a[p * n + k];
                    const double aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
            }
        }
    }
    double m = a[0];
    for (int i = 1; i < n; i++) m = std::min(m, a[i * n + i]);
    return static_cast<float>(m);
}

}  // namespace oracle
