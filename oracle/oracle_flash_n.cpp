// CPU restatement of MixtureComp's IsoT_FIMn flash (secondary variables S
// and n_ij): FlashFIMn / InitFlashFIMn and the derivatives they call.
//
// TEST INFRASTRUCTURE ONLY: the oracle the HIP kernels are checked against.
// Nothing in the product path may include, link or call it.
//
// Restates src/MixtureComp.cpp:355-442 (InitFlashFIMn), :575-695
// (FlashFIMn), :2581-2757 (CalVjpVfpVfn_partial, CalXiPn_partial,
// CalRhoPn_partial, CalMuPnLBC_partial), :4495-4632 (CaldXsdXpAPI03),
// :4634-4820 (CaldXsdXp03) and :4822-4852 (CalVfiVfp_full03), in the
// reference's arithmetic order.  CaldXsdXp03's dgesv is the restated
// dgetf2 + dgetrs (LUSolveColMajor).
#include "oracle_flash.hpp"

#include <algorithm>

namespace oracle {

static inline real deltaN(int i, int j) { return i == j ? 1.0 : 0.0; }

// MixtureComp::FlashFIMn, src/MixtureComp.cpp:575-695.
void Mixture::FlashFIMn(double Pin, double Tin, const double* Niin, const double* Sjin, int lastNP,
                        const double* xijin, SkipState& skip)
{
    for (auto& it : iters) it = 0;
    lNP = lastNP > 0 ? lastNP - 1 : 0;
    if (lNP == 2) {
        for (int i = 0; i < NC; i++) lKs[i] = xijin[i] / xijin[i + numCom];
    }
    InitPTN(Pin, Tin + CONV5, Niin);
    CalFtypeFIM(Sjin, skip);
    CalFlash();
    CalVjpVfpVfn_partial();
    WaterProperties(false, Sjin, 0);
    CalSaturation();
    CaldXsdXpAPI03();
    CalXiPn_partial();
    CalRhoPn_partial();
    CalMuPnLBC_partial();
    CalVfiVfp_full03();
    SetDerivFlags();
    CalSkipForNextStep(skip);
}

// MixtureComp::InitFlashFIMn, src/MixtureComp.cpp:355-442.
void Mixture::InitFlashFIMn(double Pin, double Tin, const double* Sjin, double Vpore, const double* Ziin,
                            SkipState& skip)
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
    for (int i = 0; i < NC; i++) Ni[i] = zi[i] * Nh;
    CalVjpVfpVfn_partial();
    WaterProperties(true, Sjin, Vpore);
    CalSaturation();
    CaldXsdXpAPI03();
    CalXiPn_partial();
    CalRhoPn_partial();
    CalMuPnLBC_partial();
    CalVfiVfp_full03();
    SetDerivFlags();
    CalSkipForNextStep(skip);
}

// MixtureComp::CalVjpVfpVfn_partial, src/MixtureComp.cpp:2581-2626: the
// phase volumes' derivatives at fixed phase moles (vji: by n_ij, vjp: by P);
// vfi is cleared and vfP holds the hydrocarbon sum.
void Mixture::CalVjpVfpVfn_partial()
{
    real CgTP = GAS_CONSTANT * T / P;
    for (int i = 0; i < numCom; i++) vfi[i] = 0;
    vfP = 0;
    for (int j = 0; j < NP; j++) {
        const real& aj = Aj[j];
        const real& bj = Bj[j];
        const real& zj = Zj[j];
        const real* xj = x[j];
        real* Znj = Zn[j];
        real tmp;
        const int j1 = phaseLabel[j];
        for (int i = 0; i < NC; i++) {
            tmp = 0;
            for (int m = 0; m < NC; m++) tmp += (1 - BIC[i * NC + m]) * sqrt(Ai[i] * Ai[m]) * xj[m];
            An[i] = 2 / nu[j] * (tmp - aj);
            Bn[i] = 1 / nu[j] * (Bi[i] - bj);
            Znj[i] = ((bj - zj) * An[i] +
                      ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                       ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                       (delta1 + delta2 - 1) * zj * zj) *
                          Bn[i]) /
                     (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                      (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
            vji[j1][i] = CgTP * (zj + nu[j] * Znj[i]) - Vshift[i];
        }
        Zp[j] = ((bj - zj) * aj +
                 ((aj + delta1 * delta2 * (3 * bj * bj + 2 * bj)) +
                  ((delta1 + delta2) * (2 * bj + 1) - 2 * delta1 * delta2 * bj) * zj -
                  (delta1 + delta2 - 1) * zj * zj) *
                     bj) /
                P /
                (3 * zj * zj + 2 * ((delta1 + delta2 - 1) * bj - 1) * zj +
                 (aj + delta1 * delta2 * bj * bj - (delta1 + delta2) * bj * (bj + 1)));
        vjp[j1] = CgTP * nu[j] * (Zp[j] - zj / P);
        vfP += vjp[j1];
    }
}

// MixtureComp::CaldXsdXpAPI03, src/MixtureComp.cpp:4495-4632: compact
// d(S, n_ij) / d(P, N) rows and the secondary variables' correction res.
void Mixture::CaldXsdXpAPI03()
{
    const int nrowMax = NPH + NPH * numCom;
    const int ncol = numCom + 1;
    for (int k = 0; k < nrowMax * ncol; k++) dXsdXp[k] = 0;
    for (int k = 0; k < NPH * (MAXNC + 2); k++) resN[k] = 0;
    resPc = 0;
    const real vf2 = vf * vf;
    const real vwp = vjp[numPhase - 1];
    const real vw = vj[numPhase - 1];
    sysDim = 0;
    if (NP == 1) {
        const int j0 = phaseLabel[0];
        real* bId = dXsdXp;
        bId[0] = ((vfP - vwp) * vf - vfP * vj[j0]) / vf2;
        bId++;
        for (int m = 0; m < NC; m++) bId[m] = vji[j0][m] * vw / vf2;
        bId += NC;
        bId[0] = -vfi[numCom - 1] * vj[j0] / vf2;
        bId++;
        for (int m = 0; m < ncol; m++) bId[m] = -dXsdXp[m];
        bId = &dXsdXp[2 * ncol + 1];
        for (int i = 0; i < NC; i++) {
            bId[i] = 1;
            bId += ncol;
        }
    } else {
        CalFugNAll(false);
        CalFugPAll(false);
        CaldXsdXp03();
        int sIndex[NPH] = {0, 0, 0};
        for (int j = 0; j < numPhase - 1; j++) {
            if (phaseExist[j]) {
                for (int j1 = j + 1; j1 < numPhase; j1++) sIndex[j1]++;
            }
        }
        const real* STmp = rhsDerN;
        const int nrhs = numCom + 1;
        const int wNP = NP + 1;
        int bId;
        for (int i = 0; i < nrhs; i++) {
            for (int j = 0; j < NP; j++) dXsdXp[sIndex[phaseLabel[j]] * nrhs + i] = STmp[j];
            STmp += NP;
            dXsdXp[NP * nrhs + i] = STmp[0];
            STmp++;
            for (int j = 0; j < NP; j++) {
                bId = wNP + sIndex[phaseLabel[j]] * NC;
                for (int k = 0; k < NC; k++) dXsdXp[(bId + k) * nrhs + i] = STmp[k];
                STmp += NC;
            }
        }
        for (int j = 0; j < NP; j++) resN[sIndex[phaseLabel[j]]] = STmp[j];
        STmp += NP;
        resN[NP] = STmp[0];
        STmp++;
        for (int j = 0; j < NP; j++) {
            bId = wNP + sIndex[phaseLabel[j]] * NC;
            for (int k = 0; k < NC; k++) resN[bId + k] = STmp[k];
            STmp += NC;
        }
        // The reference pairs the j-th stored block of n_ij corrections
        // (stored in oil-then-gas order) with vji of internal phase j's
        // label: when internal phase 0 is the gas phase, each phase's
        // volume derivatives multiply the other phase's corrections.
        const real* myres = &resN[NP + 1];
        for (int j = 0; j < NP; j++) {
            const int j1 = phaseLabel[j];
            for (int i = 0; i < NC; i++) resPc += vji[j1][i] * myres[i];
            myres += NC;
        }
    }
}

// MixtureComp::CaldXsdXp03, src/MixtureComp.cpp:4634-4820: the flash's
// equations F (fugacity equality, component balance, phase volumes) are
// linearised in s = (S_j, S_w, n_ij); -dF/ds X = [dF/dp, F] is solved by
// dgesv for the NC + 2 primary columns and the residual column.
void Mixture::CaldXsdXp03()
{
    const int dim = NP + 1 + NP * NC;
    for (int k = 0; k < dim * dim; k++) JmatTmpN[k] = 0;
    real* MTmp = JmatTmpN;
    const real vf2 = vf * vf;
    // -dFf / ds
    for (int j = 0; j < NP - 1; j++) {
        const real* fugNJ = fugN[j];
        const real* fugNNP = fugN[NP - 1];
        for (int i = 0; i < NC; i++) {
            MTmp += (NP + 1);
            MTmp += j * NC;
            for (int k = 0; k < NC; k++) MTmp[k] = -fugNJ[k];
            fugNJ += NC;
            MTmp += (NP - 1 - j) * NC;
            for (int k = 0; k < NC; k++) MTmp[k] = fugNNP[k];
            fugNNP += NC;
            MTmp += NC;
        }
    }
    // -dFn / ds
    for (int i = 0; i < NC; i++) {
        MTmp += (NP + 1);
        for (int j = 0; j < NP; j++) {
            MTmp[i] = 1;
            MTmp += NC;
        }
    }
    // -dFs / ds, hydrocarbon phases
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        MTmp[j] = -1;
        MTmp += (NP + 1);
        for (int m = 0; m < NP; m++) {
            const int m1 = phaseLabel[m];
            if (m1 == j1) {
                for (int k = 0; k < NC; k++) MTmp[k] = vji[j1][k] * (vf - vj[j1]) / vf2;
            } else {
                for (int k = 0; k < NC; k++) MTmp[k] = -vji[m1][k] * vj[j1] / vf2;
            }
            MTmp += NC;
        }
    }
    // -dFsw / ds
    MTmp[NP] = -1;
    MTmp += (NP + 1);
    for (int m = 0; m < NP; m++) {
        const int m1 = phaseLabel[m];
        for (int k = 0; k < NC; k++) MTmp[k] = -vji[m1][k] * vj[numPhase - 1] / vf2;
        MTmp += NC;
    }
    // column-major copy for dgesv
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) JmatDerN[i * dim + j] = JmatTmpN[j * dim + i];

    for (int k = 0; k < dim * dim; k++) JmatTmpN[k] = 0;
    MTmp = JmatTmpN;
    const int ncol2 = numCom + 1;
    // dFf / dp
    for (int j = 0; j < NP - 1; j++) {
        for (int i = 0; i < NC; i++) {
            MTmp[0] = fugP[j][i] - fugP[NP - 1][i];
            MTmp += ncol2;
        }
    }
    // dFn / dp
    for (int i = 0; i < NC; i++) {
        MTmp++;
        MTmp[i] = 1;
        MTmp += numCom;
    }
    // dFs / dp, hydrocarbon phases
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        MTmp[0] = (vfP * vj[j1] - vjp[j1] * vf) / vf2;
        MTmp += numCom;
        MTmp[0] = vfi[numCom - 1] * vj[j1] / vf2;
        MTmp++;
    }
    // dFsw / dp
    MTmp[0] = (vfP * vj[numPhase - 1] - vjp[numPhase - 1] * vf) / vf2;
    MTmp += numCom;
    MTmp[0] = vfi[numCom - 1] * (vj[numPhase - 1] - vf) / vf2;
    MTmp++;

    const int nrhs = numCom + 1;
    const int nrow = NP * NC + NP + 1;
    for (int i = 0; i < nrhs; i++)
        for (int j = 0; j < nrow; j++) rhsDerN[i * nrow + j] = JmatTmpN[j * nrhs + i];

    // the residual column: resF, resN, resS
    real* myRes = &rhsDerN[nrhs * nrow];
    for (int j = 0; j < NP - 1; j++) {
        const real* fugJ = fug[j];
        const real* fugNP = fug[NP - 1];
        for (int i = 0; i < NC; i++) myRes[i] = fugJ[i] - fugNP[i];
        myRes += NC;
    }
    for (int i = 0; i < NC; i++) {
        myRes[i] = Ni[i];
        for (int j = 0; j < NP; j++) myRes[i] -= x[j][i] * nu[j];
    }
    myRes += NC;
    for (int j = 0; j < NP; j++) myRes[j] = S[phaseLabel[j]] - vj[phaseLabel[j]] / vf;
    myRes[NP] = S[numPhase - 1] - vj[numPhase - 1] / vf;

    sysDim = nrow;
    for (int k = 0; k < nrow * nrow; k++) sysM[k] = JmatDerN[k];
    for (int k = 0; k < nrow * (numCom + 2); k++) sysB[k] = rhsDerN[k];
    LUSolveColMajor(nrow, numCom + 1 + 1, JmatDerN, rhsDerN);
}

// MixtureComp::CalXiPn_partial, src/MixtureComp.cpp:2628-2646.
void Mixture::CalXiPn_partial()
{
    real CgTP = GAS_CONSTANT * T / P;
    real tmp;
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        tmp = -xi[j1] * xi[j1] * CgTP;
        for (int i = 0; i < NC; i++) xix[j1 * numCom + i] = tmp * Zn[j][i];
        xiP[j1] = tmp * (Zp[j] - Zj[j] / P);
    }
}

// MixtureComp::CalRhoPn_partial, src/MixtureComp.cpp:2648-2663.
void Mixture::CalRhoPn_partial()
{
    for (int j = 0; j < NP; j++) {
        const int j1 = phaseLabel[j];
        rhoP[j1] = xiP[j1] * MW[j];
        for (int m = 0; m < NC; m++)
            rhox[j1 * numCom + m] = xix[j1 * numCom + m] * MW[j] + xi[j1] / nu[j] * (MWC[m] - MW[j]);
    }
}

// MixtureComp::CalMuPnLBC_partial, src/MixtureComp.cpp:2667-2757 (the
// reference's der1IJ = -(1 / 2) * ... is an integer zero, kept as such).
void Mixture::CalMuPnLBC_partial()
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
            derxTj = (Tc[k] - xTj) / nu[j];
            derxPj = (Pc[k] - xPj) / nu[j];
            derMWj = (MWC[k] - MW[j]) / nu[j];
            der3J = 0;
            der4J = sqrtMWi[k];
            for (int i = 0; i < NC; i++) {
                val1IJ = muAux1I[i] / sqrt(MW[j]);
#ifdef OCP_ORACLE_EXACT_LBC
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
                der3J += sqrtMWi[i] * (xj[i] * der2IJ + (deltaN(i, k) - xj[i]) * val2IJ / nu[j]);
                der4J -= xj[i] * sqrtMWi[i];
            }
            der4J /= nu[j];
            der6J = 5.4402 *
                    (1.0 / 6 * pow(xTj, -5.0 / 6) * derxTj -
                     pow(xTj, 1.0 / 6) * (0.5 / MW[j] * derMWj + 2.0 / 3 / xPj * derxPj)) /
                    (sqrt(MW[j]) * pow(xPj, 2.0 / 3));
            der7J = xix[j1 * numCom + k] * xVj + (Vcvis[k] - xVj) * xi[j1] / nu[j];
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

// MixtureComp::CalVfiVfp_full03, src/MixtureComp.cpp:4822-4852: total
// derivatives of vf from the n_ij rows of dXsdXp.
void Mixture::CalVfiVfp_full03()
{
    if (NP == 1) {
        const int j1 = phaseLabel[0];
        for (int i = 0; i < NC; i++) vfi[i] = vji[j1][i];
    } else {
        const int ncol = numCom + 1;
        const real* nijPN = &dXsdXp[(NP + 1) * ncol];
        for (int j = 0; j < numPhase - 1; j++) {
            if (!phaseExist[j]) continue;
            for (int i = 0; i < NC; i++) {
                vfP += vji[j][i] * nijPN[0];
                nijPN++;
                for (int m = 0; m < NC; m++) vfi[m] += vji[j][i] * nijPN[m];
                nijPN += numCom;
            }
        }
    }
}

void Mixture::PhaseProps(double Pin, double Tin, const double* xin, double* out3)
{
    InitPTZ(Pin, Tin + CONV5, xin);
    NP = 1;
    CalAiBi();
    for (int i = 0; i < NC; i++) x[0][i] = zi[i];
    nu[0] = 1;
    CalAjBj(Aj[0], Bj[0], x[0]);
    SolEoS(Zj[0], Aj[0], Bj[0]);
    CalMW();
    CalVfXiRho();
    CalViscoLBC();
    out3[0] = todbl(xiC[0]);
    out3[1] = todbl(rhoC[0]);
    out3[2] = todbl(muC[0]);
}

}  // namespace oracle
