// Flash kernel templates: IsoT_FIM::CalFlash (FlashFIM + PassFlashValue per cell,
// src/OCPFluidMethod.cpp:1108-1187) and IsoT_FIM::InitFlash (InitFlashFIM,
// :1095-1106), one lane per cell over the SoA state of device_state.hpp.
#pragma once

#include "flash_coop.hpp"
#include "flash_device.hpp"
#include "flash_kernel.hpp"

namespace ocpg {

template <int NC>
struct FlashKernel {
    static constexpr int nc = NC + 1;
    static constexpr int ncol = nc + 1;

    // Everything after PhaseEquilibrium: MixtureComp::CalFlash tail
    // (src/MixtureComp.cpp:697-708), CalVfiVfp_full02 (:3911-3989), the
    // water block (:521-546 / :321-345), CalSaturation (:1178-1186),
    // CaldXsdXpAPI02p (:4386-4493), CalXiPNX_partial / CalRhoPX_partial /
    // CalMuPXLBC_partial (:3037-3213), pSderExist/pVnumCom (:559-570),
    // CalSkipForNextStep (:5155-5254), then PassFlashValue's stores.
    //
    // Inlined into its kernels for NC <= 6 (the cell object then stays in
    // registers); the larger instantiations keep it out of line.
    template <int NPT>
    __device__ __forceinline__ static void finish(FlashCell<NC>& F, const DevEos& E, const DevBulk& B, int64_t cell,
                                                  bool init, double Sw_in, double Vpore)
    {
        if constexpr (NC <= 6)
            finishBody<NPT>(F, E, B, cell, init, Sw_in, Vpore);
        else
            finishCall<NPT>(F, E, B, cell, init, Sw_in, Vpore);
    }

    template <int NPT>
    __device__ __noinline__ static void finishCall(FlashCell<NC>& F, const DevEos& E, const DevBulk& B,
                                                   int64_t cell, bool init, double Sw_in, double Vpore)
    {
        finishBody<NPT>(F, E, B, cell, init, Sw_in, Vpore);
    }

    template <int NPT>
    __device__ __forceinline__ static void finishBody(FlashCell<NC>& F, const DevEos& E, const DevBulk& B,
                                                      int64_t cell, bool init, double Sw_in, double Vpore)
    {
        const int64_t ld = B.ld;
        constexpr int NP = NPT;
        const double P = F.P, T = F.T;
        // ---- CalMW, CalVfXiRho, CalViscoLBC (per internal phase) ----
        double MW[2], vC[2], xiC[2], rhoC[2], muC[2], muA[2][5];
        double vf = 0;
        // The LBC terms' powers of Tc, Pc and Tr: the same expressions the
        // reference evaluates per phase and, in the derivatives, per (k, i)
        // pair, evaluated once (pow is a pure function: the same values).
        double visT[NC], tc16[NC], pc23[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            tc16[i] = pow(E.Tc[i], 1.0 / 6);
            pc23[i] = pow(E.Pc[i], 2.0 / 3);
            const double Tri = T / E.Tc[i];
            if (Tri <= 1.5)
                visT[i] = 34 * 1E-5 * pow(Tri, 0.94);
            else
                visT[i] = 17.78 * 1E-5 * pow((4.58 * Tri - 1.67), 0.625);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            MW[j] = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) MW[j] += F.x[j][i] * E.MWC[i];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            double tmp = F.Zj[j] * kGasConst * T / P;
#pragma unroll
            for (int i = 0; i < NC; ++i) tmp -= F.x[j][i] * E.Vshift[i];
            vC[j] = tmp * F.nu[j];
            vf += vC[j];
            xiC[j] = 1 / tmp;
            rhoC[j] = MW[j] * xiC[j];
        }
        double xT16[2], xP23[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            double m0 = 0, m1 = 0, xT = 0, xP = 0, xV = 0;
            const double sMW = sqrt(MW[j]);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                double tmp = 5.4402 * tc16[i] / sMW / pc23[i];
                tmp = visT[i] / tmp;
                m0 += F.x[j][i] * E.sqrtMWi[i] * tmp;
                m1 += F.x[j][i] * E.sqrtMWi[i];
                xT += F.x[j][i] * E.Tc[i];
                xP += F.x[j][i] * E.Pc[i];
                xV += F.x[j][i] * E.Vcvis[i];
            }
            muA[j][0] = m0;
            muA[j][1] = m1;
            xT16[j] = pow(xT, 1.0 / 6);
            xP23[j] = pow(xP, 2.0 / 3);
            muA[j][2] = 5.4402 * xT16[j] / sMW / xP23[j];
            muA[j][3] = xiC[j] * xV;
            const double a3 = muA[j][3];
            const double* L = E.LBCcoef;
            muA[j][4] = a3 * (a3 * (a3 * (L[4] * a3 + L[3]) + L[2]) + L[1]) + L[0];
            muC[j] = muA[j][0] / muA[j][1] + 1E-4 * (pow(muA[j][4], 4.0) - 1) / muA[j][2];
        }
        // ---- IdentifyPhase (src/MixtureComp.cpp:2352-2384) ----
        int lab[2];
        if (NP == 1) {
            double A = 0, Bv = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                A += F.x[0][i] * E.Vc[i] * E.Tc[i];
                Bv += F.x[0][i] * E.Vc[i];
            }
            lab[0] = (T > A / Bv) ? 1 : 0;
            lab[1] = 1 - lab[0];
        } else {
            lab[0] = (MW[0] > MW[1]) ? 0 : 1;
            lab[1] = 1 - lab[0];
        }
        // ---- InitFlashFIM: scale to moles from the pore volume ----
        double Nh = F.Nh;
        if (init) {
            Nh = Vpore * (1 - Sw_in) / vf;
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (j < NP) F.nu[j] = Nh * F.nu[j];
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (j < NP) vC[j] = Nh * vC[j];
            vf *= Nh;
            F.Nh = Nh;
#pragma unroll
            for (int i = 0; i < NC; ++i) F.Ni[i] = F.zi[i] * Nh;
        }
        // ---- CalVfiVfp_full02 ----
        const double CgTP = kGasConst * T / P;
        double vfi[nc], vfP;
        double vjp[2], vji[2][NC];  // internal phase order
        double dn[NC * (NC + 1)];   // rhsDer after the solve (column-major)
        double Zp[2];
        if (NP == 1) {
            const double aj = F.Aj[0], bj = F.Bj[0], zj = F.Zj[0];
            const double den = dZden(aj, bj, zj);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double tmp = F.aSum1(i, F.x[0]);
                const double An = 2 / F.nu[0] * (tmp - aj);
                const double Bn = 1 / F.nu[0] * (F.Bi[i] - bj);
                const double Zn = dZnum(aj, bj, zj, An, Bn) / den;
                vfi[i] = CgTP * (zj + F.nu[0] * Zn) - E.Vshift[i];
            }
            Zp[0] = dZnum(aj, bj, zj, aj, bj) / P / den;
            vfP = CgTP * F.nu[0] * (Zp[0] - zj / P);
        } else {
            double Zn[2][NC];
            double A[NC * NC];
            {
                double f1[FlashCell<NC>::NT];
                F.calFugN(0, f1, Zn[0]);
#pragma unroll
                for (int i = 0; i < NC; ++i)
#pragma unroll
                    for (int m = 0; m <= i; ++m) A[i * NC + m] = f1[FlashCell<NC>::tri(i, m)];
                F.calFugN(1, f1, Zn[1]);
#pragma unroll
                for (int i = 0; i < NC; ++i)
#pragma unroll
                    for (int m = 0; m <= i; ++m) {
                        const double v1 = f1[FlashCell<NC>::tri(i, m)];
                        const double s = A[i * NC + m] + v1;
                        A[i * NC + m] = s;
                        A[m * NC + i] = s;
                        // rhs column k+1 = fugN[1] row k: entries i of row m and m of row i
                        dn[(m + 1) * NC + i] = v1;
                        dn[(i + 1) * NC + m] = v1;
                    }
            }
            // CalFugPAll (src/MixtureComp.cpp:2519-2579)
            double fugP[2][NC];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const double aj = F.Aj[j], bj = F.Bj[j], zj = F.Zj[j];
                const double Ap = aj / P;
                const double Bp = bj / P;
                Zp[j] = dZnum(aj, bj, zj, Ap, Bp) / dZden(aj, bj, zj);
                const double G = (zj + kDelta1 * bj) / (zj + kDelta2 * bj);
                const double Gp = (kDelta1 - kDelta2) / ((zj + kDelta2 * bj) * (zj + kDelta2 * bj)) *
                                  (Bp * zj - Zp[j] * bj);
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double C = P / (zj - bj);
                    const double tmp = F.aSum1(i, F.x[j]);
                    const double Ei = -aj / ((kDelta1 - kDelta2) * bj) * (2 * tmp / aj - F.Bi[i] / bj);
                    const double Cp = ((zj - bj) - P * (Zp[j] - Bp)) / ((zj - bj) * (zj - bj));
                    const double Dp = F.Bi[i] / bj * Zp[j];
                    fugP[j][i] = Cp / C + Dp + Ei / G * Gp;
                }
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) dn[i] = fugP[1][i] - fugP[0][i];
            luSolve<NC, NC + 1>(A, dn);
            vjp[0] = CgTP * F.nu[0] * (Zp[0] - F.Zj[0] / P);
            vjp[1] = CgTP * F.nu[1] * (Zp[1] - F.Zj[1] / P);
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                vjp[0] += (CgTP * (F.Zj[0] + F.nu[0] * Zn[0][k]) - E.Vshift[k]) * dn[k];
                vjp[1] -= (CgTP * (F.Zj[1] + F.nu[1] * Zn[1][k]) - E.Vshift[k]) * dn[k];
            }
            vfP = vjp[0] + vjp[1];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double* d = dn + (i + 1) * NC;
                double a0 = 0, a1 = 0;
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    a0 += (CgTP * (F.Zj[0] + F.nu[0] * Zn[0][k]) - E.Vshift[k]) * d[k];
                    a1 += (CgTP * (F.Zj[1] + F.nu[1] * Zn[1][k]) - E.Vshift[k]) * ((i == k ? 1.0 : 0.0) - d[k]);
                }
                vji[0][i] = a0;
                vji[1][i] = a1;
                vfi[i] = a0 + a1;
            }
        }
        // ---- water (PVTW, OCPTable::Eval_All on column 0) ----
        double pw[5], cw[5];
        {
            const int nRow = E.pvtwRows;
            int seg = -1;
            for (int r = 0; r + 1 < nRow; ++r)
                if (P >= E.pvtw[r][0] && P < E.pvtw[r + 1][0]) {
                    seg = r;
                    break;
                }
            if (seg < 0) {
                const int r = (P < E.pvtw[0][0]) ? 0 : nRow - 1;
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    cw[q] = 0;
                    pw[q] = E.pvtw[r][q];
                }
            } else {
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    cw[q] = (E.pvtw[seg + 1][q] - E.pvtw[seg][q]) / (E.pvtw[seg + 1][0] - E.pvtw[seg][0]);
                    pw[q] = E.pvtw[seg][q] + cw[q] * (P - E.pvtw[seg][0]);
                }
            }
        }
        const double bw = pw[1] * (1 - pw[2] * (P - pw[0]));
        const double bwp = -pw[2] * pw[1];
        const double muW = pw[3];
        const double xiW = 1 / (kCONV1 * bw);
        const double rhoW = E.stdRhoW / bw;
        const double muPW = cw[3];
        const double xiPW = -bwp / (bw * bw * kCONV1);
        const double rhoPW = kCONV1 * xiPW * E.stdRhoW;
        if (init) F.Ni[NC] = Vpore * Sw_in * xiW;
        const double NiW = F.Ni[NC];
        const double Nt = Nh + NiW;
        const double vw = kCONV1 * NiW * bw;
        vfi[NC] = kCONV1 * bw;
        const double vwp = kCONV1 * NiW * bwp;
        vf += vw;
        vfP += vwp;
        // ---- CalSaturation ----
        double S[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (j < NP) S[lab[j]] = vC[j] / vf;
        S[2] = vw / vf;

        // ---- stores shared by every cell (PassFlashValue) ----
        const int64_t n = cell;
        B.phaseNum[n] = NP + 1;
        B.Nt[n] = Nt;
        B.vf[n] = vf;
        B.vfP[n] = vfP;
#pragma unroll
        for (int j = 0; j < 3; ++j) B.S[j * ld + n] = S[j];
        // non-finite results (a NaN or infinite input) are flagged apart
        // from non-convergence
        if (B.nonconv && !(isfinite(vf) && isfinite(S[0]) && isfinite(S[1]) && isfinite(S[2])))
            B.nonconv[n] = 2;
#pragma unroll
        for (int i = 0; i < nc; ++i) B.vfi[i * ld + n] = vfi[i];
        if (init) {
#pragma unroll
            for (int i = 0; i < nc; ++i) B.Ni[i * ld + n] = F.Ni[i];
        }
        const bool exist0 = (NP == 2) || lab[0] == 0;
        const bool exist1 = (NP == 2) || lab[0] == 1;
        B.phaseExist[0 * ld + n] = exist0;
        B.phaseExist[1 * ld + n] = exist1;
        B.phaseExist[2 * ld + n] = 1;
        B.pSderExist[0 * ld + n] = exist0;
        B.pSderExist[1 * ld + n] = exist1;
        B.pSderExist[2 * ld + n] = 1;
        B.pVnumCom[0 * ld + n] = exist0 ? NC : 0;
        B.pVnumCom[1 * ld + n] = exist1 ? NC : 0;
        B.pVnumCom[2 * ld + n] = 0;
        B.bRowSizedSdP[n] = (NP == 2) ? 3 + 2 * NC : 2 + NC;
        // water phase slot
        B.rho[2 * ld + n] = rhoW;
        B.xi[2 * ld + n] = xiW;
        B.mu[2 * ld + n] = muW;
        B.rhoP[2 * ld + n] = rhoPW;
        B.xiP[2 * ld + n] = xiPW;
        B.muP[2 * ld + n] = muPW;
#pragma unroll
        for (int i = 0; i < nc; ++i) {
            B.xij[(2 * nc + i) * ld + n] = (i == NC) ? 1.0 : 0.0;
            B.rhox[(2 * nc + i) * ld + n] = 0.0;
            B.xix[(2 * nc + i) * ld + n] = 0.0;
            B.mux[(2 * nc + i) * ld + n] = 0.0;
        }

        // ---- CalXiPNX_partial, CalRhoPX_partial, CalMuPXLBC_partial, stores ----
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            const int s = lab[j];
            const double aj = F.Aj[j], bj = F.Bj[j], zj = F.Zj[j];
            const double den = dZden(aj, bj, zj);
            const double mxi = -xiC[j] * xiC[j];
            double xix[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double Ax = 2 * F.aSumX(i, F.x[j]);
                const double Zx = dZnum(aj, bj, zj, Ax, F.Bi[i]) / den;
                xix[i] = mxi * (CgTP * Zx - E.Vshift[i]);
            }
            const double xiP = mxi * CgTP * (Zp[j] - zj / P);
            const double rhoP = xiP * MW[j];
            // LBC derivatives
            double xTj = 0, xPj = 0, xVj = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                xTj += F.x[j][i] * E.Tc[i];
                xPj += F.x[j][i] * E.Pc[i];
                xVj += F.x[j][i] * E.Vcvis[i];
            }
            const double* L = E.LBCcoef;
            const double a3 = muA[j][3];
            const double poly = (L[1] + a3 * (2 * L[2] + a3 * (3 * L[3] + a3 * 4 * L[4])));
            const double der7P = xVj * xiP;
            const double der8P = der7P * poly;
            const double a4c = pow(muA[j][4], 3.0);
            const double muP = (4 * 1E-4) * a4c * der8P / muA[j][2];
            const double sMW = sqrt(MW[j]);
            const double pMW15 = pow(MW[j], -1.5);
            const double a4q = pow(muA[j][4], 4.0);
            B.rho[s * ld + n] = rhoC[j];
            B.xi[s * ld + n] = xiC[j];
            B.mu[s * ld + n] = muC[j];
            B.rhoP[s * ld + n] = rhoP;
            B.xiP[s * ld + n] = xiP;
            B.muP[s * ld + n] = muP;
            const double xTm56 = pow(xTj, -5.0 / 6);
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                // der1IJ of the reference is -(1 / 2) * ... = 0 (integer division)
                double der3J = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double val1IJ = E.muAux1I[i] / sMW;
                    const double der1IJ = 0 * E.muAux1I[i] * pMW15 * E.MWC[k];
                    const double tmp = visT[i];
                    const double val2IJ = tmp / val1IJ;
                    const double der2IJ = -tmp * der1IJ / (val1IJ * val1IJ);
                    der3J += F.x[j][i] * E.sqrtMWi[i] * der2IJ + (i == k ? 1.0 : 0.0) * E.sqrtMWi[k] * val2IJ;
                }
                const double der4J = E.sqrtMWi[k];
                const double der6J = 5.4402 *
                                     (1.0 / 6 * xTm56 * E.Tc[k] -
                                      xT16[j] * (0.5 / MW[j] * E.MWC[k] + 2.0 / 3 / xPj * E.Pc[k])) /
                                     (sMW * xP23[j]);
                const double der7J = xix[k] * xVj + xiC[j] * E.Vcvis[k];
                const double der8J = der7J * poly;
                const double mux = (der3J * muA[j][1] - muA[j][0] * der4J) / (muA[j][1] * muA[j][1]) +
                                   1E-4 * (4 * a4c * der8J * muA[j][2] - (a4q - 1) * der6J) /
                                       (muA[j][2] * muA[j][2]);
                const double rhox = xix[k] * MW[j] + xiC[j] * E.MWC[k];
                B.xij[(s * nc + k) * ld + n] = F.x[j][k];
                B.rhox[(s * nc + k) * ld + n] = rhox;
                B.xix[(s * nc + k) * ld + n] = xix[k];
                B.mux[(s * nc + k) * ld + n] = mux;
            }
            B.xij[(s * nc + NC) * ld + n] = 0.0;
            B.rhox[(s * nc + NC) * ld + n] = 0.0;
            B.xix[(s * nc + NC) * ld + n] = 0.0;
            B.mux[(s * nc + NC) * ld + n] = 0.0;
        }

        // ---- CaldXsdXpAPI02p: compact d(S, x)/d(P, N) rows ----
        double* D = B.dSec;
        const double vf2 = vf * vf;
        if (NP == 1) {
            const double vjh = vC[0];
            double r0[ncol];
            r0[0] = ((vfP - vwp) * vf - vfP * vjh) / vf2;
#pragma unroll
            for (int m = 0; m < NC; ++m) r0[1 + m] = vfi[m] * vw / vf2;
            r0[1 + NC] = -vfi[NC] * vjh / vf2;
#pragma unroll
            for (int c = 0; c < ncol; ++c) {
                D[(0 * ncol + c) * ld + n] = r0[c];
                D[(1 * ncol + c) * ld + n] = -r0[c];
            }
            const double Nh2 = Nh * Nh;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                D[((2 + i) * ncol + 0) * ld + n] = 0.0;
#pragma unroll
                for (int m = 0; m < NC; ++m)
                    D[((2 + i) * ncol + 1 + m) * ld + n] = ((i == m ? 1.0 : 0.0) * Nh - F.Ni[i]) / Nh2;
                D[((2 + i) * ncol + 1 + NC) * ld + n] = 0.0;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int row = lab[j];
                D[(row * ncol + 0) * ld + n] = (vjp[j] * vf - vfP * vC[j]) / vf2;
#pragma unroll
                for (int m = 0; m < NC; ++m)
                    D[(row * ncol + 1 + m) * ld + n] = (vji[j][m] * vf - vfi[m] * vC[j]) / vf2;
                D[(row * ncol + 1 + NC) * ld + n] = (0.0 * vf - vfi[NC] * vC[j]) / vf2;
            }
            D[(2 * ncol + 0) * ld + n] = (vwp * vf - vfP * vw) / vf2;
#pragma unroll
            for (int m = 0; m < NC; ++m) D[(2 * ncol + 1 + m) * ld + n] = (0.0 * vf - vfi[m] * vw) / vf2;
            D[(2 * ncol + 1 + NC) * ld + n] = (vfi[NC] * vf - vfi[NC] * vw) / vf2;
            const int base0 = 3 + lab[0] * NC;
            const int base1 = 3 + lab[1] * NC;
            const double nu0s = F.nu[0] * F.nu[0];
            const double nu1s = F.nu[1] * F.nu[1];
#pragma unroll
            for (int c = 0; c <= NC; ++c) {
                const double* d = dn + c * NC;
                double sum = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) sum += d[i];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double v0 = (d[i] * F.nu[0] - sum * Nh * F.n[0][i]) / nu0s;
                    double v1;
                    if (c == 0)
                        v1 = (-d[i] * F.nu[1] + sum * Nh * F.n[1][i]) / nu1s;
                    else
                        v1 = (((i == c - 1) ? 1.0 : 0.0) - d[i]) * F.nu[1] - (1 - sum) * Nh * F.n[1][i];
                    if (c != 0) v1 = v1 / nu1s;
                    D[((base0 + i) * ncol + c) * ld + n] = v0;
                    D[((base1 + i) * ncol + c) * ld + n] = v1;
                }
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                D[((base0 + i) * ncol + 1 + NC) * ld + n] = 0.0;
                D[((base1 + i) * ncol + 1 + NC) * ld + n] = 0.0;
            }
        }

        // ---- CalSkipForNextStep (only a one-phase result keeps the skip
        // flag: PhaseEquilibrium clears it when NP > 1) ----
        if constexpr (NPT == 1) if (F.flagSkip && F.ftype == 0) {
            // CalPhiNSkip (src/MixtureComp.cpp:5155-5217) + AssembleSkipMatSTA
            const double aj = F.Aj[0], bj = F.Bj[0], zj = F.Zj[0], nu0 = F.nu[0];
            const double* xj = F.x[0];
            double An[NC], Bn[NC], Zn[NC], tI[NC];
            const double den = dZden(aj, bj, zj);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                tI[i] = F.aSum1(i, xj);
                An[i] = 2 / nu0 * (tI[i] - aj);
                Bn[i] = 1 / nu0 * (F.Bi[i] - bj);
                Zn[i] = dZnum(aj, bj, zj, An[i], Bn[i]) / den;
            }
            const double G = (zj + kDelta1 * bj) / (zj + kDelta2 * bj);
            const double lG = log(G);
            float M[NC * NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double C = 1 / (zj - bj);
                const double tmp = tI[i];
                const double Ei = -aj / ((kDelta1 - kDelta2) * bj) * (2 * tmp / aj - F.Bi[i] / bj);
#pragma unroll
                for (int k = 0; k <= i; ++k) {
                    const double aik = F.omb(i, k) * F.sq(i, k);
                    const double Cnk = (Bn[k] - Zn[k]) / ((zj - bj) * (zj - bj));
                    const double Dnk = F.Bi[i] / bj * (Zn[k] - (F.Bi[k] - bj) * (zj - 1) / (nu0 * bj));
                    const double Gnk = (kDelta1 - kDelta2) / ((zj + kDelta2 * bj) * (zj + kDelta2 * bj)) *
                                       (Bn[k] * zj - Zn[k] * bj);
                    double Enk = -1 / (kDelta1 - kDelta2) / (bj * bj) *
                                 (2 * (bj * aik / nu0 + Bn[k] * (F.Bi[i] * aj / bj - tmp)) - An[k] * F.Bi[i] -
                                  aj * F.Bi[i] / nu0);
                    Enk -= Ei / nu0;
                    const double phiN = 1 / C * Cnk + Dnk + Enk * lG + Ei / G * Gnk;
                    const float v = static_cast<float>((i == k ? 1.0 : 0.0) + nu0 * sqrt(xj[i] * xj[k]) * phiN);
                    M[i * NC + k] = v;
                    M[k * NC + i] = v;
                }
            }
            const float lam = minEigenSym<NC>(M);
            B.skipMinEig[n] = static_cast<double>(lam);
            B.skipP[n] = P;
            B.skipT[n] = T;
#pragma unroll
            for (int i = 0; i < NC; ++i) B.skipZi[i * ld + n] = F.zi[i];
        }
        B.skipFlag[n] = F.flagSkip ? 1 : 0;
    }
};

// SkipStaAnaly::IfSkip + CalFtypeFIM, src/AcceleratePVT.cpp:54-120.
template <int NC>
__device__ __forceinline__ int calFtypeFIM(const FlashCell<NC>& F, const DevBulk& B, int64_t n, const double* S)
{
    const int64_t ld = B.ld;
    if (B.skipFlag[n]) {
        const double lam = B.skipMinEig[n];
        bool skip = true;
        if (fabs(1 - B.skipP[n] / F.P) >= lam / 10) skip = false;
        else if (fabs(B.skipT[n] - F.T) >= lam * 10) skip = false;
        else {
#pragma unroll
            for (int i = 0; i < NC; ++i)
                if (fabs(F.Ni[i] / F.Nh - B.skipZi[i * ld + n]) >= lam / 10) skip = false;
        }
        if (skip) return 1;
    }
    if (F.lNP >= 2) {
        if (S[0] < 1E-4 || S[1] < 1E-4) return 0;
        return 2;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Staged flash.  One call of IsoT_FIM::CalFlash runs five kernels over the
// cells:
//   classify  : ftype of every cell (SkipStaAnaly::CalFtypeFIM), lists
//   stability : PhaseStable on the ftype-0 list
//   split     : PhaseSplit on the cells that need it (ftype 2, unstable)
//   finish1/2 : CalFlash tail + derivatives + stores for the cells that end
//               with one / two hydrocarbon phases
// Each cell's result is what one FlashFIM call computes; the stages only
// regroup cells so that a wavefront runs one kind of work, and each kernel
// carries only its own working set.  State between stages lives in the
// FlashWork buffers (SoA, leading dimension ld).
//
// Overlapped form (launchFlashT with side streams): the few cells whose
// stability test or split takes far longer than the rest (hundreds of
// iterations, Newton restarts) would hold the whole pass.  There the
// stability test and the split of the cells it finds unstable run on side
// stream B, the split slices' leftovers on side stream C and the cells
// whose first Newton split fails on side stream D, each to the end of its
// split, while the main stream finishes the other cells and goes on to the
// caller's next kernels.  The caller joins the side streams, finishes their
// cells (launchFlashSideFinishT: one small launch once the chip is free,
// where a side-stream launch would wait for the main stream's kernels to
// drain) and recomputes what read them (flashSideCells).  Every cell still
// runs the same arithmetic on the same inputs.
// ---------------------------------------------------------------------------

enum { FL_SKIP = 1, FL_SSMSTA = 2, FL_NRSTA = 4, FL_SSMSP = 8, FL_NRSP = 16, FL_GUIDED = 32, FL_DEFER = 64 };
// Blocks of a work-queue stage: 256 CUs x 4 SIMDs x 2 waves, in blocks of
// kFlashBlock / 64 waves.
constexpr unsigned kQueueBlocks = 256 * 4 * 2 / (kFlashBlock / 64);
// SplitSSM2 iterations of the first two successive-substitution slices (the
// warm pass needs 8 or fewer in half of its two-phase cells, 15 at the 90th
// percentile); the third slice runs to the end.
constexpr int kSplitBudget0 = 8;
// Whether the split starts from the learned K where a trial from it proved
// the feed unstable (measured on the 1.1M-cell cold pass: fewer split
// iterations in total but a longer tail than Wilson's start, so off).
constexpr bool kSplitFromGuess = false;
constexpr int kSplitBudget1 = 24;
// Overlapped form: the third slice stops after this many more iterations
// (the warm pass: 0.3% of the two-phase cells need more than 96 in all) and
// hands the rest to side stream C.
constexpr int kSplitBudget2 = 64;
// Longest lists the side streams take (longer ones stay on the main
// stream), and their grids: the per-lane side kernels hold a few waves;
// the one-cell-per-wavefront kernels start every cell of a short list at
// once.
constexpr int kSideMaxStab = 4096;
constexpr int kSideMaxTail = 2048;
constexpr int kSideMaxRestart = 1 << 30;  // restarts are the slowest cells: always beside
constexpr unsigned kSideBlocks = 32;
constexpr unsigned kCoopBlocks = 256;
// The restart list can hold thousands of cells, each milliseconds long (a
// faulted 10-component field): its kernel takes one wave on every SIMD
// (256 CUs x 4), so that once the main stream is done the whole chip drains
// it; a short list leaves the extra waves nothing to take.
constexpr unsigned kRestartBlocks = 256 * 4 / (kFlashBlock / 64);

// Append the active lanes' cells to list k: one atomic per wavefront.
__device__ __forceinline__ void appendList(const FlashWork& W, int k, bool pred, int32_t cell)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&W.count[k], __popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) {
        const unsigned long long below = (lane == 0) ? 0ull : (mask & ((1ull << lane) - 1ull));
        W.list[k][base + __popcll(below)] = cell;
    }
}

// Work queue over list k: each wavefront takes the next 64 entries (its
// base index, or a value >= the list length when the list is exhausted).
__device__ __forceinline__ int nextChunk(const FlashWork& W, int k)
{
    int base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(&W.count[kFlashLists + k], 64);
    return __shfl(base, 0, 64);
}

// Overlapped form: a list is worked by a side stream while it is short and
// by the main stream once it is long (side streams hold only a few waves).
// gate > 0: run only when the list has at most gate cells; gate < 0: only
// when it has more than -gate; 0: always.
__device__ __forceinline__ bool gateOpen(int gate, int count)
{
    return gate == 0 || (gate > 0 ? count <= gate : count > -gate);
}

__device__ __forceinline__ unsigned long long waveSum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void addCounter(unsigned long long* dst, unsigned long long v)
{
    v = waveSum(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// (a template so each per-NC translation unit has its own symbol)
template <int NC>
__global__ void flashResetKernel(FlashWork W, DevFlashCounters* ctr)
{
    if (threadIdx.x < 2 * kFlashLists) W.count[threadIdx.x] = 0;
    if (ctr && threadIdx.x < sizeof(DevFlashCounters) / sizeof(unsigned long long))
        reinterpret_cast<unsigned long long*>(ctr)[threadIdx.x] = 0;
}

// Stage 1: zi, Nh, lNP, ftype; lists.  zin non-NULL: InitFlashFIM (the feed
// comes in, ftype 0, lNP 0).
template <int NC>
__global__ void __launch_bounds__(kFlashBlock) flashClassifyKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                   FlashWork W, DevFlashCounters* ctr,
                                                                   const double* __restrict__ zin)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = n < B.nb && cellOn(B, n);
    const int64_t ld = B.ld;
    constexpr int nc = NC + 1;
    int ftype = 0;
    if (active) {
        FlashCell<NC> F(*eos);
        int lNP = 0;
        if (zin) {
#pragma unroll
            for (int i = 0; i < NC; ++i) W.zi[i * ld + n] = zin[i * ld + n];
        } else {
            const int lastNP = B.phaseNum[n];
            lNP = lastNP > 0 ? lastNP - 1 : 0;
            F.lNP = lNP;
            F.P = B.P[n];
            F.T = B.T[n] + kCONV5;
#pragma unroll
            for (int i = 0; i < nc; ++i) F.Ni[i] = B.Ni[i * ld + n];
            double Nh = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) Nh += fabs(F.Ni[i]);
            F.Nh = Nh;
#pragma unroll
            for (int i = 0; i < NC; ++i) W.zi[i * ld + n] = F.Ni[i] / Nh;
            W.Nh[n] = Nh;
            double S[2] = {B.S[0 * ld + n], B.S[1 * ld + n]};
            ftype = calFtypeFIM<NC>(F, B, n, S);
        }
        W.ftype[n] = (int8_t)ftype;
        W.lNP[n] = (int8_t)lNP;
        W.flags[n] = (ftype == 2) ? 0 : FL_SKIP;
        if (ftype == 2) W.Yt[n] = 1.01;
        if (B.ftype) B.ftype[n] = ftype;
        if (B.nonconv) B.nonconv[n] = 0;
        if (B.iters) {
#pragma unroll
            for (int q = 0; q < 5; ++q) B.iters[q * ld + n] = 0;
        }
    }
    // cells the learned prior calls unstable get their own stability list, so
    // the waves that try the guessed K first are not mixed with full tests
    const bool guided = active && W.guess && W.guess[(int64_t)n * W.guessLd + NC] > 0.f;
    appendList(W, LIST_STAB, active && ftype == 0 && !guided, (int32_t)n);
    appendList(W, LIST_STABG, active && ftype == 0 && guided, (int32_t)n);
    appendList(W, LIST_SPLIT, active && ftype == 2, (int32_t)n);
    appendList(W, LIST_NP1, active && ftype == 1, (int32_t)n);
    if (ctr) {
        addCounter(&ctr->cells, active ? 1 : 0);
        addCounter(&ctr->ftype[0], (active && ftype == 0) ? 1 : 0);
        addCounter(&ctr->ftype[1], (active && ftype == 1) ? 1 : 0);
        addCounter(&ctr->ftype[2], (active && ftype == 2) ? 1 : 0);
    }
}

template <int NC>
__device__ __forceinline__ void loadFeed(FlashCell<NC>& F, const DevBulk& B, const FlashWork& W, int64_t n)
{
    const int64_t ld = B.ld;
    F.P = B.P[n];
    F.T = B.T[n] + kCONV5;
#pragma unroll
    for (int i = 0; i < NC; ++i) F.zi[i] = W.zi[i * ld + n];
    F.lNP = W.lNP[n];
    F.ftype = W.ftype[n];
}

// Stage 2: PhaseEquilibrium, ftype 0 part up to the stability verdict.
// Guided (a learned K guess is present): one work queue over LIST_STAB
// followed by LIST_STABG (the cells the learned prior calls unstable), each
// padded to whole wavefront chunks so a wave works on one list only; the
// cells of LIST_STABG try the guessed K first.  The plain flash runs the
// Guided = false instance, which never touches the guess.  Unstable cells go
// to list outSplit, stable ones to outNP1.
template <int NC, bool Guided>
__global__ void __launch_bounds__(kFlashBlock) flashStabilityKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                    FlashWork W, DevFlashCounters* ctr,
                                                                    int outSplit, int outNP1, int gate)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int nA = W.count[LIST_STAB];
    const int padA = Guided ? (nA + 63) & ~63 : nA;
    const int nG = Guided ? W.count[LIST_STABG] : 0;
    const int total = padA + nG;
    if (!gateOpen(gate, total)) return;
    for (int base = nextChunk(W, LIST_STAB); base < total; base = nextChunk(W, LIST_STAB)) {
    const bool guidedChunk = Guided && base >= padA;
    const int in = guidedChunk ? LIST_STABG : LIST_STAB;
    const int idx = base - (guidedChunk ? padA : 0) + (threadIdx.x & 63);
    const bool active = idx < (guidedChunk ? nG : nA);
    bool stable = false;
    int it0 = 0, it1 = 0;
    bool nonconv = false;
    int32_t n = 0;
    if (active) {
        n = W.list[in][idx];
        const int64_t ld = B.ld;
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
#pragma unroll
        for (int q = 0; q < 5; ++q) F.it[q] = 0;
        F.flagSkip = true;
        F.singlePhaseFeed();
        F.calKwilson();
        int fl = 0;
        if constexpr (Guided) {
            double Kg[NC];
            if (guidedChunk) {
#pragma unroll
                for (int i = 0; i < NC; ++i) Kg[i] = exp((double)W.guess[(int64_t)n * W.guessLd + i]);
                F.Kguide = Kg;
            }
            stable = F.template phaseStable<true>();
            if (F.guidedProof) fl |= FL_GUIDED;
        } else {
            stable = F.phaseStable();
        }
        it0 = F.it[IT_SSMSTA];
        it1 = F.it[IT_NRSTA];
        if (F.flagSkip) fl |= FL_SKIP;
        if (F.ssmStaConflag) fl |= FL_SSMSTA;
        if (F.nrStaConflag) fl |= FL_NRSTA;
        W.flags[n] = (uint8_t)fl;
        if (!stable) W.Yt[n] = F.Yt;
        if (stable) nonconv = !(F.nrStaConflag || F.ssmStaConflag);
        if (stable && B.nonconv) B.nonconv[n] = nonconv ? 1 : 0;
        if (B.iters) {
            B.iters[IT_SSMSTA * ld + n] = it0;
            B.iters[IT_NRSTA * ld + n] = it1;
        }
    }
    appendList(W, outNP1, active && stable, n);
    appendList(W, outSplit, active && !stable, n);
    if (ctr) {
        addCounter(&ctr->iters[IT_SSMSTA], (unsigned long long)it0);
        addCounter(&ctr->iters[IT_NRSTA], (unsigned long long)it1);
        addCounter(&ctr->nonconverged, nonconv ? 1 : 0);
    }
    }
}

// The split's starting K (SplitSSM2 with restart = false, src/MixtureComp.cpp:
// 1858-1870): the previous phase compositions' ratio when the cell had two
// hydrocarbon phases, else Wilson's.
template <int NC>
__device__ __forceinline__ void splitStartK(FlashCell<NC>& F, const DevBulk& B, const FlashWork& W, int64_t n,
                                            int fl)
{
    constexpr int nc = NC + 1;
    const int64_t ld = B.ld;
    if (F.lNP == 2) {
#pragma unroll
        for (int i = 0; i < NC; ++i) F.Ks[i] = B.xij[(0 * nc + i) * ld + n] / B.xij[(1 * nc + i) * ld + n];
    } else if (W.guess && kSplitFromGuess && (fl & FL_GUIDED)) {
#pragma unroll
        for (int i = 0; i < NC; ++i) F.Ks[i] = exp((double)W.guess[(int64_t)n * W.guessLd + i]);
    } else {
        F.calKwilson();
#pragma unroll
        for (int i = 0; i < NC; ++i) F.Ks[i] = F.Kw0[i];
    }
}

// Stage 3: PhaseSplit (two hydrocarbon phases), src/MixtureComp.cpp:1743-1782,
// as successive-substitution slices and a Newton stage.  A slice runs at most
// `budget` SplitSSM2 iterations per cell (budget < 0: to the end); cells
// still iterating go to list `cont` with their state (Ks, Se, iteration
// counts), finished ones to list nrOut.  Bounding the iterations per slice keeps
// a wavefront from idling on its slowest lane: the cells that need many
// iterations continue together in the next slice.
template <int NC>
__global__ void __launch_bounds__(kFlashBlock) flashSplitSSMKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                   FlashWork W, int in, int cont, int budget,
                                                                   int first, int nrOut, int gate)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int64_t ld = B.ld;
    const int total = W.count[in];
    if (!gateOpen(gate, total)) return;
    for (int base = nextChunk(W, in); base < total; base = nextChunk(W, in)) {
    const int idx = base + (threadIdx.x & 63);
    const bool active = idx < total;
    bool done = false;
    int32_t n = 0;
    if (active) {
        n = W.list[in][idx];
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
        F.Yt = W.Yt[n];
        F.NP = 2;
        F.calAiBi();
        double Se;
        int iter;
        int fl = W.flags[n];
        if (first) {
            splitStartK<NC>(F, B, W, n, fl);
            Se = 1;
            iter = 0;
            F.it[IT_RR] = 0;
            F.ssmSpConflag = true;
        } else {
#pragma unroll
            for (int i = 0; i < NC; ++i) F.Ks[i] = W.Ks[i * ld + n];
            Se = W.Se[n];
            iter = W.ssmIt[n];
            F.it[IT_RR] = W.rrIt[n];
            F.ssmSpConflag = (fl & FL_SSMSP) != 0;
        }
        done = F.splitSSM2Slice(Se, iter, budget);
#pragma unroll
        for (int i = 0; i < NC; ++i) W.Ks[i * ld + n] = F.Ks[i];
        W.Se[n] = Se;
        W.ssmIt[n] = iter;
        W.rrIt[n] = F.it[IT_RR];
        W.flags[n] = (uint8_t)(F.ssmSpConflag ? (fl | FL_SSMSP) : (fl & ~FL_SSMSP));
        if (done) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                W.nu[j * ld + n] = F.nu[j];
#pragma unroll
                for (int i = 0; i < NC; ++i) W.x[(j * NC + i) * ld + n] = F.x[j][i];
            }
        }
    }
    appendList(W, nrOut, active && done, n);
    if (cont >= 0) appendList(W, cont, active && !done, n);
    }
}

// A cell's split is over: its flags, phase count and compositions go to the
// work buffers, its iteration counts to B.iters (the end of the Newton stage
// below, shared with the side streams' kernel).
template <int NC>
__device__ __forceinline__ void splitDone(FlashCell<NC>& F, const DevBulk& B, const FlashWork& W, int64_t n,
                                          int fl, int* its, bool& nonconv)
{
    const int64_t ld = B.ld;
    const int NP = F.NP;
    if (F.ftype == 0) {
        if (NP > 1) fl &= ~FL_SKIP;
        nonconv = (NP == 1) ? !(F.nrStaConflag || F.ssmStaConflag) : !(F.nrSpConflag || F.ssmSpConflag);
    } else {
        nonconv = !(F.nrSpConflag || F.ssmSpConflag);
    }
    if (B.nonconv) B.nonconv[n] = nonconv ? 1 : 0;
    W.flags[n] = (uint8_t)fl;
    if (NP == 2) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            W.nu[j * ld + n] = F.nu[j];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                W.x[(j * NC + i) * ld + n] = F.x[j][i];
                W.n[(j * NC + i) * ld + n] = F.n[j][i];
            }
        }
    }
    its[0] = F.it[IT_SSMSP];
    its[1] = F.it[IT_NRSP];
    its[2] = F.it[IT_RR];
    if (B.iters) {
        B.iters[IT_SSMSP * ld + n] = its[0];
        B.iters[IT_NRSP * ld + n] = its[1];
        B.iters[IT_RR * ld + n] = its[2];
    }
}

// The split's Newton stage and its restarts (PhaseSplit after the first
// SplitSSM2), from the state the last slice left: the compositions and
// fractions of its last Rachford-Rice solve and the fugacities of those
// compositions (recomputed: the same arithmetic on the same inputs).
// deferList >= 0 (overlapped form): a cell whose first Newton split fails
// leaves for that list (FL_DEFER set; a side stream runs its restarts) with
// its Newton iteration count; Ks and the other counters are already in the
// buffers.
template <int NC>
__global__ void __launch_bounds__(kFlashBlock) flashSplitNRKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                  FlashWork W, DevFlashCounters* ctr, int deferList,
                                                                  int in, int outNP1, int outNP2)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int64_t ld = B.ld;
    const int total = W.count[in];
    for (int base = nextChunk(W, in); base < total; base = nextChunk(W, in)) {
    const int idx = base + (threadIdx.x & 63);
    const bool active = idx < total;
    int NP = 1;
    int its[3] = {0, 0, 0};
    bool nonconv = false;
    bool deferred = false;
    int32_t n = 0;
    if (active) {
        n = W.list[in][idx];
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
        int fl = W.flags[n];
        F.ssmStaConflag = fl & FL_SSMSTA;
        F.nrStaConflag = fl & FL_NRSTA;
        F.ssmSpConflag = (fl & FL_SSMSP) != 0;
        F.nrSpConflag = false;
        F.Yt = W.Yt[n];
        F.NP = 2;
        F.calAiBi();
#pragma unroll
        for (int i = 0; i < NC; ++i) F.Ks[i] = W.Ks[i * ld + n];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            F.nu[j] = W.nu[j * ld + n];
#pragma unroll
            for (int i = 0; i < NC; ++i) F.x[j][i] = W.x[(j * NC + i) * ld + n];
        }
        F.it[IT_SSMSTA] = 0;
        F.it[IT_NRSTA] = 0;
        F.it[IT_SSMSP] = W.ssmIt[n];
        F.it[IT_NRSP] = 0;
        F.it[IT_RR] = W.rrIt[n];
        F.calFugPhiAll();
        F.splitNR();
        if (deferList >= 0 && !F.nrSpConflag) {
            deferred = true;
            W.nrIt[n] = F.it[IT_NRSP];
            W.flags[n] = (uint8_t)(fl | FL_DEFER);
        } else {
            F.splitRestarts();
            NP = F.NP;
            splitDone<NC>(F, B, W, n, fl, its, nonconv);
        }
    }
    appendList(W, outNP2, active && !deferred && NP == 2, n);
    appendList(W, outNP1, active && !deferred && NP == 1, n);
    if (deferList >= 0) appendList(W, deferList, active && deferred, n);
    if (ctr) {
        addCounter(&ctr->iters[IT_SSMSP], (unsigned long long)its[0]);
        addCounter(&ctr->iters[IT_NRSP], (unsigned long long)its[1]);
        addCounter(&ctr->iters[IT_RR], (unsigned long long)its[2]);
        addCounter(&ctr->nonconverged, nonconv ? 1 : 0);
        addCounter(&ctr->twoPhase, (active && !deferred && NP == 2) ? 1 : 0);
    }
    }
}

// Side streams of the overlapped form: a split run to its end, from
//   TAIL_FULL    the start (cells the side stability test found unstable),
//   TAIL_SSM     where the main stream's last slice stopped, or
//   TAIL_RESTART the failed first Newton split (the restart loop),
// with the same arithmetic as the slices and the Newton stage.  Finished
// cells go to lists outNP1 / outNP2.
enum { TAIL_FULL = 0, TAIL_SSM = 1, TAIL_RESTART = 2 };

template <int NC>
__global__ void __launch_bounds__(kFlashBlock) flashSplitTailKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                    FlashWork W, DevFlashCounters* ctr, int in,
                                                                    int mode, int outNP1, int outNP2, int gate)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int64_t ld = B.ld;
    const int total = W.count[in];
    if (!gateOpen(gate, total)) return;
    for (int base = nextChunk(W, in); base < total; base = nextChunk(W, in)) {
    const int idx = base + (threadIdx.x & 63);
    const bool active = idx < total;
    int NP = 1;
    int its[3] = {0, 0, 0};
    bool nonconv = false;
    int32_t n = 0;
    if (active) {
        n = W.list[in][idx];
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
        const int fl = W.flags[n];
        F.ssmStaConflag = fl & FL_SSMSTA;
        F.nrStaConflag = fl & FL_NRSTA;
        F.nrSpConflag = false;
        F.Yt = W.Yt[n];
        F.NP = 2;
        F.calAiBi();
        F.it[IT_SSMSTA] = 0;
        F.it[IT_NRSTA] = 0;
        F.it[IT_NRSP] = 0;
        if (mode == TAIL_RESTART) {
#pragma unroll
            for (int i = 0; i < NC; ++i) F.Ks[i] = W.Ks[i * ld + n];
            F.it[IT_SSMSP] = W.ssmIt[n];
            F.it[IT_NRSP] = W.nrIt[n];
            F.it[IT_RR] = W.rrIt[n];
            F.ssmSpConflag = (fl & FL_SSMSP) != 0;
            F.splitRestarts();
        } else {
            double Se;
            int iter;
            if (mode == TAIL_FULL) {
                splitStartK<NC>(F, B, W, n, fl);
                Se = 1;
                iter = 0;
                F.it[IT_RR] = 0;
                F.ssmSpConflag = true;
            } else {
#pragma unroll
                for (int i = 0; i < NC; ++i) F.Ks[i] = W.Ks[i * ld + n];
                Se = W.Se[n];
                iter = W.ssmIt[n];
                F.it[IT_RR] = W.rrIt[n];
                F.ssmSpConflag = (fl & FL_SSMSP) != 0;
            }
            F.splitSSM2Slice(Se, iter, -1);
            F.it[IT_SSMSP] = iter;
            F.calFugPhiAll();
            F.phaseSplitAfterSSM();
        }
        NP = F.NP;
        splitDone<NC>(F, B, W, n, fl, its, nonconv);
    }
    appendList(W, outNP2, active && NP == 2, n);
    appendList(W, outNP1, active && NP == 1, n);
    if (ctr) {
        addCounter(&ctr->iters[IT_SSMSP], (unsigned long long)its[0]);
        addCounter(&ctr->iters[IT_NRSP], (unsigned long long)its[1]);
        addCounter(&ctr->iters[IT_RR], (unsigned long long)its[2]);
        addCounter(&ctr->nonconverged, nonconv ? 1 : 0);
        addCounter(&ctr->twoPhase, (active && NP == 2) ? 1 : 0);
    }
    }
}

// Side streams, one wavefront per cell (flash_coop.hpp): the few cells
// whose split runs past the slices (as TAIL_SSM) or into Newton restarts
// (FL_DEFER set: as TAIL_RESTART) continue with their successive
// substitution spread over the wavefront's lanes; SplitNR and CheckSplit
// run on every lane alike.  Same results as flashSplitTailKernel, bit for
// bit.
__device__ __forceinline__ int nextCell(const FlashWork& W, int k)
{
    int idx = 0;
    if ((threadIdx.x & 63) == 0) idx = atomicAdd(&W.count[kFlashLists + k], 1);
    return __shfl(idx, 0, 64);
}

// (one wavefront per SIMD: the few waves of this kernel may use the whole
// register file instead of spilling to scratch memory)
template <int NC>
__global__ void __launch_bounds__(kFlashBlock) __attribute__((amdgpu_waves_per_eu(1, 1)))
flashSplitTailCoopKernel(const DevEos* __restrict__ eos, DevBulk B, FlashWork W, DevFlashCounters* ctr, int in,
                         int outNP1, int outNP2, int gate)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int64_t ld = B.ld;
    const int total = W.count[in];
    if (!gateOpen(gate, total)) return;
    const bool lead = (threadIdx.x & 63) == 0;
    for (int idx = nextCell(W, in); idx < total; idx = nextCell(W, in)) {
        const int32_t n = W.list[in][idx];
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
        const int fl = W.flags[n];
        F.ssmStaConflag = fl & FL_SSMSTA;
        F.nrStaConflag = fl & FL_NRSTA;
        F.nrSpConflag = false;
        F.Yt = W.Yt[n];
        F.NP = 2;
        F.calAiBi();
        F.it[IT_SSMSTA] = 0;
        F.it[IT_NRSTA] = 0;
        F.it[IT_NRSP] = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) F.Ks[i] = W.Ks[i * ld + n];
        F.it[IT_RR] = W.rrIt[n];
        F.ssmSpConflag = (fl & FL_SSMSP) != 0;
        if (fl & FL_DEFER) {
            F.it[IT_SSMSP] = W.ssmIt[n];
            F.it[IT_NRSP] = W.nrIt[n];
            CoopSplit<NC> C(F);
            C.splitRestarts();
        } else {
            double Se = W.Se[n];
            int iter = W.ssmIt[n];
            CoopSplit<NC> C(F);
            C.splitSSM2Rest(Se, iter);
            F.it[IT_SSMSP] = iter;
            F.splitNR();
            C.splitRestarts();
        }
        const int NP = F.NP;
        int its[3] = {0, 0, 0};
        bool nonconv = false;
        if (lead) splitDone<NC>(F, B, W, n, fl, its, nonconv);
        appendList(W, outNP2, lead && NP == 2, n);
        appendList(W, outNP1, lead && NP == 1, n);
        if (ctr) {
            addCounter(&ctr->iters[IT_SSMSP], (unsigned long long)its[0]);
            addCounter(&ctr->iters[IT_NRSP], (unsigned long long)its[1]);
            addCounter(&ctr->iters[IT_RR], (unsigned long long)its[2]);
            addCounter(&ctr->nonconverged, nonconv ? 1 : 0);
            addCounter(&ctr->twoPhase, (lead && NP == 2) ? 1 : 0);
        }
    }
}

// Stage 4: the rest of FlashFIM / InitFlashFIM for the cells of `list`,
// which end with NPT hydrocarbon phases.  The EoS state the split ended with
// (Aj, Bj, Zj of the final compositions) is re-derived from the stored
// compositions.
template <int NC, int NPT>
__global__ void __launch_bounds__(kFlashBlock) flashFinishKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                 FlashWork W, int init, int list)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int total = W.count[list];
    for (int base = nextChunk(W, list); base < total; base = nextChunk(W, list)) {
    const int idx = base + (threadIdx.x & 63);
    if (idx >= total) continue;
    const int32_t n = W.list[list][idx];
    const int64_t ld = B.ld;
    constexpr int nc = NC + 1;
    FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
    loadFeed<NC>(F, B, W, n);
    F.flagSkip = (W.flags[n] & FL_SKIP) != 0;
    F.NP = NPT;
    if (!init) {
#pragma unroll
        for (int i = 0; i < nc; ++i) F.Ni[i] = B.Ni[i * ld + n];
        F.Nh = W.Nh[n];
    }
    F.calAiBi();
    if (NPT == 1) {
        F.nu[0] = 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) F.x[0][i] = F.zi[i];
    } else {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            F.nu[j] = W.nu[j * ld + n];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                F.x[j][i] = W.x[(j * NC + i) * ld + n];
                F.n[j][i] = W.n[(j * NC + i) * ld + n];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        F.calAjBj(F.Aj[j], F.Bj[j], F.x[j]);
        F.Zj[j] = solEoS(F.Aj[j], F.Bj[j]);
    }
    if (!init) {
#pragma unroll
        for (int j = 0; j < NPT; ++j) F.nu[j] = F.Nh * F.nu[j];
    }
    const double Sw = init ? B.S[2 * ld + n] : 0.0;
    const double Vp = init ? B.rockVp[n] : 0.0;
    FlashKernel<NC>::template finish<NPT>(F, *eos, B, n, init != 0, Sw, Vp);
    }
}

// The finishing stage of the side streams' cells, one launch for both
// phase counts: wavefront chunks of the one-phase lists first, then of the
// two-phase lists.
template <int NC>
__global__ void __launch_bounds__(kFlashBlock) flashFinishSideKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                     FlashWork W)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int l1[3] = {LIST_NP1B, LIST_NP1C, LIST_NP1D};
    const int l2[3] = {LIST_NP2B, LIST_NP2C, LIST_NP2D};
    int c1[3], c2[3], chunks1 = 0, chunks = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c1[q] = W.count[l1[q]];
        c2[q] = W.count[l2[q]];
        chunks1 += (c1[q] + 63) / 64;
    }
    chunks = chunks1;
#pragma unroll
    for (int q = 0; q < 3; ++q) chunks += (c2[q] + 63) / 64;
    const int64_t ld = B.ld;
    constexpr int nc = NC + 1;
    for (int ch = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64); ch < chunks;
         ch += (int)(gridDim.x * blockDim.x / 64)) {
        const bool two = ch >= chunks1;
        int r = two ? ch - chunks1 : ch, list = -1, idx = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int cnt = two ? c2[q] : c1[q];
            const int nch = (cnt + 63) / 64;
            if (list == -1) {
                if (r < nch) {
                    // the chunk is in this list; its lanes past the end idle
                    idx = r * 64 + (int)(threadIdx.x & 63);
                    list = idx < cnt ? (two ? l2[q] : l1[q]) : -2;
                } else {
                    r -= nch;
                }
            }
        }
        if (list < 0) continue;
        const int32_t n = W.list[list][idx];
        FlashCell<NC> F(*eos, &sSq[threadIdx.x]);
        loadFeed<NC>(F, B, W, n);
        F.flagSkip = (W.flags[n] & FL_SKIP) != 0;
#pragma unroll
        for (int i = 0; i < nc; ++i) F.Ni[i] = B.Ni[i * ld + n];
        F.Nh = W.Nh[n];
        F.calAiBi();
        if (!two) {
            F.NP = 1;
            F.nu[0] = 1;
#pragma unroll
            for (int i = 0; i < NC; ++i) F.x[0][i] = F.zi[i];
            F.calAjBj(F.Aj[0], F.Bj[0], F.x[0]);
            F.Zj[0] = solEoS(F.Aj[0], F.Bj[0]);
            F.nu[0] = F.Nh * F.nu[0];
            FlashKernel<NC>::template finish<1>(F, *eos, B, n, false, 0.0, 0.0);
        } else {
            F.NP = 2;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                F.nu[j] = W.nu[j * ld + n];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    F.x[j][i] = W.x[(j * NC + i) * ld + n];
                    F.n[j][i] = W.n[(j * NC + i) * ld + n];
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                F.calAjBj(F.Aj[j], F.Bj[j], F.x[j]);
                F.Zj[j] = solEoS(F.Aj[j], F.Bj[j]);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) F.nu[j] = F.Nh * F.nu[j];
            FlashKernel<NC>::template finish<2>(F, *eos, B, n, false, 0.0, 0.0);
        }
    }
}

template <int NC>
hipError_t launchFlashT(const DevEos* eos, const DevBulk& B, const FlashWork& W, DevFlashCounters* ctr,
                        const double* zin, hipStream_t s, hipEvent_t* ev, const FlashSide* side, bool finish)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = (unsigned)((B.nb + kFlashBlock - 1) / kFlashBlock);
    // the list stages run as work queues over a grid that fills the chip
    // (idle waves leave as soon as the queue is empty)
    const unsigned qgrid = std::min<unsigned>(grid, kQueueBlocks);
    const unsigned sgrid = std::min<unsigned>(grid, kSideBlocks);
    auto mark = [&](int k, hipStream_t st) {
        if (ev) (void)hipEventRecord(ev[k], st);
    };
    const int init = zin ? 1 : 0;
    hipError_t e;
    hipLaunchKernelGGL(flashResetKernel<NC>, dim3(1), dim3(64), 0, s, W, ctr);
    hipLaunchKernelGGL(flashClassifyKernel<NC>, dim3(grid), dim3(kFlashBlock), 0, s, eos, B, W, ctr, zin);
    if (!side) {
        mark(0, s);
        if (W.guess)
            hipLaunchKernelGGL((flashStabilityKernel<NC, true>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, ctr,
                               (int)LIST_SPLIT, (int)LIST_NP1, 0);
        else
            hipLaunchKernelGGL((flashStabilityKernel<NC, false>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W,
                               ctr, (int)LIST_SPLIT, (int)LIST_NP1, 0);
        mark(1, s);
        mark(2, s);
        hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W,
                           (int)LIST_SPLIT, (int)LIST_C1, kSplitBudget0, 1, (int)LIST_NR, 0);
        hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_C1,
                           (int)LIST_C2, kSplitBudget1, 0, (int)LIST_NR, 0);
        hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_C2,
                           -1, -1, 0, (int)LIST_NR, 0);
        hipLaunchKernelGGL(flashSplitNRKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, ctr, -1,
                           (int)LIST_NR, (int)LIST_NP1, (int)LIST_NP2);
        mark(3, s);
        if (!finish) return hipGetLastError();
        mark(4, s);
        hipLaunchKernelGGL((flashFinishKernel<NC, 1>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, init,
                           (int)LIST_NP1);
        hipLaunchKernelGGL((flashFinishKernel<NC, 2>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, init,
                           (int)LIST_NP2);
        mark(5, s);
        return hipGetLastError();
    }
    // Overlapped form (not with a learned guess, nor for InitFlashFIM).
    // Each side list is worked on its side stream while it is short (a
    // warm Newton pass: a few hundred cells) and on the main stream, in the
    // serial order, when it is long (a cold pass puts every cell through the
    // stability test): a kernel pair gated on the list's length on the
    // device, exactly one of which works.
    if (W.guess || zin) return hipErrorInvalidValue;
    if ((e = hipEventRecord(side->classified, s)) != hipSuccess) return e;
    // side stream B: the stability test and the split of the cells it finds unstable
    if ((e = hipStreamWaitEvent(side->sB, side->classified, 0)) != hipSuccess) return e;
    mark(0, side->sB);
    hipLaunchKernelGGL((flashStabilityKernel<NC, false>), dim3(sgrid), dim3(kFlashBlock), 0, side->sB, eos, B, W,
                       ctr, (int)LIST_SPLITB, (int)LIST_NP1B, kSideMaxStab);
    hipLaunchKernelGGL(flashSplitTailKernel<NC>, dim3(sgrid), dim3(kFlashBlock), 0, side->sB, eos, B, W, ctr,
                       (int)LIST_SPLITB, (int)TAIL_FULL, (int)LIST_NP1B, (int)LIST_NP2B, 0);
    mark(1, side->sB);
    if ((e = hipEventRecord(side->doneB, side->sB)) != hipSuccess) return e;
    // main stream: a long stability list, then the slices, the third one bounded
    hipLaunchKernelGGL((flashStabilityKernel<NC, false>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, ctr,
                       (int)LIST_SPLIT, (int)LIST_NP1, -kSideMaxStab);
    mark(2, s);
    hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_SPLIT,
                       (int)LIST_C1, kSplitBudget0, 1, (int)LIST_NR, 0);
    hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_C1,
                       (int)LIST_C2, kSplitBudget1, 0, (int)LIST_NR, 0);
    hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_C2,
                       (int)LIST_TSSM, kSplitBudget2, 0, (int)LIST_NR, 0);
    // a long leftover list stays on the main stream
    hipLaunchKernelGGL(flashSplitSSMKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, (int)LIST_TSSM,
                       -1, -1, 0, (int)LIST_NR, -kSideMaxTail);
    if ((e = hipEventRecord(side->sliced, s)) != hipSuccess) return e;
    // side stream C: a short leftover list runs to its end, one wavefront
    // per cell
    if ((e = hipStreamWaitEvent(side->sC, side->sliced, 0)) != hipSuccess) return e;
    mark(6, side->sC);
    hipLaunchKernelGGL(flashSplitTailCoopKernel<NC>, dim3(kCoopBlocks), dim3(kFlashBlock), 0, side->sC, eos, B,
                       W, ctr, (int)LIST_TSSM, (int)LIST_NP1C, (int)LIST_NP2C, kSideMaxTail);
    mark(7, side->sC);
    if ((e = hipEventRecord(side->doneC, side->sC)) != hipSuccess) return e;
    // main stream: Newton stage; cells that would restart go to list TNR
    hipLaunchKernelGGL(flashSplitNRKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, ctr,
                       (int)LIST_TNR, (int)LIST_NR, (int)LIST_NP1, (int)LIST_NP2);
    mark(3, s);
    if ((e = hipEventRecord(side->newton, s)) != hipSuccess) return e;
    // side stream D: a short restart list, one wavefront per cell
    if ((e = hipStreamWaitEvent(side->sD, side->newton, 0)) != hipSuccess) return e;
    mark(8, side->sD);
    hipLaunchKernelGGL(flashSplitTailCoopKernel<NC>, dim3(kRestartBlocks), dim3(kFlashBlock), 0, side->sD, eos, B,
                       W, ctr, (int)LIST_TNR, (int)LIST_NP1D, (int)LIST_NP2D, kSideMaxRestart);
    mark(9, side->sD);
    if ((e = hipEventRecord(side->doneD, side->sD)) != hipSuccess) return e;
    // main stream: a long restart list, then finish the cells it holds
    hipLaunchKernelGGL(flashSplitTailKernel<NC>, dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, ctr,
                       (int)LIST_TNR, (int)TAIL_RESTART, (int)LIST_NP1, (int)LIST_NP2, -kSideMaxRestart);
    if (!finish) return hipGetLastError();
    mark(4, s);
    hipLaunchKernelGGL((flashFinishKernel<NC, 1>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, 0,
                       (int)LIST_NP1);
    hipLaunchKernelGGL((flashFinishKernel<NC, 2>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, 0,
                       (int)LIST_NP2);
    mark(5, s);
    return hipGetLastError();
}

// Overlapped form, once the caller's stream s has joined the side streams:
// the finishing stage of their cells.
template <int NC>
hipError_t launchFlashSideFinishT(const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = (unsigned)((B.nb + kFlashBlock - 1) / kFlashBlock);
    hipLaunchKernelGGL(flashFinishSideKernel<NC>, dim3(std::min<unsigned>(grid, 256)), dim3(kFlashBlock), 0, s, eos,
                       B, W);
    return hipGetLastError();
}

}  // namespace ocpg
