// IsoT_FIMn finishing stage: everything FlashFIMn / InitFlashFIMn do after
// the phase equilibrium (src/MixtureComp.cpp:575-695, 355-442), then
// IsoT_FIMn::PassFlashValue's stores (src/OCPFluidMethod.cpp:3179-3196).
// The equilibrium stages (classify, stability, split) are the FIM flash's
// own kernels; only this stage differs: the secondary variables are S and
// the phase moles n_ij, so the derivatives are taken at fixed n_ij and the
// flash also hands over its Newton correction (res_n, resPc).
//
// CaldXsdXp03 (:4634-4820) solves one (NP + 1 + NP NC)-square system by
// dgesv for the NC + 3 right-hand sides [dF/dP, dF/dN, F].  Its structure
// is used here instead: the component balance gives dn_1 = b_n - dn_0, the
// fugacity rows then reduce to the symmetric NC x NC system
//     (dlnf_0/dn_0 + dlnf_1/dn_1) dn_0 = (dlnf_1/dn_1) b_n - b_f,
// the same matrix CalVfiVfp_full02 factors, and the saturation rows follow
// explicitly from the phase volumes.  Same solution, NC-square work instead
// of (3 + 2 NC)-square, results within rounding of the reference's.
#pragma once

#include "flash_impl.hpp"

namespace ocpg {

template <int NC>
struct FlashKernelN {
    static constexpr int nc = NC + 1;
    static constexpr int ncol = nc + 1;
    static constexpr int maxRows = 3 + 2 * NC;

    // Inlined into its kernels for NC <= 6; larger instantiations keep it
    // out of line (as the FIM finishing stage does).
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

    // F holds the split's phases (x, nu as fractions, Aj, Bj, Zj and, for
    // two phases, the fugacities fug of the final compositions: what the
    // reference's last CalFugPhiAll left).
    template <int NPT>
    __device__ __forceinline__ static void finishBody(FlashCell<NC>& F, const DevEos& E, const DevBulk& B,
                                                      int64_t cell, bool init, double Sw_in, double Vpore)
    {
        const int64_t ld = B.ld;
        const int64_t n = cell;
        constexpr int NP = NPT;
        const double P = F.P, T = F.T;
        // ---- CalMW, CalVfXiRho, CalViscoLBC (per internal phase) ----
        double MW[2], vC[2], xiC[2], rhoC[2], muC[2], muA[2][5];
        double vf = 0;
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
        // ---- InitFlashFIMn: scale to moles from the pore volume.  Its
        // CopyPhase runs before the scaling (src/MixtureComp.cpp:376-383), so
        // the nj it hands over are the phase fractions ----
        double Nh = F.Nh;
        const double nuFrac[2] = {F.nu[0], NP == 2 ? F.nu[1] : 0.0};
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
        // ---- CalVjpVfpVfn_partial (:2581-2626), internal phase order ----
        const double CgTP = kGasConst * T / P;
        double Zn[2][NC], vji[2][NC], vjp[2], Zp[2];
        double vfP = 0;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            const double aj = F.Aj[j], bj = F.Bj[j], zj = F.Zj[j];
            const double den = dZden(aj, bj, zj);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double tmp = F.aSum1(i, F.x[j]);
                const double An = 2 / F.nu[j] * (tmp - aj);
                const double Bn = 1 / F.nu[j] * (F.Bi[i] - bj);
                Zn[j][i] = dZnum(aj, bj, zj, An, Bn) / den;
                vji[j][i] = CgTP * (zj + F.nu[j] * Zn[j][i]) - E.Vshift[i];
            }
            Zp[j] = dZnum(aj, bj, zj, aj, bj) / P / den;
            vjp[j] = CgTP * F.nu[j] * (Zp[j] - zj / P);
            vfP += vjp[j];
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
        const double vfiW = kCONV1 * bw;
        const double vwp = kCONV1 * NiW * bwp;
        vf += vw;
        vfP += vwp;
        // ---- CalSaturation ----
        double S[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (j < NP) S[lab[j]] = vC[j] / vf;
        S[2] = vw / vf;

        // ---- CaldXsdXpAPI03 (:4495-4632) and CalVfiVfp_full03 (:4822-4852) ----
        double* D = B.dSec;
        double* R = B.resN;
        const double vf2 = vf * vf;
        double vfi[nc];
#pragma unroll
        for (int i = 0; i < NC; ++i) vfi[i] = 0;
        vfi[NC] = vfiW;
        double resPc = 0;
        if (NP == 1) {
            const double vjh = vC[0];
            double r0[ncol];
            r0[0] = ((vfP - vwp) * vf - vfP * vjh) / vf2;
#pragma unroll
            for (int m = 0; m < NC; ++m) r0[1 + m] = vji[0][m] * vw / vf2;
            r0[1 + NC] = -vfiW * vjh / vf2;
#pragma unroll
            for (int c = 0; c < ncol; ++c) {
                D[(0 * ncol + c) * ld + n] = r0[c];
                D[(1 * ncol + c) * ld + n] = -r0[c];
            }
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int c = 0; c < ncol; ++c) D[((2 + i) * ncol + c) * ld + n] = (c == 1 + i) ? 1.0 : 0.0;
#pragma unroll
            for (int r = 0; r < kNP * ncol; ++r) R[(int64_t)r * ld + n] = 0.0;
#pragma unroll
            for (int i = 0; i < NC; ++i) vfi[i] = vji[0][i];
        } else {
            // dlnf/dn of both phases (CalFugNAll at the stored Zn) and
            // dlnf/dP (CalFugPAll at the partial-derivative Zp)
            double A[NC * NC];
            double dn[NC * (NC + 2)];  // columns: P, N_1..N_NC, the residual
            {
                double f1[FlashCell<NC>::NT];
                double Zt[NC];
                F.calFugN(0, f1, Zt);
#pragma unroll
                for (int i = 0; i < NC; ++i)
#pragma unroll
                    for (int m = 0; m <= i; ++m) A[i * NC + m] = f1[FlashCell<NC>::tri(i, m)];
                F.calFugN(1, f1, Zt);
                // component balance residual b_n (CaldXsdXp03's resN)
                double rn[NC];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    double v = F.Ni[i];
                    v -= F.x[0][i] * F.nu[0];
                    v -= F.x[1][i] * F.nu[1];
                    rn[i] = v;
                }
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    double s = 0;
#pragma unroll
                    for (int m = 0; m < NC; ++m) {
                        const double v1 = (m <= i) ? f1[FlashCell<NC>::tri(i, m)] : f1[FlashCell<NC>::tri(m, i)];
                        s += v1 * rn[m];
                        if (m <= i) {
                            const double sum = A[i * NC + m] + v1;
                            A[i * NC + m] = sum;
                            A[m * NC + i] = sum;
                        }
                        dn[(m + 1) * NC + i] = v1;
                    }
                    dn[(NC + 1) * NC + i] = s - (F.fug[0][i] - F.fug[1][i]);
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const double aj = F.Aj[j], bj = F.Bj[j], zj = F.Zj[j];
                const double Ap = aj / P;
                const double Bp = bj / P;
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
                    const double fp = Cp / C + Dp + Ei / G * Gp;
                    dn[i] = (j == 0) ? -fp : dn[i] + fp;
                }
            }
            // Symmetric equilibration D A D (D = |diag A|^-1/2) before the
            // LU: A is dominated by its 1/n_ij diagonal, which spans many
            // orders of magnitude between the main and the trace components;
            // unscaled, partial pivoting loses up to ~1e-9 of the
            // heaviest component's column in near-critical cells, where the
            // reference's full-system solve keeps ~1e-15.  Scaled, the
            // solution is as accurate as the reference's.
            double Dg[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double a = fabs(A[i * NC + i]);
                Dg[i] = (a > 0 && isfinite(a)) ? 1.0 / sqrt(a) : 1.0;
            }
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int m = 0; m < NC; ++m) A[i * NC + m] *= Dg[i] * Dg[m];
#pragma unroll
            for (int c = 0; c < NC + 2; ++c)
#pragma unroll
                for (int i = 0; i < NC; ++i) dn[c * NC + i] *= Dg[i];
            luSolve<NC, NC + 2>(A, dn);
#pragma unroll
            for (int c = 0; c < NC + 2; ++c)
#pragma unroll
                for (int i = 0; i < NC; ++i) dn[c * NC + i] *= Dg[i];
            // the rows, column by column: c = 0 P, 1..NC N_k, NC + 1 the
            // water moles (no n_ij change), NC + 2 the residual (res_n)
            const int base0 = 3 + lab[0] * NC;
            const int base1 = 3 + lab[1] * NC;
#pragma unroll
            for (int c = 0; c <= ncol; ++c) {
                double d0[NC], d1[NC];
                const int dc = (c <= NC) ? c : NC + 1;  // dn column (unused for N_w)
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double a = (c == NC + 1) ? 0.0 : dn[dc * NC + i];
                    double bn = 0.0;
                    if (c >= 1 && c <= NC) bn = (i == c - 1) ? 1.0 : 0.0;
                    d0[i] = a;
                    d1[i] = (c == NC + 1) ? 0.0 : bn - a;
                }
                if (c == ncol) {
                    // residual column: b_n is the balance residual rn = F.Ni - x0 nu0 - x1 nu1
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        double v = F.Ni[i];
                        v -= F.x[0][i] * F.nu[0];
                        v -= F.x[1][i] * F.nu[1];
                        d1[i] = v - d0[i];
                    }
                }
                double dV0 = 0, dV1 = 0;
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    dV0 += vji[0][k] * d0[k];
                    dV1 += vji[1][k] * d1[k];
                }
                // b_s of this column (zero for the N_k columns and, since
                // S_j = v_j / vf exactly, for the residual)
                double bs0 = 0, bs1 = 0, bsw = 0;
                if (c == 0) {
                    bs0 = (vfP * vC[0] - vjp[0] * vf) / vf2;
                    bs1 = (vfP * vC[1] - vjp[1] * vf) / vf2;
                    bsw = (vfP * vw - vwp * vf) / vf2;
                } else if (c == NC + 1) {
                    bs0 = vfiW * vC[0] / vf2;
                    bs1 = vfiW * vC[1] / vf2;
                    bsw = vfiW * (vw - vf) / vf2;
                } else if (c == ncol) {
                    bs0 = S[lab[0]] - vC[0] / vf;
                    bs1 = S[lab[1]] - vC[1] / vf;
                    bsw = S[2] - vw / vf;
                }
                const double dS0 = (dV0 * (vf - vC[0]) - dV1 * vC[0]) / vf2 - bs0;
                const double dS1 = (dV1 * (vf - vC[1]) - dV0 * vC[1]) / vf2 - bs1;
                const double dSw = -(dV0 + dV1) * vw / vf2 - bsw;
                if (c < ncol) {
                    D[(lab[0] * ncol + c) * ld + n] = dS0;
                    D[(lab[1] * ncol + c) * ld + n] = dS1;
                    D[(2 * ncol + c) * ld + n] = dSw;
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        D[((base0 + i) * ncol + c) * ld + n] = d0[i];
                        D[((base1 + i) * ncol + c) * ld + n] = d1[i];
                    }
                    // CalVfiVfp_full03: over the phases in label order
                    if (c <= NC) {
                        const double* dA = lab[0] == 0 ? d0 : d1;
                        const double* vA = lab[0] == 0 ? vji[0] : vji[1];
                        const double* dB = lab[0] == 0 ? d1 : d0;
                        const double* vB = lab[0] == 0 ? vji[1] : vji[0];
                        double acc = (c == 0) ? vfP : vfi[c - 1];
#pragma unroll
                        for (int i = 0; i < NC; ++i) acc += vA[i] * dA[i];
#pragma unroll
                        for (int i = 0; i < NC; ++i) acc += vB[i] * dB[i];
                        if (c == 0)
                            vfP = acc;
                        else
                            vfi[c - 1] = acc;
                    }
                } else {
                    R[(int64_t)lab[0] * ld + n] = dS0;
                    R[(int64_t)lab[1] * ld + n] = dS1;
                    R[(int64_t)2 * ld + n] = dSw;
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        R[(int64_t)(base0 + i) * ld + n] = d0[i];
                        R[(int64_t)(base1 + i) * ld + n] = d1[i];
                    }
#pragma unroll
                    for (int r = maxRows; r < kNP * ncol; ++r) R[(int64_t)r * ld + n] = 0.0;
                    // resPc: the reference pairs the b-th stored block (label
                    // b) with vji of internal phase b (:4612-4620)
                    const double* blk0 = lab[0] == 0 ? d0 : d1;  // label 0's corrections
                    const double* blk1 = lab[0] == 0 ? d1 : d0;
#pragma unroll
                    for (int i = 0; i < NC; ++i) resPc += vji[0][i] * blk0[i];
#pragma unroll
                    for (int i = 0; i < NC; ++i) resPc += vji[1][i] * blk1[i];
                }
            }
        }

        // ---- stores shared by every cell (PassFlashValue) ----
        B.phaseNum[n] = NP + 1;
        B.Nt[n] = Nt;
        B.vf[n] = vf;
        B.vfP[n] = vfP;
#pragma unroll
        for (int j = 0; j < 3; ++j) B.S[j * ld + n] = S[j];
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
        B.resPc[n] = resPc;
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (j < NP) B.nj[(int64_t)lab[j] * ld + n] = init ? nuFrac[j] : F.nu[j];
        B.nj[(int64_t)2 * ld + n] = NiW;
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

        // ---- CalXiPn_partial, CalRhoPn_partial, CalMuPnLBC_partial
        // (:2628-2757): derivatives by n_ij at fixed P, by P at fixed n_ij ----
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            const int s = lab[j];
            const double zj = F.Zj[j];
            const double nuj = F.nu[j];
            const double mxi = -xiC[j] * xiC[j] * CgTP;
            double xix[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) xix[i] = mxi * Zn[j][i];
            const double xiP = mxi * (Zp[j] - zj / P);
            const double rhoP = xiP * MW[j];
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
            const double xTm56 = pow(xTj, -5.0 / 6);
            B.rho[s * ld + n] = rhoC[j];
            B.xi[s * ld + n] = xiC[j];
            B.mu[s * ld + n] = muC[j];
            B.rhoP[s * ld + n] = rhoP;
            B.xiP[s * ld + n] = xiP;
            B.muP[s * ld + n] = muP;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const double derxTj = (E.Tc[k] - xTj) / nuj;
                const double derxPj = (E.Pc[k] - xPj) / nuj;
                const double derMWj = (E.MWC[k] - MW[j]) / nuj;
                double der3J = 0;
                double der4J = E.sqrtMWi[k];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double val1IJ = E.muAux1I[i] / sMW;
                    // the reference's -(1 / 2) is an integer zero
                    const double der1IJ = 0 * E.muAux1I[i] * pMW15 * derMWj;
                    const double tmp = visT[i];
                    const double val2IJ = tmp / val1IJ;
                    const double der2IJ = -tmp * der1IJ / (val1IJ * val1IJ);
                    der3J += E.sqrtMWi[i] *
                             (F.x[j][i] * der2IJ + ((i == k ? 1.0 : 0.0) - F.x[j][i]) * val2IJ / nuj);
                    der4J -= F.x[j][i] * E.sqrtMWi[i];
                }
                der4J /= nuj;
                const double der6J = 5.4402 *
                                     (1.0 / 6 * xTm56 * derxTj -
                                      xT16[j] * (0.5 / MW[j] * derMWj + 2.0 / 3 / xPj * derxPj)) /
                                     (sMW * xP23[j]);
                const double der7J = xix[k] * xVj + (E.Vcvis[k] - xVj) * xiC[j] / nuj;
                const double der8J = der7J * poly;
                const double mux = (der3J * muA[j][1] - muA[j][0] * der4J) / (muA[j][1] * muA[j][1]) +
                                   1E-4 * (4 * a4c * der8J * muA[j][2] - (a4q - 1) * der6J) /
                                       (muA[j][2] * muA[j][2]);
                const double rhox = xix[k] * MW[j] + xiC[j] / nuj * (E.MWC[k] - MW[j]);
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

        // ---- CalSkipForNextStep (as FlashFIM's) ----
        if constexpr (NPT == 1) if (F.flagSkip && F.ftype == 0) {
            const double aj = F.Aj[0], bj = F.Bj[0], zj = F.Zj[0], nu0 = F.nu[0];
            const double* xj = F.x[0];
            double An[NC], Bn[NC], Znk[NC], tI[NC];
            const double den = dZden(aj, bj, zj);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                tI[i] = F.aSum1(i, xj);
                An[i] = 2 / nu0 * (tI[i] - aj);
                Bn[i] = 1 / nu0 * (F.Bi[i] - bj);
                Znk[i] = dZnum(aj, bj, zj, An[i], Bn[i]) / den;
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
                    const double Cnk = (Bn[k] - Znk[k]) / ((zj - bj) * (zj - bj));
                    const double Dnk = F.Bi[i] / bj * (Znk[k] - (F.Bi[k] - bj) * (zj - 1) / (nu0 * bj));
                    const double Gnk = (kDelta1 - kDelta2) / ((zj + kDelta2 * bj) * (zj + kDelta2 * bj)) *
                                       (Bn[k] * zj - Znk[k] * bj);
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

// The IsoT_FIMn finishing stage for the cells of `list`, which end with NPT
// hydrocarbon phases (the FIM stage's flashFinishKernel, with the final
// fugacities re-derived for two phases).
template <int NC, int NPT>
__global__ void __launch_bounds__(kFlashBlock) flashFinishNKernel(const DevEos* __restrict__ eos, DevBulk B,
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
            F.calAjBj(F.Aj[0], F.Bj[0], F.x[0]);
            F.Zj[0] = solEoS(F.Aj[0], F.Bj[0]);
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                F.nu[j] = W.nu[j * ld + n];
#pragma unroll
                for (int i = 0; i < NC; ++i) F.x[j][i] = W.x[(j * NC + i) * ld + n];
            }
            // Aj, Bj, Zj and the fugacities of the final compositions
            F.calFugPhiAll();
        }
        if (!init) {
#pragma unroll
            for (int j = 0; j < NPT; ++j) F.nu[j] = F.Nh * F.nu[j];
        }
        const double Sw = init ? B.S[2 * ld + n] : 0.0;
        const double Vp = init ? B.rockVp[n] : 0.0;
        FlashKernelN<NC>::template finish<NPT>(F, *eos, B, n, init != 0, Sw, Vp);
    }
}

// The IsoT_FIMn finishing stage of the overlapped flash's side-stream cells
// (after the join) that end with NPT hydrocarbon phases: the wavefront
// chunks of lists B, C and D in turn, each cell as flashFinishNKernel
// finishes it.  One instance per phase count keeps each kernel's register
// use that of the main stream's finishing kernel.
template <int NC, int NPT>
__global__ void __launch_bounds__(kFlashBlock) flashFinishSideNKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                                      FlashWork W)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int lists[3] = {NPT == 1 ? LIST_NP1B : LIST_NP2B, NPT == 1 ? LIST_NP1C : LIST_NP2C,
                          NPT == 1 ? LIST_NP1D : LIST_NP2D};
    int cnt[3], chunks = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        cnt[q] = W.count[lists[q]];
        chunks += (cnt[q] + 63) / 64;
    }
    const int64_t ld = B.ld;
    constexpr int nc = NC + 1;
    for (int ch = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64); ch < chunks;
         ch += (int)(gridDim.x * blockDim.x / 64)) {
        int r = ch, list = -1, idx = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int nch = (cnt[q] + 63) / 64;
            if (list == -1) {
                if (r < nch) {
                    idx = r * 64 + (int)(threadIdx.x & 63);
                    list = idx < cnt[q] ? lists[q] : -2;
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
        F.NP = NPT;
#pragma unroll
        for (int i = 0; i < nc; ++i) F.Ni[i] = B.Ni[i * ld + n];
        F.Nh = W.Nh[n];
        F.calAiBi();
        if (NPT == 1) {
            F.nu[0] = 1;
#pragma unroll
            for (int i = 0; i < NC; ++i) F.x[0][i] = F.zi[i];
            F.calAjBj(F.Aj[0], F.Bj[0], F.x[0]);
            F.Zj[0] = solEoS(F.Aj[0], F.Bj[0]);
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                F.nu[j] = W.nu[j * ld + n];
#pragma unroll
                for (int i = 0; i < NC; ++i) F.x[j][i] = W.x[(j * NC + i) * ld + n];
            }
            F.calFugPhiAll();
        }
#pragma unroll
        for (int j = 0; j < NPT; ++j) F.nu[j] = F.Nh * F.nu[j];
        FlashKernelN<NC>::template finish<NPT>(F, *eos, B, n, false, 0.0, 0.0);
    }
}

template <int NC>
hipError_t launchFlashSideFinishNT(const DevEos* eos, const DevBulk& B, const FlashWork& W, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = std::min<unsigned>((unsigned)((B.nb + kFlashBlock - 1) / kFlashBlock), 256);
    hipLaunchKernelGGL((flashFinishSideNKernel<NC, 1>), dim3(grid), dim3(kFlashBlock), 0, s, eos, B, W);
    hipLaunchKernelGGL((flashFinishSideNKernel<NC, 2>), dim3(grid), dim3(kFlashBlock), 0, s, eos, B, W);
    return hipGetLastError();
}

template <int NC>
hipError_t launchFlashFinishNT(const DevEos* eos, const DevBulk& B, const FlashWork& W, int init, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = (unsigned)((B.nb + kFlashBlock - 1) / kFlashBlock);
    const unsigned qgrid = std::min<unsigned>(grid, kQueueBlocks);
    hipLaunchKernelGGL((flashFinishNKernel<NC, 1>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, init,
                       (int)LIST_NP1);
    hipLaunchKernelGGL((flashFinishNKernel<NC, 2>), dim3(qgrid), dim3(kFlashBlock), 0, s, eos, B, W, init,
                       (int)LIST_NP2);
    return hipGetLastError();
}

}  // namespace ocpg
