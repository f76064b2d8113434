// Well kernels, one instantiation per hydrocarbon component count (included
// by flash_nc.hip, since the producer's surface weight needs the flash):
//
//   wellFluxKernel     : Well::CalTrans (src/Well.cpp:164-202), Well::CalFlux
//                        with ReCalXi = false (:204-258), Well::CalProdWeight
//                        (:829-862 -> MixtureComp::CalProdWeight, src/
//                        MixtureComp.cpp:830-853) and the well rows of
//                        IsoT_FIM::CalRes (src/OCPFluidMethod.cpp:1296-1352).
//                        One wavefront per well, a lane per perforation.
//   wellAssembleKernel : the per-perforation blocks of AssembleMatInjWellsNew /
//                        AssembleMatProdWellsNew (src/OCPFluidMethod.cpp:
//                        2438-2733).  One wavefront per perforation.
//   wellDiagKernel     : the AddDiag calls of those functions, in the
//                        reference's order: each perforated cell adds its
//                        perforations' blocks to its diagonal (wells in order,
//                        perforations in order), each well row sums its own.
#pragma once

#include "fim_kernels.hpp"
#include "flash_device.hpp"
#include "well_state.hpp"

namespace ocpg {

constexpr int kWellBlock = 64;

// OCPTable::Eval_All on PVTW column 0 (src/OCPTable.cpp:47-86), values only.
__device__ __forceinline__ void pvtwValues(const DevEos& E, double P, double* pw)
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
        for (int q = 0; q < 5; ++q) pw[q] = E.pvtw[r][q];
    } else {
        for (int q = 0; q < 5; ++q) {
            const double cw = (E.pvtw[seg + 1][q] - E.pvtw[seg][q]) / (E.pvtw[seg + 1][0] - E.pvtw[seg][0]);
            pw[q] = E.pvtw[seg][q] + cw * (P - E.pvtw[seg][0]);
        }
    }
}

// MixtureComp::Flash (src/MixtureComp.cpp:167-193) of the moles Ni (nc,
// water last) at P (psia) and TF (F): a cold flash (ftype 0, no previous
// phase count), water from PVTW.  Per phase (oil, gas, water): whether it
// exists, its volume vj, molar density xi and mass density rho; a phase the
// flash does not produce gets zeros.  Nt is the total moles.
struct FlashPhases {
    bool ex[3];
    double vj[3], xi[3], rho[3];
    double Nt;
};

template <int NC>
__device__ FlashPhases flashPhases(const DevEos& E, double* sqLds, double P0, double TF, const double* Ni)
{
    FlashCell<NC> F(E, sqLds);
    F.P = P0;
    F.T = TF + kCONV5;
    double Nh = 0;
    for (int i = 0; i < NC; ++i) Nh += fabs(Ni[i]);
    F.Nh = Nh;
    for (int i = 0; i < NC; ++i) F.zi[i] = Ni[i] / Nh;
    F.ftype = 0;
    F.lNP = 0;
    F.NP = 1;
    for (int q = 0; q < 5; ++q) F.it[q] = 0;
    F.phaseEquilibrium();
    const int NP = F.NP;
    const double P = F.P, T = F.T;
    FlashPhases out;
    for (int j = 0; j < 3; ++j) {
        out.ex[j] = false;
        out.vj[j] = out.xi[j] = out.rho[j] = 0;
    }
    double MW[2], vC[2], xiC[2], rhoC[2];
    for (int j = 0; j < NP; ++j) {
        F.nu[j] = Nh * F.nu[j];
        MW[j] = 0;
        for (int i = 0; i < NC; ++i) MW[j] += F.x[j][i] * E.MWC[i];
    }
    // CalVfXiRho (src/MixtureComp.cpp:1158-1176)
    for (int j = 0; j < NP; ++j) {
        double tmp = F.Zj[j] * kGasConst * T / P;
        for (int i = 0; i < NC; ++i) tmp -= F.x[j][i] * E.Vshift[i];
        vC[j] = tmp * F.nu[j];
        xiC[j] = 1 / tmp;
        rhoC[j] = MW[j] * xiC[j];
    }
    // IdentifyPhase (src/MixtureComp.cpp:2352-2384) + CopyPhase (:2386-2398)
    int label[2] = {0, 1};
    if (NP == 1) {
        double A = 0, B = 0;
        for (int i = 0; i < NC; ++i) {
            A += F.x[0][i] * E.Vc[i] * E.Tc[i];
            B += F.x[0][i] * E.Vc[i];
        }
        const double Tc = A / B;
        label[0] = T > Tc ? 1 : 0;
    } else if (!(MW[0] > MW[1])) {
        label[0] = 1;
        label[1] = 0;
    }
    for (int j = 0; j < NP; ++j) {
        const int j1 = label[j];
        out.ex[j1] = true;
        out.vj[j1] = vC[j];
        out.xi[j1] = xiC[j];
        out.rho[j1] = rhoC[j];
    }
    double pw[5];
    pvtwValues(E, P, pw);
    const double bw = pw[1] * (1 - pw[2] * (P - pw[0]));
    out.ex[2] = true;
    out.xi[2] = 1 / (kCONV1 * bw);
    out.rho[2] = E.stdRhoW / bw;
    out.vj[2] = kCONV1 * Ni[NC] * bw;
    out.Nt = Nh + Ni[NC];
    return out;
}

// MixtureComp::CalProdWeight (src/MixtureComp.cpp:830-853): the stream's
// volume per mole in the rate units of prodPhase after the surface flash.
template <int NC>
__device__ double surfaceProdWeight(const DevEos& E, double* sqLds, double Psurf, double Tsurf, const double* Ni,
                                    const double* prodPhase)
{
    const FlashPhases f = flashPhases<NC>(E, sqLds, Psurf, Tsurf, Ni);
    const double qt = f.Nt;
    double factor[3];
    factor[0] = f.vj[0] / qt / kCONV1;
    factor[1] = f.vj[1] / qt / 1000;
    factor[2] = f.xi[2] * f.vj[2] / qt;
    double tmp = 0;
    for (int i = 0; i < 3; ++i) tmp += factor[i] * prodPhase[i];
    return tmp;
}

// MixtureComp::CalProdRate (src/MixtureComp.cpp:855-865): the stream's
// surface rates (stb oil, Mscf gas, stb water).
template <int NC>
__device__ void surfaceProdRate(const DevEos& E, double* sqLds, double Psurf, double Tsurf, const double* Ni,
                                double* rate)
{
    const FlashPhases f = flashPhases<NC>(E, sqLds, Psurf, Tsurf, Ni);
    rate[0] = f.vj[0] / kCONV1;
    rate[1] = f.vj[1] / 1000;
    rate[2] = f.vj[2] * f.xi[2];
}

// MixtureComp::XiPhase / RhoPhase (src/MixtureComp.cpp:711-768): molar and
// mass density of the single phase of feed zi (nc entries) at P, TF; phase 2
// is water (PVTW), otherwise the hydrocarbon feed's one EoS root.
template <int NC>
__device__ void phaseDensity(const DevEos& E, double* sqLds, double P, double TF, const double* zin, int phase,
                             double* xiOut, double* rhoOut)
{
    if (phase == 2) {
        double pw[5];
        pvtwValues(E, P, pw);
        const double bw = pw[1] * (1 - pw[2] * (P - pw[0]));
        *xiOut = 1 / (kCONV1 * bw);
        *rhoOut = E.stdRhoW / bw;
        return;
    }
    FlashCell<NC> F(E, sqLds);
    F.P = P;
    F.T = TF + kCONV5;
    for (int i = 0; i < NC; ++i) F.zi[i] = zin[i];
    F.calAiBi();
    double A, B;
    F.calAjBj(A, B, F.zi);
    const double Z = solEoS(A, B);
    double v = Z * kGasConst * F.T / F.P;
    for (int i = 0; i < NC; ++i) v -= F.zi[i] * E.Vshift[i];
    const double xi = 1 / v;
    double MW = 0;
    for (int i = 0; i < NC; ++i) MW += F.zi[i] * E.MWC[i];
    *xiOut = xi;
    *rhoOut = MW * xi;
}

template <int NC>
__global__ void __launch_bounds__(kWellBlock) wellFluxKernel(const DevEos* __restrict__ eos, DevBulk B, FimDevice F,
                                                            WellDevice W)
{
    constexpr int nc = NC + 1;
    constexpr int len = nc + 1;
    constexpr int np = 3;
    constexpr double kTiny = 1E-8;  // include/OCPConst.hpp:38
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t ld = B.ld;
    double* qW = W.wellQi + (int64_t)w * nc;
    if (!W.open[w]) {
        if (lane < nc) qW[lane] = 0;
        if (lane == 0) {
            W.prodWeight[w] = 0;
            W.wellRelRes[w] = -1;
        }
        return;
    }
    const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
    const bool inj = W.type[w] == 0;
    const double bhp = W.bhp[w];
    for (int p = p0 + lane; p < p1; p += kWellBlock) {
        const int64_t k = W.perfCell[p];
        const double temp = kCONV1 * W.perfWI[p] * W.perfMult[p];
        double tj[np];
        double transINJ = 0;
        for (int j = 0; j < np; ++j) {
            tj[j] = 0;
            if (B.phaseExist[j * ld + k]) {
                tj[j] = temp * B.kr[j * ld + k] / B.mu[j * ld + k];
                transINJ += tj[j];
            }
            W.perfTransj[(int64_t)p * np + j] = tj[j];
        }
        const double Pperf = bhp + W.perfDG[p];
        W.perfP[p] = Pperf;
        double* q = W.perfQi + (int64_t)p * nc;
        if (inj) {
            if (W.unweight[w]) transINJ = W.perfWI[p];
            const double dP = B.P[k] - Pperf;
            const double qt = transINJ * dP;
            for (int i = 0; i < nc; ++i) q[i] = qt * W.perfXi[p] * W.injZi[(int64_t)w * nc + i];
        } else {
            double qi[nc];
            for (int i = 0; i < nc; ++i) qi[i] = 0;
            for (int j = 0; j < np; ++j) {
                if (!B.phaseExist[j * ld + k]) continue;
                const double dP = B.Pj[j * ld + k] - Pperf;
                const double qj = tj[j] * dP;
                const double xi = B.xi[j * ld + k];
                for (int i = 0; i < nc; ++i) qi[i] += qj * xi * B.xij[(j * nc + i) * ld + k];
            }
            for (int i = 0; i < nc; ++i) q[i] = qi[i];
        }
    }
    __syncthreads();
    if (lane != 0) return;
    double qs[nc];
    for (int i = 0; i < nc; ++i) qs[i] = 0;
    for (int p = p0; p < p1; ++p)
        for (int i = 0; i < nc; ++i) qs[i] += W.perfQi[(int64_t)p * nc + i];
    for (int i = 0; i < nc; ++i) qW[i] = qs[i];
    const bool rate = W.mode[w] != 0;
    double weight = 0;
    if (!inj && rate) {
        double qt = 0;
        bool flag = true;
        for (int i = 0; i < nc; ++i) {
            qt += qs[i];
            if (qs[i] < 0) flag = false;
        }
        double Nin[nc];
        if (qt > kTiny && flag) {
            for (int i = 0; i < nc; ++i) Nin[i] = qs[i];
        } else {
            for (int i = 0; i < nc; ++i) Nin[i] = 0;
            for (int p = p0; p < p1; ++p) {
                const int64_t n = W.perfCell[p];
                for (int j = 0; j < np; ++j) {
                    if (!B.phaseExist[j * ld + n]) continue;
                    const double tjp = W.perfTransj[(int64_t)p * np + j];
                    const double xi = B.xi[j * ld + n];
                    for (int i = 0; i < nc; ++i) Nin[i] += tjp * xi * B.xij[(j * nc + i) * ld + n];
                }
            }
        }
        weight = surfaceProdWeight<NC>(*eos, &sSq[0], W.Psurf, W.Tsurf, Nin, W.prodPW + (int64_t)w * 3);
    }
    W.prodWeight[w] = weight;
    // the well's equation (CalRes), stored negated like every rhs entry
    double r;
    double rel = -1;
    if (!rate) {
        r = bhp - W.ctrlBHP[w];
    } else if (inj) {
        r = W.ctrlRate[w];
        for (int i = 0; i < nc; ++i) r += qs[i];
        rel = fabs(r / W.ctrlRate[w]);
    } else {
        r = -W.ctrlRate[w];
        for (int i = 0; i < nc; ++i) r += qs[i] * weight;
        rel = fabs(r / W.ctrlRate[w]);
    }
    W.wellRelRes[w] = rel;
    double* res = F.res + (F.nb + W.row[w]) * len;
    res[0] = -r;
    for (int e = 1; e < len; ++e) res[e] = -0.0;
}

// Per perforation: dQ/dXp and dQ/dXs of the perforated cell (own cell only;
// a perforation's flow is always upwinded at its cell), condensed with the
// cell's dSec_dPri.  With W.fimn, IsoT_FIMn's (src/OCPFluidMethod.cpp:
// 3524-3845): a producer's n_ij columns gain -transIJ dP / nj and its
// diagonal term becomes transj xi dP / nj, and the perforation's
// right-hand-side term -dQdXsB res_n (cells with two hydrocarbon phases)
// goes to W.perfRhs.
template <int NC>
__global__ void __launch_bounds__(kWellBlock) wellAssembleKernel(DevBulk B, FimDevice F, WellDevice W, double dt)
{
    constexpr int nc = NC + 1;
    constexpr int ncol = nc + 1;
    constexpr int bsize = ncol * ncol;
    constexpr int np = 3;
    constexpr int maxRows = 3 + 2 * NC;
    constexpr int S = maxRows + 1;
    constexpr int perLane = (bsize + kWellBlock - 1) / kWellBlock;
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t ld = B.ld;
    // perforation -> well
    int w = 0;
    while (W.perfStart[w + 1] <= p) ++w;
    if (!W.open[w]) return;
    const int64_t n = W.perfCell[p];
    const bool inj = W.type[w] == 0;
    const bool rate = W.mode[w] != 0;
    const double bhp = W.bhp[w], dG = W.perfDG[p];

    __shared__ double sSec[maxRows * ncol];
    __shared__ double sDF[nc * S];
    __shared__ double sM[bsize];
    __shared__ double sDQW[nc];

    const int rowsOwn = min(max(B.bRowSizedSdP[n], 0), maxRows);
    for (int q = lane; q < rowsOwn * ncol; q += kWellBlock) sSec[q] = B.dSec[(int64_t)q * ld + n];
    bool ex[np], sd[np];
    int pv[np];
    int nsat = 0;
    for (int j = 0; j < np; ++j) {
        ex[j] = B.phaseExist[j * ld + n];
        sd[j] = B.pSderExist[j * ld + n];
        pv[j] = min(max(B.pVnumCom[j * ld + n], 0), NC);
        nsat += sd[j] ? 1 : 0;
    }
    const int ncolB = min(nsat + pv[0] + pv[1] + pv[2], maxRows);
    const double* tj = W.perfTransj + (int64_t)p * np;
    const double WI = W.perfWI[p], mult = W.perfMult[p];
    const double pxi = W.perfXi[p];

    for (int q = lane; q < nc * (ncolB + 1); q += kWellBlock) {
        const int i = q / (ncolB + 1);
        const int col = q % (ncolB + 1) - 1;  // -1: pressure column
        int j1S = -1, jX = -1, kX = -1;
        if (col >= 0 && col < nsat) {
            int s = 0;
            for (int j1 = 0; j1 < np; ++j1)
                if (sd[j1]) {
                    if (s == col) j1S = j1;
                    s++;
                }
        } else if (col >= nsat) {
            int start = nsat;
            for (int j = 0; j < np; ++j) {
                if (col >= start && col < start + pv[j]) {
                    jX = j;
                    kX = col - start;
                }
                start += pv[j];
            }
        }
        double v = 0, vw = 0;
        for (int j = 0; j < np; ++j) {
            if (!ex[j]) continue;
            const double mu = B.mu[j * ld + n];
            if (inj) {
                const double zi = W.injZi[(int64_t)w * nc + i];
                const double dP = B.P[n] - bhp - dG;
                const double transIJ = tj[j] * pxi * zi;
                if (col < 0) {
                    v += transIJ * (1 - dP * B.muP[j * ld + n] / mu);
                    vw += -transIJ;
                } else if (j1S >= 0) {
                    v += kCONV1 * WI * mult * pxi * zi * B.dKr_dS[(j * 3 + j1S) * ld + n] * dP / mu;
                } else if (jX == j) {
                    v += -transIJ * dP / mu * B.mux[(j * nc + kX) * ld + n];
                }
            } else {
                const double dP = B.Pj[j * ld + n] - bhp - dG;
                const double xi = B.xi[j * ld + n];
                const double xij = B.xij[(j * nc + i) * ld + n];
                const double transIJ = tj[j] * xi * xij;
                if (col < 0) {
                    v += transIJ * (1 - dP * B.muP[j * ld + n] / mu) + dP * tj[j] * xij * B.xiP[j * ld + n];
                    vw += -transIJ;
                } else if (j1S >= 0) {
                    double tmp = kCONV1 * WI * mult * dP / mu * xi * xij * B.dKr_dS[(j * 3 + j1S) * ld + n];
                    tmp += transIJ * B.dPcj_dS[(j * 3 + j1S) * ld + n];
                    v += tmp;
                } else if (jX == j) {
                    double tmp = dP * tj[j] * xij *
                                 (B.xix[(j * nc + kX) * ld + n] - xi / mu * B.mux[(j * nc + kX) * ld + n]);
                    if (W.fimn) tmp -= transIJ * dP / B.nj[j * ld + n];
                    v += tmp;
                    if (kX == i) v += W.fimn ? tj[j] * xi * dP / B.nj[j * ld + n] : tj[j] * xi * dP;
                }
            }
        }
        sDF[i * S + col + 1] = v;
        if (col < 0) sDQW[i] = vw;
    }
    __syncthreads();

    // M = dQdXpB + dQdXsB * dSec (rows 1..nc; row 0 is zero)
    for (int t = 0; t < perLane; ++t) {
        const int e = lane + t * kWellBlock;
        if (e < bsize) {
            const int r = e / ncol, cc = e % ncol;
            double v = 0;
            if (r > 0) {
                const double* rw = &sDF[(r - 1) * S];
                v = (cc == 0) ? rw[0] : 0.0;
                for (int kq = 0; kq < ncolB; ++kq) v = v + rw[1 + kq] * sSec[kq * ncol + cc];
            }
            sM[e] = v;
            W.perfBlock[(int64_t)p * bsize + e] = dt * v;
            // cell -> well block: dt * dQdXpW
            const double bw = (r > 0 && cc == 0) ? sDQW[r - 1] : 0.0;
            F.val[(int64_t)W.slotBW[p] * bsize + e] = dt * bw;
        }
    }
    if (lane < nc) W.perfDQW[(int64_t)p * nc + lane] = sDQW[lane];
    if (W.fimn && lane < ncol) {
        // DaAxpby(ncol, ncolB, -1, dQdXsB, res_n, 1, tmpb): row 0 of dQdXsB is zero
        double y = 0.0;
        if (lane > 0 && B.phaseNum[n] > 2) {
            y = 1.0 * y;
            const double* rw = &sDF[(lane - 1) * S];
            for (int kq = 0; kq < ncolB; ++kq) y += -1.0 * rw[1 + kq] * B.resN[(int64_t)kq * ld + n];
        }
        W.perfRhs[(int64_t)p * ncol + lane] = y;
    }
    __syncthreads();
    // well -> cell block: sum_i weight * row i+1 of M (rate control), zero (BHP)
    const double weight = inj ? 1.0 : W.prodWeight[w];
    double* dst = F.val + (int64_t)W.slotWB[p] * bsize;
    for (int t = 0; t < perLane; ++t) {
        const int e = lane + t * kWellBlock;
        if (e < bsize) {
            double v = 0;
            if (rate && e < ncol) {
                for (int i = 0; i < nc; ++i) v += weight * sM[(i + 1) * ncol + e];
            }
            dst[e] = v;
        }
    }
}

// Blocks [0, npcell): a perforated cell adds its perforations' blocks to its
// diagonal.  Blocks [npcell, npcell + nw): a well row's diagonal.
template <int NC>
__global__ void __launch_bounds__(kWellBlock) wellDiagKernel(FimDevice F, WellDevice W)
{
    constexpr int nc = NC + 1;
    constexpr int ncol = nc + 1;
    constexpr int bsize = ncol * ncol;
    const int lane = threadIdx.x;
    if ((int)blockIdx.x < W.npcell) {
        const int64_t n = W.pcell[blockIdx.x];
        double* d = F.val + (int64_t)F.IA[n] * bsize;
        const int q0 = W.cellPerfPtr[n], q1 = W.cellPerfPtr[n + 1];
        for (int e = lane; e < bsize; e += kWellBlock) {
            double v = d[e];
            for (int q = q0; q < q1; ++q) v += W.perfBlock[(int64_t)W.cellPerfIdx[q] * bsize + e];
            d[e] = v;
        }
        // IsoT_FIMn: the perforations' right-hand-side terms, in well order
        if (W.fimn && lane < ncol) {
            double v = F.rhsN[n * ncol + lane];
            for (int q = q0; q < q1; ++q) v += W.perfRhs[(int64_t)W.cellPerfIdx[q] * ncol + lane];
            F.rhsN[n * ncol + lane] = v;
        }
        return;
    }
    const int w = blockIdx.x - W.npcell;
    if (w >= W.nw || !W.open[w]) return;
    const bool inj = W.type[w] == 0;
    const bool rate = W.mode[w] != 0;
    const double weight = inj ? 1.0 : W.prodWeight[w];
    double* d = F.val + (int64_t)F.IA[F.nb + W.row[w]] * bsize;
    for (int e = lane; e < bsize; e += kWellBlock) {
        const int r = e / ncol, cc = e % ncol;
        double v = 0;
        for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; ++p) {
            if (rate) {
                if (e == 0) {
                    double b0 = 0;
                    for (int i = 0; i < nc; ++i)
                        b0 += inj ? W.perfDQW[(int64_t)p * nc + i] : W.perfDQW[(int64_t)p * nc + i] * weight;
                    v += b0;
                } else if (r == cc) {
                    v += 1;
                } else {
                    v += 0;
                }
            } else {
                v += (r == cc) ? 1.0 : 0.0;
            }
        }
        d[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Per-step well preparation and checks (lane 0 of one block per well; the
// wells' work is a few serial flashes each).

constexpr double kGravity = 0.00694444;   // include/OCPConst.hpp:56
constexpr double kWellTiny = 1E-8;        // include/OCPConst.hpp:38
constexpr double kSegLen = 10;            // maxlen of Well::CalInjdG / CalProddG01

// Well::CalTrans (src/Well.cpp:164-202): perforation phase transmissibilities.
__device__ inline void wellCalTrans(const DevBulk& B, const WellDevice& W, int w)
{
    const int64_t ld = B.ld;
    for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; ++p) {
        const int64_t k = W.perfCell[p];
        const double temp = kCONV1 * W.perfWI[p] * W.perfMult[p];
        for (int j = 0; j < 3; ++j) {
            double tj = 0;
            if (B.phaseExist[j * ld + k]) tj = temp * B.kr[j * ld + k] / B.mu[j * ld + k];
            W.perfTransj[(int64_t)p * 3 + j] = tj;
        }
    }
}

// WellPerf::transINJ as Well::CalTrans leaves it for an injector.
__device__ inline double wellTransInj(const DevBulk& B, const WellDevice& W, int w, int p)
{
    if (W.unweight[w]) return W.perfWI[p];
    const int64_t k = W.perfCell[p];
    double t = 0;
    for (int j = 0; j < 3; ++j)
        if (B.phaseExist[j * B.ld + k]) t += W.perfTransj[(int64_t)p * 3 + j];
    return t;
}

// Segments of the interval between two depths (Well::CalInjdG's seg_num,
// seg_len): 0 segments for a zero-length interval.
__device__ inline int wellSegments(double d, double* segLen)
{
    const int n = (int)ceil(fabs(d / kSegLen));
    *segLen = n > 0 ? d / n : 0.0;
    return n;
}

// Well::CaldG (src/Well.cpp:357-365): Well::CalInjdG (:367-435) for an
// injector, Well::CalProddG01 (:438-555) for a producer.  The pressure at a
// perforation starts from bhp + the previous dG there; W.perfTmp holds the
// per-interval increments.
template <int NC>
__device__ void wellCaldG(const DevEos& E, double* sq, const DevBulk& B, const WellDevice& W, int w)
{
    constexpr int nc = NC + 1;
    const int64_t ld = B.ld;
    const int p0 = W.perfStart[w], np = W.perfStart[w + 1] - p0;
    if (np <= 0) return;
    const double depth = W.wellDepth[w];
    const double bhp = W.bhp[w];
    const bool inj = W.type[w] == 0;
    auto pdepth = [&](int q) { return B.depth[W.perfCell[p0 + q]]; };
    double* dGp = W.perfTmp + p0;
    double* dG = W.perfDG + p0;
    for (int q = 0; q < np; ++q) dGp[q] = 0;
    const bool higher = depth <= pdepth(0);
    const bool lower = !higher && depth >= pdepth(np - 1);
    if (!higher && !lower) return;
    double tmpNi[nc];
    for (int i = 0; i < nc; ++i) tmpNi[i] = 0;
    const double injT = W.injTemp[w];
    const double* zInj = W.injZi + (int64_t)w * nc;
    const int injPh = W.injPhase[w];
    for (int t = 0; t < np; ++t) {
        // "higher": perforations from the deepest up; "lower": from the top down
        const int q = higher ? np - 1 - t : t;
        double dz;
        if (higher)
            dz = q == 0 ? pdepth(0) - depth : pdepth(q) - pdepth(q - 1);
        else
            dz = q == np - 1 ? depth - pdepth(np - 1) : pdepth(q + 1) - pdepth(q);
        double segLen;
        const int seg = wellSegments(dz, &segLen);
        if (seg == 0) continue;
        const int p = p0 + q;
        const int64_t n = W.perfCell[p];
        W.perfP[p] = bhp + dG[q];
        const double Pperf = W.perfP[p];
        double Ptmp = Pperf;
        if (inj) {
            for (int i = 0; i < seg; ++i) {
                double xi, rho;
                phaseDensity<NC>(E, sq, Ptmp, injT, zInj, injPh, &xi, &rho);
                if (higher)
                    Ptmp -= rho * kGravity * segLen;
                else
                    Ptmp += rho * kGravity * segLen;
            }
        } else {
            for (int j = 0; j < 3; ++j) {
                if (!B.phaseExist[j * ld + n]) continue;
                for (int k = 0; k < nc; ++k)
                    tmpNi[k] += (B.P[n] - W.perfP[p]) * W.perfTransj[(int64_t)p * 3 + j] * B.xi[j * ld + n] *
                                B.xij[(j * nc + k) * ld + n];
            }
            double tmpSum = 0;
            for (int k = 0; k < nc; ++k) tmpSum += fabs(tmpNi[k]);
            if (tmpSum < kWellTiny)
                for (int k = 0; k < nc; ++k) tmpNi[k] = B.Ni[k * ld + n];
            for (int i = 0; i < seg; ++i) {
                const FlashPhases f = flashPhases<NC>(E, sq, Ptmp, B.T[n], tmpNi);
                double qtacc = 0, rhoacc = 0;
                for (int j = 0; j < 3; ++j) {
                    if (!f.ex[j]) continue;
                    qtacc += f.vj[j];
                    rhoacc += f.vj[j] * f.rho[j];
                }
                if (higher)
                    Ptmp -= rhoacc / qtacc * kGravity * segLen;
                else
                    Ptmp += rhoacc / qtacc * kGravity * segLen;
            }
        }
        dGp[q] = higher ? Pperf - Ptmp : Ptmp - Pperf;
    }
    if (higher) {
        dG[0] = dGp[0];
        for (int q = 1; q < np; ++q) dG[q] = dG[q - 1] + dGp[q];
    } else {
        dG[np - 1] = dGp[np - 1];
        for (int q = np - 2; q >= 0; --q) dG[q] = dG[q + 1] + dGp[q];
    }
}

// Well::CheckOptMode (src/Well.cpp:893-944) with Well::CalInjRateMaxBHP
// (:263-284) and Well::CalProdRateMinBHP (:289-320).
template <int NC>
__device__ void wellCheckOptMode(const DevEos& E, double* sq, const DevBulk& B, const WellDevice& W, int w)
{
    constexpr int nc = NC + 1;
    const int64_t ld = B.ld;
    const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
    const bool inj = W.type[w] == 0;
    const int initMode = W.initMode[w];
    double q = 0;
    if (inj) {
        const double Pwell = W.ctrlBHP[w];
        for (int p = p0; p < p1; ++p) {
            const double Pperf = Pwell + W.perfDG[p];
            const int64_t k = W.perfCell[p];
            double xi, rho;
            phaseDensity<NC>(E, sq, Pperf, W.injTemp[w], W.injZi + (int64_t)w * nc, W.injPhase[w], &xi, &rho);
            const double dP = Pperf - B.P[k];
            q += wellTransInj(B, W, w, p) * xi * dP;
        }
        if (initMode == 0) {
            if (q > W.ctrlRate[w]) {
                W.mode[w] = 1;
            } else {
                W.mode[w] = 0;
                W.bhp[w] = W.ctrlBHP[w];
            }
        } else if (q > W.ctrlRate[w]) {
            W.mode[w] = initMode;
        } else {
            W.mode[w] = 0;
            W.bhp[w] = W.ctrlBHP[w];
        }
        return;
    }
    if (initMode == 0) {
        W.mode[w] = 0;
        W.bhp[w] = W.ctrlBHP[w];
        return;
    }
    const double Pwell = W.ctrlBHP[w];
    double qi[nc];
    for (int i = 0; i < nc; ++i) qi[i] = 0;
    for (int p = p0; p < p1; ++p) {
        const double Pperf = Pwell + W.perfDG[p];
        const int64_t k = W.perfCell[p];
        for (int j = 0; j < 3; ++j) {
            if (!B.phaseExist[j * ld + k]) continue;
            const double dP = B.Pj[j * ld + k] - Pperf;
            const double temp = W.perfTransj[(int64_t)p * 3 + j] * B.xi[j * ld + k] * dP;
            for (int i = 0; i < nc; ++i) qi[i] += B.xij[(j * nc + i) * ld + k] * temp;
        }
    }
    double rate[3];
    surfaceProdRate<NC>(E, sq, W.Psurf, W.Tsurf, qi, rate);
    for (int j = 0; j < 3; ++j) q += rate[j] * W.prodPW[(int64_t)w * 3 + j];
    if (q > W.ctrlRate[w]) {
        W.mode[w] = initMode;
    } else {
        W.mode[w] = 0;
        W.bhp[w] = W.ctrlBHP[w];
    }
}

// what = 0: AllWells::PrepareWell (src/AllWells.cpp:201-214): per open well
//           CalTrans, CaldG, CheckOptMode, then CalFlux's perforation
//           pressures and, with ReCalXi, the injected fluid's molar density
//           at each perforation (src/Well.cpp:204-228).
// what = 1: the well part of IsoT_FIM::ResetToLastTimeStep (src/
//           OCPFluidMethod.cpp:2897-2902): AllWells::ResetBHP (bhp = lbhp,
//           Well::CorrectBHP, src/Well.cpp:882-890), CalTrans, CaldG and
//           CalFlux's perforation pressures.
template <int NC>
__global__ void __launch_bounds__(kWellBlock) wellPrepareKernel(const DevEos* __restrict__ eos, DevBulk B,
                                                               WellDevice W, int what)
{
    constexpr int nc = NC + 1;
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    const int w = blockIdx.x;
    if (threadIdx.x != 0 || w >= W.nw || !W.open[w]) return;
    const DevEos& E = *eos;
    const bool inj = W.type[w] == 0;
    if (what == 1) {
        W.bhp[w] = W.lbhp[w];
        if (W.mode[w] == 0) W.bhp[w] = W.ctrlBHP[w];
    }
    wellCalTrans(B, W, w);
    if (W.hasSetup) wellCaldG<NC>(E, &sSq[0], B, W, w);
    if (what == 0) wellCheckOptMode<NC>(E, &sSq[0], B, W, w);
    const double bhp = W.bhp[w];
    for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; ++p) {
        W.perfP[p] = bhp + W.perfDG[p];
        if (what == 0 && inj) {
            double xi, rho;
            phaseDensity<NC>(E, &sSq[0], W.perfP[p], W.injTemp[w], W.injZi + (int64_t)w * nc, W.injPhase[w], &xi,
                             &rho);
            W.perfXi[p] = xi;
        }
    }
}

// AllWells::CheckP (src/AllWells.cpp:335-369): Well::CheckP and
// Well::CheckCrossFlow (src/Well.cpp:946-1066) of the open wells in order,
// one thread, stopping at the first well with a negative pressure.  Writes
// W.status[0]: 0 success, 1 negative pressure, 2 a control switched to BHP or
// a perforation closed by cross flow.
template <int NC>
__global__ void wellCheckKernel(DevBulk B, WellDevice W)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bool flagSwitch = false, flagCross = false;
    int result = 0;
    for (int w = 0; w < W.nw && result == 0; ++w) {
        if (!W.open[w]) continue;
        const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
        const bool inj = W.type[w] == 0;
        bool neg = W.bhp[w] < 0;
        for (int p = p0; p < p1 && !neg; ++p)
            if (W.perfState[p] == 1 && W.perfP[p] < 0) neg = true;
        if (neg) {
            result = 1;
            break;
        }
        if (W.mode[w] != 0 && (inj ? W.bhp[w] > W.ctrlBHP[w] : W.bhp[w] < W.ctrlBHP[w])) {
            W.mode[w] = 0;
            W.bhp[w] = W.ctrlBHP[w];
            flagSwitch = true;
            continue;
        }
        // Well::CheckCrossFlow
        bool flagC = true;
        for (int p = p0; p < p1; ++p) {
            const double Pk = B.P[W.perfCell[p]];
            const bool cross = inj ? Pk > W.perfP[p] : Pk < W.perfP[p];
            const bool back = inj ? Pk < W.perfP[p] : Pk > W.perfP[p];
            if (W.perfState[p] == 1 && cross) {
                W.perfState[p] = 0;
                W.perfMult[p] = 0;
                flagC = false;
                break;
            } else if (W.perfState[p] == 0 && back) {
                W.perfState[p] = 1;
                W.perfMult[p] = 1;
            }
        }
        bool anyOpen = false;
        for (int p = p0; p < p1; ++p) anyOpen = anyOpen || W.perfState[p] == 1;
        if (!anyOpen && p1 > p0) {
            W.perfState[p1 - 1] = 1;
            W.perfMult[p1 - 1] = 1;
        }
        if (!flagC) {
            wellCalTrans(B, W, w);
            flagCross = true;
        }
    }
    if (result == 0 && (flagSwitch || flagCross)) result = 2;
    W.status[0] = result;
}

// MixtureComp::XiPhase and RhoPhase of one feed (ocpg_phase_density).
template <int NC>
__global__ void phaseDensityKernel(const DevEos* __restrict__ eos, double P, double TF, const double* zi, int phase,
                                   double* out)
{
    __shared__ double sSq[FlashCell<NC>::NT * kFlashBlock];
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    phaseDensity<NC>(*eos, &sSq[0], P, TF, zi, phase, &out[0], &out[1]);
}

template <int NC>
hipError_t launchWellPrepareT(const DevEos* eos, const DevBulk& B, WellDevice& W, int what, hipStream_t s)
{
    if (W.nw == 0) return hipSuccess;
    hipLaunchKernelGGL(wellPrepareKernel<NC>, dim3(W.nw), dim3(kWellBlock), 0, s, eos, B, W, what);
    return hipGetLastError();
}

template <int NC>
hipError_t launchWellCheckT(const DevBulk& B, WellDevice& W, hipStream_t s)
{
    if (W.nw == 0) return hipSuccess;
    hipLaunchKernelGGL(wellCheckKernel<NC>, dim3(1), dim3(kWellBlock), 0, s, B, W);
    return hipGetLastError();
}

template <int NC>
hipError_t launchPhaseDensityT(const DevEos* eos, double P, double TF, const double* zi, int phase, double* out,
                               hipStream_t s)
{
    hipLaunchKernelGGL(phaseDensityKernel<NC>, dim3(1), dim3(kWellBlock), 0, s, eos, P, TF, zi, phase, out);
    return hipGetLastError();
}

template <int NC>
hipError_t launchWellFluxT(const DevEos* eos, const DevBulk& B, FimDevice& F, WellDevice& W, hipStream_t s)
{
    if (W.nw == 0) return hipSuccess;
    hipLaunchKernelGGL(wellFluxKernel<NC>, dim3(W.nw), dim3(kWellBlock), 0, s, eos, B, F, W);
    return hipGetLastError();
}

template <int NC>
hipError_t launchWellAssembleT(const DevBulk& B, FimDevice& F, WellDevice& W, double dt, hipStream_t s)
{
    if (W.nw == 0 || W.nperf == 0) return hipSuccess;
    hipLaunchKernelGGL(wellAssembleKernel<NC>, dim3(W.nperf), dim3(kWellBlock), 0, s, B, F, W, dt);
    hipLaunchKernelGGL(wellDiagKernel<NC>, dim3(W.npcell + W.nw), dim3(kWellBlock), 0, s, F, W);
    return hipGetLastError();
}

}  // namespace ocpg
