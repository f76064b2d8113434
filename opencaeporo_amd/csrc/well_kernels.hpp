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

// MixtureComp::CalProdWeight: MixtureComp::Flash (src/MixtureComp.cpp:
// 167-193) of the stream Ni at surface conditions, then the stream's volume
// per mole in the rate units of prodPhase.  A hydrocarbon phase the flash
// does not produce contributes no volume.
template <int NC>
__device__ double surfaceProdWeight(const DevEos& E, double* sqLds, double Psurf, double Tsurf, const double* Ni,
                                    const double* prodPhase)
{
    FlashCell<NC> F(E, sqLds);
    F.P = Psurf;
    F.T = Tsurf + kCONV5;
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
    double vj[3] = {0, 0, 0};
    double MW[2], vC[2];
    for (int j = 0; j < NP; ++j) {
        F.nu[j] = Nh * F.nu[j];
        MW[j] = 0;
        for (int i = 0; i < NC; ++i) MW[j] += F.x[j][i] * E.MWC[i];
    }
    for (int j = 0; j < NP; ++j) {
        double tmp = F.Zj[j] * kGasConst * T / P;
        for (int i = 0; i < NC; ++i) tmp -= F.x[j][i] * E.Vshift[i];
        vC[j] = tmp * F.nu[j];
    }
    // IdentifyPhase (src/MixtureComp.cpp:2352-2384) + CopyPhase
    if (NP == 1) {
        double A = 0, B = 0;
        for (int i = 0; i < NC; ++i) {
            A += F.x[0][i] * E.Vc[i] * E.Tc[i];
            B += F.x[0][i] * E.Vc[i];
        }
        const double Tc = A / B;
        vj[T > Tc ? 1 : 0] = vC[0];
    } else if (MW[0] > MW[1]) {
        vj[0] = vC[0];
        vj[1] = vC[1];
    } else {
        vj[1] = vC[0];
        vj[0] = vC[1];
    }
    double pw[5];
    pvtwValues(E, P, pw);
    const double bw = pw[1] * (1 - pw[2] * (P - pw[0]));
    const double xiW = 1 / (kCONV1 * bw);
    vj[2] = kCONV1 * Ni[NC] * bw;
    const double qt = Nh + Ni[NC];
    double factor[3];
    factor[0] = vj[0] / qt / kCONV1;
    factor[1] = vj[1] / qt / 1000;
    factor[2] = xiW * vj[2] / qt;
    double tmp = 0;
    for (int i = 0; i < 3; ++i) tmp += factor[i] * prodPhase[i];
    return tmp;
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
// cell's dSec_dPri.
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
                    const double tmp = dP * tj[j] * xij *
                                       (B.xix[(j * nc + kX) * ld + n] - xi / mu * B.mux[(j * nc + kX) * ld + n]);
                    v += tmp;
                    if (kX == i) v += tj[j] * xi * dP;
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
