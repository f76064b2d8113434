// TEST INFRASTRUCTURE ONLY: CPU restatement of the well terms of IsoT_FIM
// (oracle + CPU baseline).  Not part of the product.
//
//   CalWellFlux   : Well::CalTrans (src/Well.cpp:164-202), Well::CalFlux with
//                   ReCalXi = false (:204-258), Well::CalProdWeight (:829-862)
//   AssembleWells : IsoT_FIM::AssembleMatWellsNew (src/OCPFluidMethod.cpp:
//                   2417-2436) -> AssembleMatInjWellsNew (:2438-2580),
//                   AssembleMatProdWellsNew (:2582-2733); IsoT_FIMn's
//                   (:3503-3845) with fimn
//   PrepareWells  : AllWells::PrepareWell (src/AllWells.cpp:201-214):
//                   Well::CalTrans, CaldG (CalInjdG :367-435, CalProddG01
//                   :438-555), CheckOptMode (:893-944), CalFlux's ReCalXi
//   CheckWells    : AllWells::CheckP (src/AllWells.cpp:335-369), Well::CheckP
//                   and Well::CheckCrossFlow (src/Well.cpp:946-1066)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle_fim.hpp"

namespace oracle {

int OpenWells(const ocpg_wells* w)
{
    if (!w) return 0;
    int k = 0;
    for (int i = 0; i < w->nw; i++) k += w->open[i] ? 1 : 0;
    return k;
}

void CalWellFlux(const ocpg_eos& eos, const ocpg_wells& W, const ocpg_bulk& bk, WellFlux& out)
{
    const int np = NPH;
    const int nc = eos.nc + 1;
    const int nperf = W.nw > 0 ? W.perfStart[W.nw] : 0;
    out.perfTransj.assign((size_t)nperf * np, 0.0);
    out.perfQi.assign((size_t)nperf * nc, 0.0);
    out.wellQi.assign((size_t)W.nw * nc, 0.0);
    out.prodWeight.assign(W.nw, 0.0);
    for (int w = 0; w < W.nw; w++) {
        if (!W.open[w]) continue;
        const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
        double* qiW = &out.wellQi[(size_t)w * nc];
        if (W.type[w] == OCPG_WELL_INJ) {
            std::vector<double> transINJ(p1 - p0, 0.0);
            for (int p = p0; p < p1; p++) {
                const int64_t k = W.perfCell[p];
                const double temp = CONV1 * W.perfWI[p] * W.perfMult[p];
                for (int j = 0; j < np; j++) {
                    double& tj = out.perfTransj[(size_t)p * np + j];
                    tj = 0;
                    const int64_t id = k * np + j;
                    if (bk.phaseExist[id]) {
                        tj = temp * bk.kr[id] / bk.mu[id];
                        transINJ[p - p0] += tj;
                    }
                }
                if (W.unweight && W.unweight[w]) transINJ[p - p0] = W.perfWI[p];
            }
            for (int p = p0; p < p1; p++) {
                const double Pperf = W.bhp[w] + W.perfDG[p];
                const int64_t k = W.perfCell[p];
                const double dP = bk.P[k] - Pperf;
                const double qt = transINJ[p - p0] * dP;
                for (int i = 0; i < nc; i++) {
                    double& q = out.perfQi[(size_t)p * nc + i];
                    q = qt * W.perfXi[p] * W.injZi[(size_t)w * nc + i];
                    qiW[i] += q;
                }
            }
        } else {
            for (int p = p0; p < p1; p++) {
                const int64_t k = W.perfCell[p];
                const double temp = CONV1 * W.perfWI[p] * W.perfMult[p];
                for (int j = 0; j < np; j++) {
                    double& tj = out.perfTransj[(size_t)p * np + j];
                    tj = 0;
                    const int64_t id = k * np + j;
                    if (bk.phaseExist[id]) tj = temp * bk.kr[id] / bk.mu[id];
                }
            }
            for (int p = p0; p < p1; p++) {
                const double Pperf = W.bhp[w] + W.perfDG[p];
                const int64_t k = W.perfCell[p];
                double* q = &out.perfQi[(size_t)p * nc];
                for (int j = 0; j < np; j++) {
                    const int64_t id = k * np + j;
                    if (!bk.phaseExist[id]) continue;
                    const double dP = bk.Pj[id] - Pperf;
                    const double qj = out.perfTransj[(size_t)p * np + j] * dP;
                    const double xi = bk.xi[id];
                    for (int i = 0; i < nc; i++) q[i] += qj * xi * bk.xij[id * nc + i];
                }
                for (int i = 0; i < nc; i++) qiW[i] += q[i];
            }
            if (W.mode[w] != OCPG_WELL_BHP) {
                // Well::CalProdWeight
                double qt = 0;
                bool flag = true;
                for (int i = 0; i < nc; i++) {
                    qt += qiW[i];
                    if (qiW[i] < 0) flag = false;
                }
                Mixture mix(eos);
                if (qt > TINY && flag) {
                    out.prodWeight[w] = mix.ProdWeight(W.Psurf, W.Tsurf, qiW, &W.prodPhaseWeight[(size_t)w * 3]);
                } else {
                    std::vector<double> tmpNi(nc, 0.0);
                    for (int p = p0; p < p1; p++) {
                        const int64_t n = W.perfCell[p];
                        for (int j = 0; j < np; j++) {
                            const int64_t id = n * np + j;
                            if (!bk.phaseExist[id]) continue;
                            for (int k = 0; k < nc; k++)
                                tmpNi[k] += out.perfTransj[(size_t)p * np + j] * bk.xi[id] * bk.xij[id * nc + k];
                        }
                    }
                    out.prodWeight[w] =
                        mix.ProdWeight(W.Psurf, W.Tsurf, tmpNi.data(), &W.prodPhaseWeight[(size_t)w * 3]);
                }
            }
        }
    }
}

// Well::CalTrans, src/Well.cpp:164-202 (transj per perforation and phase).
static void WellTrans(const ocpg_wells& W, const ocpg_bulk& bk, int w, double* transj)
{
    for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; p++) {
        const int64_t k = W.perfCell[p];
        const double temp = CONV1 * W.perfWI[p] * W.perfMult[p];
        for (int j = 0; j < NPH; j++) {
            transj[(size_t)p * NPH + j] = 0;
            const int64_t id = k * NPH + j;
            if (bk.phaseExist[id]) transj[(size_t)p * NPH + j] = temp * bk.kr[id] / bk.mu[id];
        }
    }
}

static double TransInj(const ocpg_wells& W, const ocpg_bulk& bk, const double* transj, int w, int p)
{
    if (W.unweight && W.unweight[w]) return W.perfWI[p];
    double t = 0;
    const int64_t k = W.perfCell[p];
    for (int j = 0; j < NPH; j++)
        if (bk.phaseExist[k * NPH + j]) t += transj[(size_t)p * NPH + j];
    return t;
}

// Well::CaldG -> CalInjdG / CalProddG01 (src/Well.cpp:357-555).
static void WellCaldG(Mixture& mix, const ocpg_wells& W, const ocpg_well_setup& S, const ocpg_bulk& bk, int w,
                      WellStep& st)
{
    const int nc = mix.numCom;
    const int p0 = W.perfStart[w];
    const int numPerf = W.perfStart[w + 1] - p0;
    if (numPerf <= 0) return;
    double* dG = const_cast<double*>(W.perfDG) + p0;
    const double bhp = W.bhp[w];
    const double depth = S.depth[w];
    const double maxlen = 10;
    std::vector<double> dGperf(numPerf, 0.0);
    std::vector<double> tmpNi(nc, 0.0);
    auto pdepth = [&](int q) { return bk.depth[W.perfCell[p0 + q]]; };
    const bool inj = W.type[w] == OCPG_WELL_INJ;
    const double* zInj = W.injZi + (size_t)w * nc;
    // one segment's pressure change: the injected fluid's density, or the
    // volume-weighted density of the flashed perforation stream
    auto rhoAt = [&](double Ptmp, int64_t n) {
        if (inj) {
            double xi, rho;
            mix.PhaseDensity(Ptmp, S.injTemp[w], zInj, S.injPhase[w] == 2 ? 2 : 1, &xi, &rho);
            return rho;
        }
        bool ex[NPH];
        double vj[NPH], xi[NPH], rho[NPH];
        mix.FlashPhases(Ptmp, bk.T[n], tmpNi.data(), ex, vj, xi, rho);
        double qtacc = 0, rhoacc = 0;
        for (int j = 0; j < NPH; j++) {
            if (!ex[j]) continue;
            qtacc += vj[j];
            rhoacc += vj[j] * rho[j];
        }
        return rhoacc / qtacc;
    };
    auto stream = [&](int p, int64_t n) {
        for (int j = 0; j < NPH; j++) {
            const int64_t id = n * NPH + j;
            if (!bk.phaseExist[id]) continue;
            for (int k = 0; k < nc; k++)
                tmpNi[k] += (bk.P[n] - st.perfP[p]) * st.perfTransj[(size_t)p * NPH + j] * bk.xi[id] *
                            bk.xij[id * nc + k];
        }
        double tmpSum = 0;
        for (int k = 0; k < nc; k++) tmpSum += std::fabs(tmpNi[k]);
        if (tmpSum < TINY)
            for (int i = 0; i < nc; i++) tmpNi[i] = bk.Ni[n * nc + i];
    };
    if (depth <= pdepth(0)) {
        // well above its perforations
        for (int q = numPerf - 1; q >= 0; q--) {
            const double dz = q == 0 ? pdepth(0) - depth : pdepth(q) - pdepth(q - 1);
            const int seg_num = (int)std::ceil(std::fabs(dz / maxlen));
            if (seg_num == 0) continue;
            const double seg_len = dz / seg_num;
            const int p = p0 + q;
            const int64_t n = W.perfCell[p];
            st.perfP[p] = bhp + dG[q];
            const double Pperf = st.perfP[p];
            double Ptmp = Pperf;
            if (!inj) stream(p, n);
            for (int i = 0; i < seg_num; i++) Ptmp -= rhoAt(Ptmp, n) * GRAVITY_FACTOR * seg_len;
            dGperf[q] = Pperf - Ptmp;
        }
        dG[0] = dGperf[0];
        for (int q = 1; q < numPerf; q++) dG[q] = dG[q - 1] + dGperf[q];
    } else if (depth >= pdepth(numPerf - 1)) {
        // well below its perforations
        for (int q = 0; q < numPerf; q++) {
            const double dz = q == numPerf - 1 ? depth - pdepth(numPerf - 1) : pdepth(q + 1) - pdepth(q);
            const int seg_num = (int)std::ceil(std::fabs(dz / maxlen));
            if (seg_num == 0) continue;
            const double seg_len = dz / seg_num;
            const int p = p0 + q;
            const int64_t n = W.perfCell[p];
            st.perfP[p] = bhp + dG[q];
            const double Pperf = st.perfP[p];
            double Ptmp = Pperf;
            if (!inj) stream(p, n);
            for (int i = 0; i < seg_num; i++) Ptmp += rhoAt(Ptmp, n) * GRAVITY_FACTOR * seg_len;
            dGperf[q] = Ptmp - Pperf;
        }
        dG[numPerf - 1] = dGperf[numPerf - 1];
        for (int q = numPerf - 2; q >= 0; q--) dG[q] = dG[q + 1] + dGperf[q];
    }
}

// Well::CheckOptMode (src/Well.cpp:893-944).
static void WellCheckOptMode(Mixture& mix, ocpg_wells& W, const ocpg_well_setup& S, const ocpg_bulk& bk, int w,
                             WellStep& st)
{
    const int nc = mix.numCom;
    int32_t* mode = const_cast<int32_t*>(W.mode);
    double* bhp = const_cast<double*>(W.bhp);
    const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
    const bool inj = W.type[w] == OCPG_WELL_INJ;
    const int initMode = S.initMode[w];
    if (inj) {
        // Well::CalInjRateMaxBHP (:263-284)
        double q = 0;
        for (int p = p0; p < p1; p++) {
            const double Pperf = W.ctrlBHP[w] + W.perfDG[p];
            const int64_t k = W.perfCell[p];
            double xi, rho;
            mix.PhaseDensity(Pperf, S.injTemp[w], W.injZi + (size_t)w * nc, S.injPhase[w] == 2 ? 2 : 1, &xi, &rho);
            const double dP = Pperf - bk.P[k];
            q += TransInj(W, bk, st.perfTransj, w, p) * xi * dP;
        }
        if (q > W.ctrlRate[w]) {
            mode[w] = initMode == OCPG_WELL_BHP ? OCPG_WELL_RATE : initMode;
        } else {
            mode[w] = OCPG_WELL_BHP;
            bhp[w] = W.ctrlBHP[w];
        }
        return;
    }
    if (initMode == OCPG_WELL_BHP) {
        mode[w] = OCPG_WELL_BHP;
        bhp[w] = W.ctrlBHP[w];
        return;
    }
    // Well::CalProdRateMinBHP (:289-320)
    std::vector<double> qi(nc, 0.0);
    for (int p = p0; p < p1; p++) {
        const double Pperf = W.ctrlBHP[w] + W.perfDG[p];
        const int64_t k = W.perfCell[p];
        for (int j = 0; j < NPH; j++) {
            const int64_t id = k * NPH + j;
            if (!bk.phaseExist[id]) continue;
            const double dP = bk.Pj[id] - Pperf;
            const double temp = st.perfTransj[(size_t)p * NPH + j] * bk.xi[id] * dP;
            for (int i = 0; i < nc; i++) qi[i] += bk.xij[id * nc + i] * temp;
        }
    }
    double rate[NPH];
    mix.ProdRate(W.Psurf, W.Tsurf, qi.data(), rate);
    double q = 0;
    for (int j = 0; j < NPH; j++) q += rate[j] * W.prodPhaseWeight[(size_t)w * 3 + j];
    if (q > W.ctrlRate[w]) {
        mode[w] = initMode;
    } else {
        mode[w] = OCPG_WELL_BHP;
        bhp[w] = W.ctrlBHP[w];
    }
}

void PrepareWells(const ocpg_eos& eos, ocpg_wells& W, const ocpg_well_setup* S, const ocpg_bulk& bk,
                  WellStep& st, int what)
{
    Mixture mix(eos);
    const int nc = eos.nc + 1;
    double* bhp = const_cast<double*>(W.bhp);
    double* perfXi = const_cast<double*>(W.perfXi);
    for (int w = 0; w < W.nw; w++) {
        if (!W.open[w]) continue;
        if (what == 1) {
            // AllWells::ResetBHP: bhp = lbhp, Well::CorrectBHP
            bhp[w] = st.lbhp[w];
            if (W.mode[w] == OCPG_WELL_BHP) bhp[w] = W.ctrlBHP[w];
        }
        WellTrans(W, bk, w, st.perfTransj);
        if (S) WellCaldG(mix, W, *S, bk, w, st);
        if (what == 0 && S) WellCheckOptMode(mix, W, *S, bk, w, st);
        for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; p++) {
            st.perfP[p] = bhp[w] + W.perfDG[p];
            if (what == 0 && S && W.type[w] == OCPG_WELL_INJ) {
                double xi, rho;
                mix.PhaseDensity(st.perfP[p], S->injTemp[w], W.injZi + (size_t)w * nc, S->injPhase[w] == 2 ? 2 : 1,
                                 &xi, &rho);
                perfXi[p] = xi;
            }
        }
    }
}

int CheckWells(ocpg_wells& W, const ocpg_bulk& bk, WellStep& st)
{
    int32_t* mode = const_cast<int32_t*>(W.mode);
    double* bhp = const_cast<double*>(W.bhp);
    double* mult = const_cast<double*>(W.perfMult);
    bool flagSwitch = false, flagCross = false;
    for (int w = 0; w < W.nw; w++) {
        if (!W.open[w]) continue;
        const int p0 = W.perfStart[w], p1 = W.perfStart[w + 1];
        const bool inj = W.type[w] == OCPG_WELL_INJ;
        // Well::CheckP
        if (bhp[w] < 0) return 1;
        for (int p = p0; p < p1; p++)
            if (st.perfState[p] == 1 && st.perfP[p] < 0) return 1;
        if (inj) {
            if (mode[w] != OCPG_WELL_BHP && bhp[w] > W.ctrlBHP[w]) {
                mode[w] = OCPG_WELL_BHP;
                bhp[w] = W.ctrlBHP[w];
                flagSwitch = true;
                continue;
            }
        } else if (mode[w] != OCPG_WELL_BHP && bhp[w] < W.ctrlBHP[w]) {
            mode[w] = OCPG_WELL_BHP;
            bhp[w] = W.ctrlBHP[w];
            flagSwitch = true;
            continue;
        }
        // Well::CheckCrossFlow
        bool flagC = true;
        for (int p = p0; p < p1; p++) {
            const double Pk = bk.P[W.perfCell[p]];
            if (!inj) {
                if (st.perfState[p] == 1 && Pk < st.perfP[p]) {
                    st.perfState[p] = 0;
                    mult[p] = 0;
                    flagC = false;
                    break;
                } else if (st.perfState[p] == 0 && Pk > st.perfP[p]) {
                    st.perfState[p] = 1;
                    mult[p] = 1;
                }
            } else {
                if (st.perfState[p] == 1 && Pk > st.perfP[p]) {
                    st.perfState[p] = 0;
                    mult[p] = 0;
                    flagC = false;
                    break;
                } else if (st.perfState[p] == 0 && Pk < st.perfP[p]) {
                    st.perfState[p] = 1;
                    mult[p] = 1;
                }
            }
        }
        bool flag = false;
        for (int p = p0; p < p1; p++)
            if (st.perfState[p] == 1) flag = true;
        if (!flag && p1 > p0) {
            st.perfState[p1 - 1] = 1;
            mult[p1 - 1] = 1;
        }
        if (!flagC) {
            WellTrans(W, bk, w, st.perfTransj);
            flagCross = true;
        }
    }
    return (flagSwitch || flagCross) ? 2 : 0;
}

// DaABpbC, src/DenseMat.cpp:75-94 (row-major, accumulated over k in order).
static void DaABpbCw(int m, int n, int k, const double* A, const double* B, double* C)
{
    for (int r = 0; r < m; r++)
        for (int p = 0; p < k; p++) {
            const double a = A[r * k + p];
            for (int c = 0; c < n; c++) C[r * n + c] = C[r * n + c] + a * B[p * n + c];
        }
}

void AssembleWells(int nc, const ocpg_conn& conn, const ocpg_wells& W, const WellFlux& wf,
                   const ocpg_bulk& bk, double dt, const int32_t* IA, double* val, bool fimn, double* rhs)
{
    const int np = NPH;
    const int ncol = nc + 1;
    const int ncol2 = np * nc + np;
    const int bsize = ncol * ncol;
    const int bsize2 = ncol * ncol2;
    const int lendSdP = maxLendSdP(nc);
    const int64_t nb = bk.nb;
    // next free slot of each row after the bulk connections
    std::vector<int64_t> fill(nb + OpenWells(&W));
    for (int64_t n = 0; n < nb; n++) fill[n] = IA[n] + 1;
    for (int64_t c = 0; c < conn.nconn; c++) {
        fill[conn.bId[c]]++;
        fill[conn.eId[c]]++;
    }

    std::vector<double> bmat(bsize), bmat2(bsize), dQdXpB(bsize), dQdXpW(bsize), dQdXsB(bsize2);
    bool phaseExistB[NPH], phasedS_B[NPH];
    int pVnumComB[NPH];
    int64_t wId = nb;
    for (int w = 0; w < W.nw; w++) {
        if (!W.open[w]) continue;
        const bool inj = W.type[w] == OCPG_WELL_INJ;
        const double weight = inj ? 1.0 : wf.prodWeight[w];  // ProdWeight(i): one value for all i
        // NewDiag(wId, 0)
        fill[wId] = IA[wId];
        std::memset(&val[(size_t)IA[wId] * bsize], 0, sizeof(double) * bsize);
        fill[wId]++;
        double* wdiag = &val[(size_t)IA[wId] * bsize];
        for (int p = W.perfStart[w]; p < W.perfStart[w + 1]; p++) {
            const int64_t n = W.perfCell[p];
            std::fill(dQdXpB.begin(), dQdXpB.end(), 0.0);
            std::fill(dQdXpW.begin(), dQdXpW.end(), 0.0);
            std::fill(dQdXsB.begin(), dQdXsB.end(), 0.0);
            int jxB = 0, ncolB = 0;
            for (int j = 0; j < np; j++) {
                phaseExistB[j] = bk.phaseExist[n * np + j];
                phasedS_B[j] = bk.pSderExist[n * np + j];
                if (phasedS_B[j]) jxB++;
                pVnumComB[j] = bk.pVnumCom[n * np + j];
                ncolB += pVnumComB[j];
            }
            ncolB += jxB;
            const double* transj = &wf.perfTransj[(size_t)p * np];
            if (inj) {
                const double dP = bk.P[n] - W.bhp[w] - W.perfDG[p];
                for (int j = 0; j < np; j++) {
                    if (!phaseExistB[j]) {
                        jxB += pVnumComB[j];
                        continue;
                    }
                    const int64_t n_np_j = n * np + j;
                    const double mu = bk.mu[n_np_j];
                    const double muP = bk.muP[n_np_j];
                    for (int i = 0; i < nc; i++) {
                        const double zi = W.injZi[(size_t)w * nc + i];
                        const double transIJ = transj[j] * W.perfXi[p] * zi;
                        dQdXpB[(i + 1) * ncol] += transIJ * (1 - dP * muP / mu);
                        dQdXpW[(i + 1) * ncol] += -transIJ;
                        int j1B = 0;
                        for (int j1 = 0; j1 < np; j1++) {
                            if (phasedS_B[j1]) {
                                dQdXsB[(i + 1) * ncolB + j1B] += CONV1 * W.perfWI[p] * W.perfMult[p] * W.perfXi[p] *
                                                                 zi * bk.dKr_dS[n_np_j * np + j1] * dP / mu;
                                j1B++;
                            }
                        }
                        for (int k = 0; k < pVnumComB[j]; k++)
                            dQdXsB[(i + 1) * ncolB + jxB + k] += -transIJ * dP / mu * bk.mux[n_np_j * nc + k];
                    }
                    jxB += pVnumComB[j];
                }
            } else {
                for (int j = 0; j < np; j++) {
                    if (!phaseExistB[j]) {
                        jxB += pVnumComB[j];
                        continue;
                    }
                    const int64_t n_np_j = n * np + j;
                    const double dP = bk.Pj[n_np_j] - W.bhp[w] - W.perfDG[p];
                    const double xi = bk.xi[n_np_j];
                    const double mu = bk.mu[n_np_j];
                    const double muP = bk.muP[n_np_j];
                    const double xiP = bk.xiP[n_np_j];
                    for (int i = 0; i < nc; i++) {
                        const double xij = bk.xij[n_np_j * nc + i];
                        const double transIJ = transj[j] * xi * xij;
                        dQdXpB[(i + 1) * ncol] += transIJ * (1 - dP * muP / mu) + dP * transj[j] * xij * xiP;
                        dQdXpW[(i + 1) * ncol] += -transIJ;
                        int j1B = 0;
                        for (int j1 = 0; j1 < np; j1++) {
                            if (phasedS_B[j1]) {
                                double tmp = CONV1 * W.perfWI[p] * W.perfMult[p] * dP / mu * xi * xij *
                                             bk.dKr_dS[n_np_j * np + j1];
                                tmp += transIJ * bk.dPcj_dS[n_np_j * np + j1];
                                dQdXsB[(i + 1) * ncolB + j1B] += tmp;
                                j1B++;
                            }
                        }
                        for (int k = 0; k < pVnumComB[j]; k++) {
                            double tmp = dP * transj[j] * xij *
                                         (bk.xix[n_np_j * nc + k] - xi / mu * bk.mux[n_np_j * nc + k]);
                            if (fimn) tmp -= transIJ * dP / bk.nj[n_np_j];
                            dQdXsB[(i + 1) * ncolB + jxB + k] += tmp;
                        }
                        if (i < pVnumComB[j])
                            dQdXsB[(i + 1) * ncolB + jxB + i] +=
                                fimn ? transj[j] * xi * dP / bk.nj[n_np_j] : transj[j] * xi * dP;
                    }
                    jxB += pVnumComB[j];
                }
            }
            // IsoT_FIMn: AddRhs(n, -dQdXsB res_n) where the cell has two
            // hydrocarbon phases (src/OCPFluidMethod.cpp:3610-3616, 3780-3786)
            if (fimn && rhs && bk.phaseNum[n] > 2) {
                std::vector<double> tmpb(ncol, 0.0);
                for (int r = 0; r < ncol; r++) {
                    tmpb[r] = 1.0 * tmpb[r];
                    for (int c = 0; c < ncolB; c++)
                        tmpb[r] += -1.0 * dQdXsB[r * ncolB + c] * bk.res_n[n * ncol2 + c];
                }
                for (int r = 0; r < ncol; r++) rhs[n * ncol + r] += tmpb[r];
            }
            // Bulk - Bulk: AddDiag(n, dt * (dQdXpB + dQdXsB * dSec))
            bmat = dQdXpB;
            DaABpbCw(ncol, ncol, ncolB, dQdXsB.data(), &bk.dSec_dPri[n * lendSdP], bmat.data());
            for (auto& v : bmat) v = dt * v;
            {
                double* d = &val[(size_t)IA[n] * bsize];
                for (int q = 0; q < bsize; q++) d[q] += bmat[q];
            }
            // Bulk - Well: NewOffDiag(n, wId, dt * dQdXpW)
            bmat = dQdXpW;
            for (auto& v : bmat) v = dt * v;
            std::memcpy(&val[(size_t)fill[n]++ * bsize], bmat.data(), sizeof(double) * bsize);
            if (W.mode[w] != OCPG_WELL_BHP) {
                // Well - Well: AddDiag(wId, ...)
                std::fill(bmat.begin(), bmat.end(), 0.0);
                for (int i = 0; i < nc; i++) {
                    bmat[0] += inj ? dQdXpW[(i + 1) * ncol] : dQdXpW[(i + 1) * ncol] * weight;
                    bmat[(i + 1) * ncol + i + 1] = 1;
                }
                for (int q = 0; q < bsize; q++) wdiag[q] += bmat[q];
                // Well - Bulk: NewOffDiag(wId, n, sum_i weight * row i+1 of (dQdXpB + dQdXsB * dSec))
                bmat = dQdXpB;
                DaABpbCw(ncol, ncol, ncolB, dQdXsB.data(), &bk.dSec_dPri[n * lendSdP], bmat.data());
                std::fill(bmat2.begin(), bmat2.end(), 0.0);
                for (int i = 0; i < nc; i++)
                    for (int c = 0; c < ncol; c++) bmat2[c] += weight * bmat[(i + 1) * ncol + c];
                std::memcpy(&val[(size_t)fill[wId]++ * bsize], bmat2.data(), sizeof(double) * bsize);
            } else {
                std::fill(bmat.begin(), bmat.end(), 0.0);
                for (int i = 0; i < ncol; i++) bmat[i * ncol + i] = 1;
                for (int q = 0; q < bsize; q++) wdiag[q] += bmat[q];
                std::memset(&val[(size_t)fill[wId]++ * bsize], 0, sizeof(double) * bsize);
            }
        }
        wId++;
    }
}

}  // namespace oracle
