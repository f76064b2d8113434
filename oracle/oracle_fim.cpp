// CPU restatement of the FIM property pass and assembly of IsoT_FIM.
// TEST INFRASTRUCTURE ONLY (oracle + CPU baseline); never in the product path.
//
// Restates, on the reference's cell-major Bulk layout:
//   IsoT_FIM::CalFlash + PassFlashValue   src/OCPFluidMethod.cpp:1108-1187
//   IsoT_FIM::InitFlash                    src/OCPFluidMethod.cpp:1095-1106
//   IsoT_FIM::CalKrPc -> FlowUnit_ODGW01   src/OCPFluidMethod.cpp:1189-1199,
//                                          src/FlowUnit.cpp:235-328
//   IsothermalMethod::CalRock              src/OCPFluidMethod.cpp:25-32, src/Rock.cpp:19-27
//   IsoT_FIM::CalRes (bulk part)           src/OCPFluidMethod.cpp:1201-1389
//   IsoT_FIM::AssembleMatBulksNew          src/OCPFluidMethod.cpp:1569-1839
#include "oracle_fim.hpp"

#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {



// IsoT_FIM::PassFlashValue, src/OCPFluidMethod.cpp:1122-1187 (Newton
// bookkeeping dSNR/maxNRdSSP is not part of this path).
static void PassFlashValue(const Mixture& mix, const ocpg_bulk& bk, int64_t n)
{
    const int np = NPH;
    const int nc = mix.numCom;
    const int64_t bIdp = n * np;
    int len = 0;
    bk.phaseNum[n] = 0;
    bk.Nt[n] = todbl(mix.Nt);
    bk.vf[n] = todbl(mix.vf);
    for (int j = 0; j < np; j++) {
        bk.S[bIdp + j] = todbl(mix.S[j]);
        bk.phaseExist[bIdp + j] = mix.phaseExist[j];
        if (mix.phaseExist[j]) {
            bk.phaseNum[n]++;
            bk.rho[bIdp + j] = todbl(mix.rho[j]);
            bk.xi[bIdp + j] = todbl(mix.xi[j]);
            bk.mu[bIdp + j] = todbl(mix.mu[j]);
            bk.rhoP[bIdp + j] = todbl(mix.rhoP[j]);
            bk.xiP[bIdp + j] = todbl(mix.xiP[j]);
            bk.muP[bIdp + j] = todbl(mix.muP[j]);
            for (int i = 0; i < nc; i++) {
                bk.xij[bIdp * nc + j * nc + i] = todbl(mix.xij[j * nc + i]);
                bk.rhox[bIdp * nc + j * nc + i] = todbl(mix.rhox[j * nc + i]);
                bk.xix[bIdp * nc + j * nc + i] = todbl(mix.xix[j * nc + i]);
                bk.mux[bIdp * nc + j * nc + i] = todbl(mix.mux[j * nc + i]);
            }
        }
        bk.pSderExist[bIdp + j] = mix.pSderExist[j];
        bk.pVnumCom[bIdp + j] = mix.pVnumCom[j];
        if (mix.pSderExist[j]) len++;
        len += mix.pVnumCom[j];
    }
    bk.vfP[n] = todbl(mix.vfP);
    for (int i = 0; i < nc; i++) bk.vfi[n * nc + i] = todbl(mix.vfi[i]);
    bk.bRowSizedSdP[n] = len;
    const int L = maxLendSdP(nc);
    for (int k = 0; k < len * (nc + 1); k++) bk.dSec_dPri[n * L + k] = todbl(mix.dXsdXp[k]);
    if (bk.ftype) bk.ftype[n] = mix.ftype;
    if (bk.iters)
        for (int q = 0; q < 5; q++) bk.iters[n * 5 + q] = (int32_t)mix.iters[q];
    // MixtureComp's ePEC bookkeeping (src/MixtureComp.cpp:1301-1346): 1e8
    // where neither the Newton nor the SSM stage of the outcome converged;
    // 2 where the results are not finite (the product's own flag)
    if (bk.nonconv) {
        bk.nonconv[n] = (todbl(mix.ePEC) >= 1e8) ? 1 : 0;
        bool fin = std::isfinite(bk.vf[n]);
        for (int j = 0; j < np; j++) fin = fin && std::isfinite(bk.S[bIdp + j]);
        if (!fin) bk.nonconv[n] = 2;
    }
}

static void LoadSkip(const ocpg_bulk& bk, int NC, int64_t n, SkipState& s)
{
    s.flag = bk.skipFlag[n] != 0;
    s.minEigen = bk.skipMinEig[n];
    s.P = bk.skipP[n];
    s.T = bk.skipT[n];
    for (int i = 0; i < NC; i++) s.zi[i] = bk.skipZi[n * NC + i];
}

static void StoreSkip(const ocpg_bulk& bk, int NC, int64_t n, const SkipState& s)
{
    bk.skipFlag[n] = s.flag ? 1 : 0;
    bk.skipMinEig[n] = s.minEigen;
    bk.skipP[n] = s.P;
    bk.skipT[n] = s.T;
    for (int i = 0; i < NC; i++) bk.skipZi[n * NC + i] = s.zi[i];
}

void CalFlash(const ocpg_eos& eos, const ocpg_bulk& bk, int nthreads)
{
    const int NC = eos.nc;
    const int nc = NC + 1;
    const int np = NPH;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        Mixture mix(eos);
#pragma omp for schedule(dynamic, 256)
        for (int64_t n = 0; n < bk.nb; n++) {
            SkipState s;
            LoadSkip(bk, NC, n, s);
            mix.FlashFIM(bk.P[n], bk.T[n], &bk.Ni[n * nc], &bk.S[n * np], bk.phaseNum[n],
                         &bk.xij[n * np * nc], s);
            StoreSkip(bk, NC, n, s);
            PassFlashValue(mix, bk, n);
        }
    }
}

// IsoT_FIM::InitFlash: Ni is an output (from z, Sw and the pore volume).
void InitFlash(const ocpg_eos& eos, const ocpg_bulk& bk, const double* zi, int nthreads)
{
    const int NC = eos.nc;
    const int nc = NC + 1;
    const int np = NPH;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        Mixture mix(eos);
#pragma omp for schedule(dynamic, 256)
        for (int64_t n = 0; n < bk.nb; n++) {
            SkipState s;
            LoadSkip(bk, NC, n, s);
            mix.InitFlashFIM(bk.P[n], bk.T[n], &bk.S[n * np], bk.rockVp[n], &zi[n * NC], s);
            StoreSkip(bk, NC, n, s);
            for (int i = 0; i < nc; i++) bk.Ni[n * nc + i] = todbl(mix.Ni[i]);
            PassFlashValue(mix, bk, n);
        }
    }
}

// IsoT_FIMn::PassFlashValue, src/OCPFluidMethod.cpp:3179-3196: IsoT_FIM's,
// then nj of each existing phase, res (bRowSizedSdP entries) and resPc.
static void PassFlashValueN(const Mixture& mix, const ocpg_bulk& bk, int64_t n)
{
    PassFlashValue(mix, bk, n);
    const int np = NPH;
    const int nc = mix.numCom;
    for (int j = 0; j < np; j++)
        if (bk.phaseExist[n * np + j]) bk.nj[n * np + j] = todbl(mix.nj[j]);
    const int len = bk.bRowSizedSdP[n];
    for (int k = 0; k < len; k++) bk.res_n[n * (np * (nc + 1)) + k] = todbl(mix.resN[k]);
    bk.resPc[n] = todbl(mix.resPc);
}

void CalFlashN(const ocpg_eos& eos, const ocpg_bulk& bk, int nthreads)
{
    const int NC = eos.nc;
    const int nc = NC + 1;
    const int np = NPH;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        Mixture mix(eos);
#pragma omp for schedule(dynamic, 256)
        for (int64_t n = 0; n < bk.nb; n++) {
            SkipState s;
            LoadSkip(bk, NC, n, s);
            mix.FlashFIMn(bk.P[n], bk.T[n], &bk.Ni[n * nc], &bk.S[n * np], bk.phaseNum[n],
                          &bk.xij[n * np * nc], s);
            StoreSkip(bk, NC, n, s);
            PassFlashValueN(mix, bk, n);
        }
    }
}

void InitFlashN(const ocpg_eos& eos, const ocpg_bulk& bk, const double* zi, int nthreads)
{
    const int NC = eos.nc;
    const int nc = NC + 1;
    const int np = NPH;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        Mixture mix(eos);
#pragma omp for schedule(dynamic, 256)
        for (int64_t n = 0; n < bk.nb; n++) {
            SkipState s;
            LoadSkip(bk, NC, n, s);
            mix.InitFlashFIMn(bk.P[n], bk.T[n], &bk.S[n * np], bk.rockVp[n], &zi[n * NC], s);
            StoreSkip(bk, NC, n, s);
            for (int i = 0; i < nc; i++) bk.Ni[n * nc + i] = todbl(mix.Ni[i]);
            PassFlashValueN(mix, bk, n);
        }
    }
}

// OCPTable::Eval_All on column 0 of a 4-column table.
static void TableEval4(const double (*tab)[4], int nRow, double val, double* out, double* slope)
{
    int seg = -1;
    for (int i = 0; i + 1 < nRow; i++) {
        if (val >= tab[i][0] && val < tab[i + 1][0]) {
            seg = i;
            break;
        }
    }
    if (seg < 0) {
        const int r = (val < tab[0][0]) ? 0 : nRow - 1;
        for (int k = 0; k < 4; k++) {
            slope[k] = 0;
            out[k] = tab[r][k];
        }
        return;
    }
    for (int k = 0; k < 4; k++) {
        slope[k] = (tab[seg + 1][k] - tab[seg][k]) / (tab[seg + 1][0] - tab[seg][0]);
        out[k] = tab[seg][k] + slope[k] * (val - tab[seg][0]);
    }
}

// FlowUnit_ODGW01::CalKrPcDeriv + CalKro_Stone2Der, src/FlowUnit.cpp:235-328.
static void KrPcDeriv(const ocpg_satfunc& sat, const double* S_in, double* kr_out, double* pc_out,
                      double* dkrdS, double* dPcjdS)
{
    double data[4], cdata[4];
    const double Sg = S_in[1];
    const double Sw = S_in[2];
    TableEval4(sat.swof, sat.swof_rows, Sw, data, cdata);
    const double krw = data[1];
    const double dKrwdSw = cdata[1];
    const double krow = data[2];
    const double dKrowdSw = cdata[2];
    const double Pcwo = -data[3];
    const double dPcwdSw = -cdata[3];
    TableEval4(sat.sgof, sat.sgof_rows, Sg, data, cdata);
    const double krg = data[1];
    const double dKrgdSg = cdata[1];
    const double krog = data[2];
    const double dKrogdSg = cdata[2];
    const double Pcgo = data[3];
    const double dPcgdSg = cdata[3];

    const double kroMax = sat.swof[0][2];
    double kro = kroMax * ((krow / kroMax + krw) * (krog / kroMax + krg) - (krw + krg));
    double dKrodSw = kroMax * ((dKrowdSw / kroMax + dKrwdSw) * (krog / kroMax + krg) - (dKrwdSw));
    double dKrodSg = kroMax * ((krow / kroMax + krw) * (dKrogdSg / kroMax + dKrgdSg) - (dKrgdSg));
    if (kro < 0) {
        kro = 0;
        dKrodSg = 0;
        dKrodSw = 0;
    }
    kr_out[0] = kro;
    kr_out[1] = krg;
    kr_out[2] = krw;
    pc_out[0] = 0;
    pc_out[1] = Pcgo;
    pc_out[2] = Pcwo;
    dkrdS[0] = 0;
    dkrdS[1] = dKrodSg;
    dkrdS[2] = dKrodSw;
    dkrdS[3] = 0;
    dkrdS[4] = dKrgdSg;
    dkrdS[5] = 0;
    dkrdS[6] = 0;
    dkrdS[7] = 0;
    dkrdS[8] = dKrwdSw;
    dPcjdS[0] = 0;
    dPcjdS[1] = 0;
    dPcjdS[2] = 0;
    dPcjdS[3] = 0;
    dPcjdS[4] = dPcgdSg;
    dPcjdS[5] = 0;
    dPcjdS[6] = 0;
    dPcjdS[7] = 0;
    dPcjdS[8] = dPcwdSw;
}

void CalKrPc(const ocpg_satfunc& sat, const ocpg_bulk& bk, int nthreads)
{
    const int np = NPH;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t n = 0; n < bk.nb; n++) {
        const int64_t bId = n * np;
        KrPcDeriv(sat, &bk.S[bId], &bk.kr[bId], &bk.Pc[bId], &bk.dKr_dS[bId * np],
                  &bk.dPcj_dS[bId * np]);
        for (int j = 0; j < np; j++) bk.Pj[bId + j] = bk.P[n] + bk.Pc[bId + j];
    }
}

void CalRock(const ocpg_rock& rock, const ocpg_bulk& bk, int nthreads)
{
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t n = 0; n < bk.nb; n++) {
        const double dP = (bk.P[n] - rock.Pref);
        bk.poro[n] = bk.poroInit[n] * (1 + (rock.cp1 + rock.cp2 / 2 * dP) * dP);
        bk.poroP[n] = bk.poroInit[n] * (rock.cp1 + rock.cp2 * dP);
        bk.rockVp[n] = bk.v[n] * bk.poro[n];
    }
}

// Each cell's connections in ascending index order (CSR), and each
// connection's position in its two cells' lists: the slot order in which
// LinearSystem::NewOffDiag fills a row while the connection loop runs.
CellConn CellConnections(int64_t nb, const ocpg_conn& conn)
{
    CellConn cc;
    const int64_t nconn = conn.nconn;
    cc.ptr.assign(nb + 1, 0);
    for (int64_t c = 0; c < nconn; c++) {
        cc.ptr[conn.bId[c] + 1]++;
        cc.ptr[conn.eId[c] + 1]++;
    }
    for (int64_t n = 0; n < nb; n++) cc.ptr[n + 1] += cc.ptr[n];
    cc.conn.resize(cc.ptr[nb]);
    cc.posB.resize(nconn);
    cc.posE.resize(nconn);
    std::vector<int64_t> fill(cc.ptr.begin(), cc.ptr.end() - 1);
    for (int64_t c = 0; c < nconn; c++) {
        const int64_t b = conn.bId[c], e = conn.eId[c];
        cc.posE[c] = (int32_t)(fill[e] - cc.ptr[e]);
        cc.conn[fill[e]++] = c;
        cc.posB[c] = (int32_t)(fill[b] - cc.ptr[b]);
        cc.conn[fill[b]++] = c;
    }
    return cc;
}

// IsoT_FIM::CalRes (src/OCPFluidMethod.cpp:1201-1389): bulk-to-bulk part
// serial in connection order as the reference accumulates, then the well
// sources and well equations (:1284-1352), norms and negation.
void CalRes(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt, double* resAbs,
            int32_t* upblock, double* upblock_Rho, double* upblock_Velocity, ocpg_resnorm* norms,
            const ocpg_wells* wells, const WellFlux* wf, int nthreads)
{
    const int np = NPH;
    const int len = nc + 1;
    const int64_t nb = bk.nb;
    const int64_t nrow = nb + OpenWells(wells);
    std::memset(resAbs, 0, sizeof(double) * nrow * len);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t n = 0; n < nb; n++) {
        const int64_t bId = n * len;
        const int64_t bIdb = n * nc;
        resAbs[bId] = bk.rockVp[n] - bk.vf[n];
        for (int i = 0; i < nc; i++) resAbs[bId + 1 + i] = bk.Ni[bIdb + i] - bk.lNi[bIdb + i];
    }
    // Per connection (independent): upwinding, and per phase the flux factor
    // dt Akd xi kr / mu dP of the upstream cell (flag: the phase flows).
    const int64_t nconn = conn.nconn;
    std::vector<double> fac((size_t)nconn * np);
    std::vector<uint8_t> flows((size_t)nconn * np);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t c = 0; c < nconn; c++) {
        const int64_t bId = conn.bId[c];
        const int64_t eId = conn.eId[c];
        const double Akd = CONV1 * CONV2 * conn.area[c];
        for (int j = 0; j < np; j++) {
            flows[c * np + j] = 0;
            const int64_t bId_np_j = bId * np + j;
            const int64_t eId_np_j = eId * np + j;
            const bool exbegin = bk.phaseExist[bId_np_j];
            const bool exend = bk.phaseExist[eId_np_j];
            double rho;
            if (exbegin && exend) {
                rho = (bk.rho[bId_np_j] + bk.rho[eId_np_j]) / 2;
            } else if (exbegin && (!exend)) {
                rho = bk.rho[bId_np_j];
            } else if ((!exbegin) && exend) {
                rho = bk.rho[eId_np_j];
            } else {
                upblock[c * np + j] = (int32_t)bId;
                upblock_Rho[c * np + j] = 0;
                continue;
            }
            int64_t uId = bId;
            const double dP = (bk.Pj[bId_np_j] - GRAVITY_FACTOR * rho * bk.depth[bId]) -
                              (bk.Pj[eId_np_j] - GRAVITY_FACTOR * rho * bk.depth[eId]);
            if (dP < 0) uId = eId;
            upblock_Rho[c * np + j] = rho;
            upblock[c * np + j] = (int32_t)uId;
            const int64_t uId_np_j = uId * np + j;
            if (bk.phaseExist[uId_np_j]) {
                upblock_Velocity[c * np + j] = Akd * bk.kr[uId_np_j] / bk.mu[uId_np_j] * dP;
            } else {
                upblock_Velocity[c * np + j] = 0;
                continue;
            }
            fac[c * np + j] = dt * Akd * bk.xi[uId_np_j] * bk.kr[uId_np_j] / bk.mu[uId_np_j] * dP;
            flows[c * np + j] = 1;
        }
    }
    // Per row: its connections in the order the reference's connection loop
    // reaches them (ascending index), each adding (bId side) or subtracting
    // (eId side) its components' fluxes -- the same sums in the same order
    // as the connection-ordered scatter, so the rows are independent.
    const CellConn cc = CellConnections(nb, conn);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t n = 0; n < nb; n++) {
        double* r = &resAbs[n * len + 1];
        for (int64_t q = cc.ptr[n]; q < cc.ptr[n + 1]; q++) {
            const int64_t c = cc.conn[q];
            const bool isB = conn.bId[c] == n;
            for (int j = 0; j < np; j++) {
                if (!flows[c * np + j]) continue;
                const double tmp = fac[c * np + j];
                const int64_t uId_np_j = (int64_t)upblock[c * np + j] * np + j;
                for (int i = 0; i < nc; i++) {
                    const double dNi = tmp * bk.xij[uId_np_j * nc + i];
                    if (isB)
                        r[i] += dNi;
                    else
                        r[i] -= dNi;
                }
            }
        }
    }
    ocpg_resnorm nr = {0, 0, 0, 0, 0};
    if (wells && wf) {
        int64_t wId = nb * len;
        for (int w = 0; w < wells->nw; w++) {
            if (!wells->open[w]) continue;
            for (int p = wells->perfStart[w]; p < wells->perfStart[w + 1]; p++) {
                const int64_t k = wells->perfCell[p];
                for (int i = 0; i < nc; i++) resAbs[k * len + 1 + i] += wf->perfQi[(size_t)p * nc + i] * dt;
            }
            if (wells->type[w] == OCPG_WELL_INJ) {
                if (wells->mode[w] == OCPG_WELL_BHP) {
                    resAbs[wId] = wells->bhp[w] - wells->ctrlBHP[w];
                } else {
                    resAbs[wId] = wells->ctrlRate[w];
                    for (int i = 0; i < nc; i++) resAbs[wId] += wf->wellQi[(size_t)w * nc + i];
                    nr.maxWellRelRes_mol = std::max(nr.maxWellRelRes_mol, std::fabs(resAbs[wId] / wells->ctrlRate[w]));
                }
            } else {
                if (wells->mode[w] == OCPG_WELL_BHP) {
                    resAbs[wId] = wells->bhp[w] - wells->ctrlBHP[w];
                } else {
                    resAbs[wId] = -wells->ctrlRate[w];
                    for (int i = 0; i < nc; i++)
                        resAbs[wId] += wf->wellQi[(size_t)w * nc + i] * wf->prodWeight[w];
                    nr.maxWellRelRes_mol = std::max(nr.maxWellRelRes_mol, std::fabs(resAbs[wId] / wells->ctrlRate[w]));
                }
            }
            wId += len;
        }
    }
    for (int64_t n = 0; n < nb; n++) {
        for (int i = 0; i < len; i++) {
            const double tmp = std::fabs(resAbs[n * len + i] / bk.rockVp[n]);
            if (nr.maxRelRes_V < tmp) {
                nr.maxRelRes_V = tmp;
                nr.maxId_V = n;
            }
        }
        for (int i = 1; i < len; i++) {
            const double tmp = std::fabs(resAbs[n * len + i] / bk.Nt[n]);
            if (nr.maxRelRes_N < tmp) {
                nr.maxRelRes_N = tmp;
                nr.maxId_N = n;
            }
        }
    }
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t k = 0; k < nrow * len; k++) resAbs[k] = -resAbs[k];
    if (norms) *norms = nr;
}

// BSR structure of LinearSystem after AssembleMatBulksNew: row n holds the
// diagonal first, then neighbours in insertion order (ascending for a sorted
// connection list).  IA has nb+1 entries.
// With wells, AssembleMatWellsNew then appends to each perforated cell's row
// the column of the well (NewOffDiag(n, wId)) and adds one row per open well:
// its diagonal, then the perforated cells in perforation order.
void BsrPattern(int64_t nb, const ocpg_conn& conn, int32_t* IA, int32_t* JA, const ocpg_wells* wells)
{
    const int nw = OpenWells(wells);
    std::vector<std::vector<int32_t>> cols(nb + nw);
    for (int64_t n = 0; n < nb; n++) cols[n].push_back((int32_t)n);
    for (int64_t c = 0; c < conn.nconn; c++) {
        cols[conn.eId[c]].push_back(conn.bId[c]);
        cols[conn.bId[c]].push_back(conn.eId[c]);
    }
    if (wells) {
        int64_t wId = nb;
        for (int w = 0; w < wells->nw; w++) {
            if (!wells->open[w]) continue;
            cols[wId].push_back((int32_t)wId);
            for (int p = wells->perfStart[w]; p < wells->perfStart[w + 1]; p++) {
                const int32_t n = wells->perfCell[p];
                cols[n].push_back((int32_t)wId);
                cols[wId].push_back(n);
            }
            wId++;
        }
    }
    IA[0] = 0;
    int64_t k = 0;
    for (int64_t n = 0; n < nb + nw; n++) {
        for (int32_t col : cols[n]) JA[k++] = col;
        IA[n + 1] = (int32_t)k;
    }
}

// DaABpbC, src/DenseMat.cpp:75-94: C(m x n) += A(m x k) B(k x n), row-major,
// accumulated over k in order (reference BLAS dgemm loop order).
static void DaABpbC(int m, int n, int k, const double* A, const double* B, double* C)
{
    for (int r = 0; r < m; r++)
        for (int p = 0; p < k; p++) {
            const double a = A[r * k + p];
            for (int c = 0; c < n; c++) C[r * n + c] = C[r * n + c] + a * B[p * n + c];
        }
}

// IsoT_FIM::AssembleMatBulksNew, src/OCPFluidMethod.cpp:1569-1839, into a BSR
// whose pattern BsrPattern() produced.  val has IA[nb]*ncol*ncol entries.
static void DaAxpby(int m, int n, double a, const double* A, const double* x, double b, double* y);

// With fimn, IsoT_FIMn::AssembleMatBulksNew (src/OCPFluidMethod.cpp:
// 3198-3501): the secondary variables are S and n_ij (x_ij = n_ij / nj, two
// more terms in the n_ij columns of the upwind side) and rhs (nb * ncol)
// receives the bulk part of the right-hand side before resAbs is added:
// -resPc on each volume row, then, in the connection loop's order, each
// side's -dFdXs res_n where that cell has two hydrocarbon phases.
void AssembleBulks(int nc, const ocpg_conn& conn, const ocpg_bulk& bk, double dt,
                   const int32_t* upblock, const double* upblock_Rho, const int32_t* IA,
                   const int32_t* JA, double* val, int nthreads, bool fimn, double* rhs)
{
    const int np = NPH;
    const int ncol = nc + 1;
    const int ncol2 = np * nc + np;
    const int bsize = ncol * ncol;
    const int bsize2 = ncol * ncol2;
    const int lendSdP = maxLendSdP(nc);
    const int64_t nb = bk.nb;
    const int nt = nthreads > 0 ? nthreads : 1;
    const CellConn cc = CellConnections(nb, conn);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t n = 0; n < nb; n++) {
        double* d = &val[(size_t)IA[n] * bsize];
        std::memset(d, 0, sizeof(double) * (size_t)(IA[n + 1] - IA[n]) * bsize);
        for (int i = 1; i < ncol; i++) d[i * ncol + i] = 1;
        d[0] = bk.v[n] * bk.poroP[n] - bk.vfP[n];
        for (int i = 0; i < nc; i++) d[i + 1] = -bk.vfi[n * nc + i];
    }

    // Per connection (independent): both sides' condensed blocks, written to
    // the two off-diagonal slots the reference's NewOffDiag calls fill
    // ((eId, bId): -b side, (bId, eId): e side).  The diagonals follow.
    std::vector<double> rhsB, rhsE;  // per connection: each side's -dFdXs res_n
    std::vector<uint8_t> hasB, hasE;
    if (fimn) {
        rhsB.assign((size_t)conn.nconn * ncol, 0.0);
        rhsE.assign((size_t)conn.nconn * ncol, 0.0);
        hasB.assign(conn.nconn, 0);
        hasE.assign(conn.nconn, 0);
    }
#pragma omp parallel num_threads(nt)
    {
    std::vector<double> bmat(bsize, 0);
    std::vector<double> dFdXpB(bsize), dFdXpE(bsize), dFdXsB(bsize2), dFdXsE(bsize2);
    bool phaseExistB[NPH], phaseExistE[NPH], phasedS_B[NPH], phasedS_E[NPH];
    int pVnumComB[NPH], pVnumComE[NPH];
#pragma omp for schedule(static)
    for (int64_t c = 0; c < conn.nconn; c++) {
        const int64_t bId = conn.bId[c];
        const int64_t eId = conn.eId[c];
        const double Akd = CONV1 * CONV2 * conn.area[c];
        std::fill(dFdXpB.begin(), dFdXpB.end(), 0.0);
        std::fill(dFdXpE.begin(), dFdXpE.end(), 0.0);
        std::fill(dFdXsB.begin(), dFdXsB.end(), 0.0);
        std::fill(dFdXsE.begin(), dFdXsE.end(), 0.0);
        const double dGamma = GRAVITY_FACTOR * (bk.depth[bId] - bk.depth[eId]);
        int jxB = 0, jxE = 0, ncolB = 0, ncolE = 0;
        for (int j = 0; j < np; j++) {
            phaseExistB[j] = bk.phaseExist[bId * np + j];
            phaseExistE[j] = bk.phaseExist[eId * np + j];
            phasedS_B[j] = bk.pSderExist[bId * np + j];
            phasedS_E[j] = bk.pSderExist[eId * np + j];
            if (phasedS_B[j]) jxB++;
            if (phasedS_E[j]) jxE++;
            pVnumComB[j] = bk.pVnumCom[bId * np + j];
            pVnumComE[j] = bk.pVnumCom[eId * np + j];
            ncolB += pVnumComB[j];
            ncolE += pVnumComE[j];
        }
        ncolB += jxB;
        ncolE += jxE;
        for (int j = 0; j < np; j++) {
            const int64_t uId = upblock[c * np + j];
            const bool phaseExistU = (uId == bId ? phaseExistB[j] : phaseExistE[j]);
            if (!phaseExistU) {
                jxB += pVnumComB[j];
                jxE += pVnumComE[j];
                continue;
            }
            const int64_t bId_np_j = bId * np + j;
            const int64_t eId_np_j = eId * np + j;
            const int64_t uId_np_j = uId * np + j;
            const double dP = bk.Pj[bId_np_j] - bk.Pj[eId_np_j] - upblock_Rho[c * np + j] * dGamma;
            const double xi = bk.xi[uId_np_j];
            const double kr = bk.kr[uId_np_j];
            const double mu = bk.mu[uId_np_j];
            const double muP = bk.muP[uId_np_j];
            const double xiP = bk.xiP[uId_np_j];
            const double rhoP = bk.rhoP[uId_np_j];
            const double transJ = Akd * kr / mu;
            for (int i = 0; i < nc; i++) {
                const double xij = bk.xij[uId_np_j * nc + i];
                const double transIJ = xij * xi * transJ;
                dFdXpB[(i + 1) * ncol] += transIJ;
                dFdXpE[(i + 1) * ncol] -= transIJ;
                double tmp = xij * transJ * xiP * dP;
                tmp += -transIJ * muP / mu * dP;
                if (!phaseExistE[j]) {
                    tmp += transIJ * (-rhoP * dGamma);
                    dFdXpB[(i + 1) * ncol] += tmp;
                } else if (!phaseExistB[j]) {
                    tmp += transIJ * (-rhoP * dGamma);
                    dFdXpE[(i + 1) * ncol] += tmp;
                } else {
                    dFdXpB[(i + 1) * ncol] += transIJ * (-bk.rhoP[bId_np_j] * dGamma) / 2;
                    dFdXpE[(i + 1) * ncol] += transIJ * (-bk.rhoP[eId_np_j] * dGamma) / 2;
                    if (bId == uId)
                        dFdXpB[(i + 1) * ncol] += tmp;
                    else
                        dFdXpE[(i + 1) * ncol] += tmp;
                }
                int j1SB = 0, j1SE = 0;
                if (bId == uId) {
                    for (int j1 = 0; j1 < np; j1++) {
                        if (phasedS_B[j1]) {
                            dFdXsB[(i + 1) * ncolB + j1SB] += transIJ * bk.dPcj_dS[bId_np_j * np + j1];
                            tmp = Akd * xij * xi / mu * bk.dKr_dS[uId_np_j * np + j1] * dP;
                            dFdXsB[(i + 1) * ncolB + j1SB] += tmp;
                            j1SB++;
                        }
                        if (phasedS_E[j1]) {
                            dFdXsE[(i + 1) * ncolE + j1SE] -= transIJ * bk.dPcj_dS[eId_np_j * np + j1];
                            j1SE++;
                        }
                    }
                    if (!phaseExistE[j]) {
                        for (int k = 0; k < pVnumComB[j]; k++) {
                            const double rhox = bk.rhox[uId_np_j * nc + k];
                            const double xix = bk.xix[uId_np_j * nc + k];
                            const double mux = bk.mux[uId_np_j * nc + k];
                            tmp = -transIJ * rhox * dGamma;
                            tmp += xij * transJ * xix * dP;
                            tmp += -transIJ * mux / mu * dP;
                            if (fimn) tmp -= transIJ * dP / bk.nj[uId_np_j];
                            dFdXsB[(i + 1) * ncolB + jxB + k] += tmp;
                        }
                        if (i < pVnumComB[j])
                            dFdXsB[(i + 1) * ncolB + jxB + i] +=
                                fimn ? xi * transJ * dP / bk.nj[uId_np_j] : xi * transJ * dP;
                    } else {
                        for (int k = 0; k < pVnumComB[j]; k++) {
                            const double rhox = bk.rhox[bId_np_j * nc + k] / 2;
                            const double xix = bk.xix[uId_np_j * nc + k];
                            const double mux = bk.mux[uId_np_j * nc + k];
                            tmp = -transIJ * rhox * dGamma;
                            tmp += xij * transJ * xix * dP;
                            tmp += -transIJ * mux / mu * dP;
                            if (fimn) tmp -= transIJ * dP / bk.nj[uId_np_j];
                            dFdXsB[(i + 1) * ncolB + jxB + k] += tmp;
                            dFdXsE[(i + 1) * ncolE + jxE + k] +=
                                -transIJ * bk.rhox[eId_np_j * nc + k] / 2 * dGamma;
                        }
                        if (i < pVnumComB[j])
                            dFdXsB[(i + 1) * ncolB + jxB + i] +=
                                fimn ? xi * transJ * dP / bk.nj[uId_np_j] : xi * transJ * dP;
                    }
                } else {
                    for (int j1 = 0; j1 < np; j1++) {
                        if (phasedS_B[j1]) {
                            dFdXsB[(i + 1) * ncolB + j1SB] += transIJ * bk.dPcj_dS[bId_np_j * np + j1];
                            j1SB++;
                        }
                        if (phasedS_E[j1]) {
                            dFdXsE[(i + 1) * ncolE + j1SE] -= transIJ * bk.dPcj_dS[eId_np_j * np + j1];
                            tmp = Akd * xij * xi / mu * bk.dKr_dS[uId_np_j * np + j1] * dP;
                            dFdXsE[(i + 1) * ncolE + j1SE] += tmp;
                            j1SE++;
                        }
                    }
                    if (!phaseExistB[j]) {
                        for (int k = 0; k < pVnumComE[j]; k++) {
                            const double rhox = bk.rhox[uId_np_j * nc + k];
                            const double xix = bk.xix[uId_np_j * nc + k];
                            const double mux = bk.mux[uId_np_j * nc + k];
                            tmp = -transIJ * rhox * dGamma;
                            tmp += xij * transJ * xix * dP;
                            tmp += -transIJ * mux / mu * dP;
                            if (fimn) tmp -= transIJ * dP / bk.nj[uId_np_j];
                            dFdXsE[(i + 1) * ncolE + jxE + k] += tmp;
                        }
                        if (i < pVnumComE[j])
                            dFdXsE[(i + 1) * ncolE + jxE + i] +=
                                fimn ? xi * transJ * dP / bk.nj[uId_np_j] : xi * transJ * dP;
                    } else {
                        for (int k = 0; k < pVnumComE[j]; k++) {
                            const double rhox = bk.rhox[eId_np_j * nc + k] / 2;
                            const double xix = bk.xix[uId_np_j * nc + k];
                            const double mux = bk.mux[uId_np_j * nc + k];
                            tmp = -transIJ * rhox * dGamma;
                            tmp += xij * transJ * xix * dP;
                            tmp += -transIJ * mux / mu * dP;
                            if (fimn) tmp -= transIJ * dP / bk.nj[uId_np_j];
                            dFdXsE[(i + 1) * ncolE + jxE + k] += tmp;
                            dFdXsB[(i + 1) * ncolB + jxB + k] +=
                                -transIJ * bk.rhox[bId_np_j * nc + k] / 2 * dGamma;
                        }
                        if (i < pVnumComE[j])
                            dFdXsE[(i + 1) * ncolE + jxE + i] +=
                                fimn ? xi * transJ * dP / bk.nj[uId_np_j] : xi * transJ * dP;
                    }
                }
            }
            jxB += pVnumComB[j];
            jxE += pVnumComE[j];
        }
        if (fimn && bk.phaseNum[bId] > 2) {
            DaAxpby(ncol, ncolB, -1.0, dFdXsB.data(), &bk.res_n[bId * ncol2], 1.0, &rhsB[(size_t)c * ncol]);
            hasB[c] = 1;
        }
        if (fimn && bk.phaseNum[eId] > 2) {
            DaAxpby(ncol, ncolE, -1.0, dFdXsE.data(), &bk.res_n[eId * ncol2], 1.0, &rhsE[(size_t)c * ncol]);
            hasE[c] = 1;
        }
        // b side: AddDiag(bId, bmat) ; NewOffDiag(eId, bId, -bmat)
        bmat = dFdXpB;
        DaABpbC(ncol, ncol, ncolB, dFdXsB.data(), &bk.dSec_dPri[bId * lendSdP], bmat.data());
        for (auto& v : bmat) v = dt * v;
        {
            double* o = &val[(size_t)(IA[eId] + 1 + cc.posE[c]) * bsize];
            for (int q = 0; q < bsize; q++) o[q] = -1.0 * bmat[q];
        }
        // e side: NewOffDiag(bId, eId, bmat) ; AddDiag(eId, -bmat)
        bmat = dFdXpE;
        DaABpbC(ncol, ncol, ncolE, dFdXsE.data(), &bk.dSec_dPri[eId * lendSdP], bmat.data());
        for (auto& v : bmat) v = dt * v;
        {
            double* o = &val[(size_t)(IA[bId] + 1 + cc.posB[c]) * bsize];
            for (int q = 0; q < bsize; q++) o[q] = bmat[q];
        }
    }
    }
    // Per row: AddDiag in the connection loop's order.  The b side added
    // bmat, the block its (eId, bId) slot holds negated; the e side added
    // -bmat, its (bId, eId) slot negated: either way the negation of the
    // mirror block, so the diagonal is the reference's sum term by term.
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t n = 0; n < nb; n++) {
        double* d = &val[(size_t)IA[n] * bsize];
        for (int64_t q = cc.ptr[n]; q < cc.ptr[n + 1]; q++) {
            const int64_t c = cc.conn[q];
            const bool isB = conn.bId[c] == n;
            const int64_t m = isB ? conn.eId[c] : conn.bId[c];
            const int32_t pos = isB ? cc.posE[c] : cc.posB[c];
            const double* o = &val[(size_t)(IA[m] + 1 + pos) * bsize];
            for (int k = 0; k < bsize; k++) d[k] += -o[k];
        }
    }
    // The right-hand side's bulk part, each row's terms in connection order.
    if (fimn && rhs) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t n = 0; n < nb; n++) {
            double* b = &rhs[n * ncol];
            for (int i = 0; i < ncol; i++) b[i] = 0;
            b[0] += -bk.resPc[n];
            for (int64_t q = cc.ptr[n]; q < cc.ptr[n + 1]; q++) {
                const int64_t c = cc.conn[q];
                const bool isB = conn.bId[c] == n;
                if (isB ? !hasB[c] : !hasE[c]) continue;
                const double* t = isB ? &rhsB[(size_t)c * ncol] : &rhsE[(size_t)c * ncol];
                for (int i = 0; i < ncol; i++) b[i] += t[i];
            }
        }
    }
    (void)JA;
}

// DaAxpby, src/DenseMat.cpp:203-219: y = a A x + b y (A m x n, row-major).
static void DaAxpby(int m, int n, double a, const double* A, const double* x, double b, double* y)
{
    for (int i = 0; i < m; i++) {
        y[i] = b * y[i];
        for (int j = 0; j < n; j++) y[i] += a * A[i * n + j] * x[j];
    }
}

// IsoT_FIM::GetSolution, bulk part (src/OCPFluidMethod.cpp:2735-2844): the
// secondary changes dSec_dPri u, the saturation chop, the updates of S,
// x_ij (two hydrocarbon phases), P and Ni, and NRdPmax / NRdNmax (signed,
// the first cell / component at the largest magnitude).  Also keeps what
// the next PassFlashValue reads (:1135-1146): the saturations before the
// update (dSNR), the predicted changes (dSNRP) and the phase flags.
void GetSolution(int nc, const ocpg_bulk& bk, const double* u, double dSmaxlim, double* dSNR, double* dSNRP,
                 uint8_t* phaseExistOld, double* NRdPmax, double* NRdNmax)
{
    const int np = NPH;
    const int NC = nc - 1;
    const int col = nc + 1;
    const int row = np * (nc + 1);
    const int lenSdP = maxLendSdP(nc);
    std::vector<double> dtmp(row, 0.0);
    *NRdPmax = 0;
    *NRdNmax = 0;
    for (int64_t k = 0; k < bk.nb * np; k++) {
        dSNR[k] = bk.S[k];
        phaseExistOld[k] = bk.phaseExist[k];
    }
    for (int64_t n = 0; n < bk.nb; n++) {
        double chopmin = 1;
        std::fill(dtmp.begin(), dtmp.end(), 0.0);
        DaAxpby(bk.bRowSizedSdP[n], col, 1, &bk.dSec_dPri[n * lenSdP], u + n * col, 1, dtmp.data());
        int js = 0;
        for (int j = 0; j < np; j++) {
            if (!bk.pSderExist[n * np + j]) continue;
            double choptmp = 1;
            if (std::fabs(dtmp[js]) > dSmaxlim) {
                choptmp = dSmaxlim / std::fabs(dtmp[js]);
            } else if (bk.S[n * np + j] + dtmp[js] < 0.0) {
                choptmp = 0.9 * bk.S[n * np + j] / std::fabs(dtmp[js]);
            }
            chopmin = std::min(chopmin, choptmp);
            js++;
        }
        js = 0;
        for (int j = 0; j < np; j++) {
            if (!bk.pSderExist[n * np + j]) {
                dSNRP[n * np + j] = 0;
                continue;
            }
            dSNRP[n * np + j] = chopmin * dtmp[js];
            bk.S[n * np + j] += dSNRP[n * np + j];
            js++;
        }
        if (bk.phaseNum[n] >= 3) {
            for (int j = 0; j < 2; j++) {
                const int64_t bId = n * np * nc + j * nc;
                for (int i = 0; i < NC; i++) {
                    bk.xij[bId + i] += chopmin * dtmp[js];
                    js++;
                }
            }
        }
        const double dP = u[n * col];
        if (std::fabs(*NRdPmax) < std::fabs(dP)) *NRdPmax = dP;
        bk.P[n] += dP;
        for (int i = 0; i < nc; i++) {
            const double dN = u[n * col + 1 + i] * chopmin;
            if (std::fabs(*NRdNmax) < std::fabs(dN) / bk.Nt[n]) *NRdNmax = dN / bk.Nt[n];
            bk.Ni[n * nc + i] += dN;
        }
    }
}

// IsoT_FIMn::GetSolution, bulk part (src/OCPFluidMethod.cpp:3847-3990):
// the secondary changes res_n + dSec_dPri u, the saturation chop (limit,
// critical saturations, non-negativity), the updates of nj and x_ij from
// the phase moles, P and Ni.  Saturations are not updated (the next flash
// sets them); dSNRP keeps the predicted changes.  The reference takes the
// count of n_ij rows of phase j from bk.pVnumCom[j], i.e. from cell 0's
// entry for every cell (:3933-3936); restated as such.
void GetSolutionN(int nc, const ocpg_bulk& bk, const double* u, double dSmaxlim, const double* scm, double* dSNR,
                  double* dSNRP, uint8_t* phaseExistOld, double* NRdPmax, double* NRdNmax)
{
    const int np = NPH;
    const int row = np * (nc + 1);
    const int col = nc + 1;
    const int lenSdP = maxLendSdP(nc);
    std::vector<double> dtmp(row, 0.0), tmpNij(np * nc, 0.0);
    *NRdPmax = 0;
    *NRdNmax = 0;
    for (int64_t k = 0; k < bk.nb * np; k++) {
        dSNR[k] = bk.S[k];
        phaseExistOld[k] = bk.phaseExist[k];
    }
    for (int64_t n = 0; n < bk.nb; n++) {
        const double dP = u[n * col];
        bk.P[n] += dP;
        if (std::fabs(*NRdPmax) < std::fabs(dP)) *NRdPmax = dP;
        double dSmax = 0;
        double chop = 1;
        const int cNp = bk.phaseNum[n];
        const int len = bk.bRowSizedSdP[n];
        for (int k = 0; k < len; k++) dtmp[k] = bk.res_n[n * row + k];
        DaAxpby(len, col, 1.0, &bk.dSec_dPri[n * lenSdP], u + n * col, 1.0, dtmp.data());
        int js = 0;
        for (int j = 0; j < np; j++) {
            if (bk.pSderExist[n * np + j]) {
                dSmax = std::max(std::fabs(dtmp[js]), dSmax);
                js++;
            }
        }
        if (dSmax > dSmaxlim) {
            chop = std::min(chop, dSmaxlim / dSmax);
            dSmax = dSmaxlim;
        }
        js = 0;
        for (int j = 0; j < np; j++) {
            const int64_t k = n * np + j;
            if (bk.pSderExist[k]) {
                if (std::fabs(bk.S[k] - scm[j]) > TINY && (bk.S[k] - scm[j]) / (chop * dtmp[js]) < 0)
                    chop *= std::min(1.0, -((bk.S[k] - scm[j]) / (chop * dtmp[js])));
                if (bk.S[k] + chop * dtmp[js] < 0) chop *= std::min(1.0, -bk.S[k] / (chop * dtmp[js]));
                js++;
            }
        }
        std::fill(tmpNij.begin(), tmpNij.end(), 0.0);
        js = 0;
        int jx = cNp;
        for (int j = 0; j < np; j++) {
            const int64_t k = n * np + j;
            if (bk.pSderExist[k]) {
                dSNRP[k] = chop * dtmp[js];
                js++;
                // Daxpy(nc, nj, xij, tmpNij)
                for (int i = 0; i < nc; i++) tmpNij[j * nc + i] += bk.nj[k] * bk.xij[k * nc + i];
                for (int i = 0; i < bk.pVnumCom[j]; i++) tmpNij[j * nc + i] += chop * dtmp[jx + i];
                jx += bk.pVnumCom[j];
                double s1 = 0;  // Dnorm1
                for (int i = 0; i < nc; i++) s1 += std::fabs(tmpNij[j * nc + i]);
                bk.nj[k] = s1;
                for (int i = 0; i < nc; i++) bk.xij[k * nc + i] = tmpNij[j * nc + i] / bk.nj[k];
            }
        }
        for (int i = 0; i < nc; i++) {
            const double dN = u[n * col + 1 + i] * chop;
            if (std::fabs(*NRdNmax) < std::fabs(dN) / bk.Nt[n]) *NRdNmax = dN / bk.Nt[n];
            bk.Ni[n * nc + i] += dN;
        }
    }
}

// After the next flash: PassFlashValue's dSNR = S - dSNR and maxNRdSSP
// (src/OCPFluidMethod.cpp:1135-1146, over the phases that existed before
// the flash) and Bulk::CalNRdSmax (src/Bulk.cpp:1235-1247).  out = {NRdSmax,
// its cell, maxNRdSSP, its cell}.
void NRChange(int64_t nb, const double* S, double* dSNR, const double* dSNRP, const uint8_t* phaseExistOld,
              double* out)
{
    const int np = NPH;
    double maxSSP = 0;
    int64_t iSSP = 0;
    for (int64_t n = 0; n < nb; n++)
        for (int j = 0; j < np; j++) {
            const int64_t k = n * np + j;
            dSNR[k] = S[k] - dSNR[k];
            if (phaseExistOld[k]) {
                if (std::fabs(maxSSP) < std::fabs(dSNR[k] - dSNRP[k])) {
                    maxSSP = dSNR[k] - dSNRP[k];
                    iSSP = n;
                }
            }
        }
    double dSmax = 0;
    int64_t index = 0;
    for (int64_t k = 0; k < nb * np; k++)
        if (std::fabs(dSmax) < std::fabs(dSNR[k])) {
            dSmax = dSNR[k];
            index = k;
        }
    out[0] = dSmax;
    out[1] = (double)(index / np);
    out[2] = maxSSP;
    out[3] = (double)iSSP;
}

}  // namespace oracle
