// FIM kernels over the SoA cell state:
//   krPcRockKernel : IsoT_FIM::CalKrPc (FlowUnit_ODGW01::CalKrPcDeriv, Stone II)
//                    and IsothermalMethod::CalRock (Rock_Linear)
//   calResKernel   : IsoT_FIM::CalRes, bulk part (accumulation + upwinded
//                    TPFA fluxes), negation; resNormKernel: its relative-
//                    residual maxima
//   assembleKernel : IsoT_FIM::AssembleMatBulksNew into the reference's BSR
// plus list-driven variants for the rows the overlapped Newton iteration
// recomputes.
#include <algorithm>

#include "fim_kernels.hpp"

namespace ocpg {

constexpr double kCONV1f = 5.61458;            // include/OCPConst.hpp:59
constexpr double kCONV2f = 1.12712E-3;         // include/OCPConst.hpp:60
constexpr double kGravityFactor = 0.00694444;  // include/OCPConst.hpp:51

// OCPTable::Eval_All on column 0 (src/OCPTable.cpp:47-86): the interval
// [r, r+1) holding val, flat extrapolation with zero slope outside.
__device__ __forceinline__ void tableEval4(const double (*tab)[4], int nRow, double val, double* out,
                                           double* slope)
{
    int seg = -1;
    for (int r = 0; r + 1 < nRow; ++r) {
        if (val >= tab[r][0] && val < tab[r + 1][0]) {
            seg = r;
            break;
        }
    }
    if (seg < 0) {
        const int r = (val < tab[0][0]) ? 0 : nRow - 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            slope[q] = 0;
            out[q] = tab[r][q];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        slope[q] = (tab[seg + 1][q] - tab[seg][q]) / (tab[seg + 1][0] - tab[seg][0]);
        out[q] = tab[seg][q] + slope[q] * (val - tab[seg][0]);
    }
}

// FlowUnit_ODGW01::CalKrPcDeriv + CalKro_Stone2Der (src/FlowUnit.cpp:235-328),
// Pj = P + Pc (src/OCPFluidMethod.cpp:1189-1199); Rock_Linear::CalPoro
// (src/Rock.cpp:19-27) and rockVp = v * poro (src/OCPFluidMethod.cpp:25-32).
__device__ __forceinline__ void krPcRockCell(const DevBulk& B, const DevSat* __restrict__ sat, const DevRock& rock,
                                             int krpc, int rockOn, int64_t n)
{
    const int64_t ld = B.ld;
    if (krpc) {
        double data[4], cdata[4];
        const double Sg = B.S[1 * ld + n];
        const double Sw = B.S[2 * ld + n];
        tableEval4(sat->swof, sat->swofRows, Sw, data, cdata);
        const double krw = data[1], dKrwdSw = cdata[1];
        const double krow = data[2], dKrowdSw = cdata[2];
        const double Pcwo = -data[3], dPcwdSw = -cdata[3];
        tableEval4(sat->sgof, sat->sgofRows, Sg, data, cdata);
        const double krg = data[1], dKrgdSg = cdata[1];
        const double krog = data[2], dKrogdSg = cdata[2];
        const double Pcgo = data[3], dPcgdSg = cdata[3];
        const double kroMax = sat->kroMax;
        double kro = kroMax * ((krow / kroMax + krw) * (krog / kroMax + krg) - (krw + krg));
        double dKrodSw = kroMax * ((dKrowdSw / kroMax + dKrwdSw) * (krog / kroMax + krg) - (dKrwdSw));
        double dKrodSg = kroMax * ((krow / kroMax + krw) * (dKrogdSg / kroMax + dKrgdSg) - (dKrgdSg));
        if (kro < 0) {
            kro = 0;
            dKrodSg = 0;
            dKrodSw = 0;
        }
        B.kr[0 * ld + n] = kro;
        B.kr[1 * ld + n] = krg;
        B.kr[2 * ld + n] = krw;
        B.Pc[0 * ld + n] = 0;
        B.Pc[1 * ld + n] = Pcgo;
        B.Pc[2 * ld + n] = Pcwo;
        const double dkr[9] = {0, dKrodSg, dKrodSw, 0, dKrgdSg, 0, 0, 0, dKrwdSw};
        const double dpc[9] = {0, 0, 0, 0, dPcgdSg, 0, 0, 0, dPcwdSw};
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            B.dKr_dS[q * ld + n] = dkr[q];
            B.dPcj_dS[q * ld + n] = dpc[q];
        }
        const double P = B.P[n];
        B.Pj[0 * ld + n] = P + 0.0;
        B.Pj[1 * ld + n] = P + Pcgo;
        B.Pj[2 * ld + n] = P + Pcwo;
    }
    if (rockOn) {
        const double dP = (B.P[n] - rock.Pref);
        const double pi = B.poroInit[n];
        const double poro = pi * (1 + (rock.cp1 + rock.cp2 / 2 * dP) * dP);
        B.poro[n] = poro;
        B.poroP[n] = pi * (rock.cp1 + rock.cp2 * dP);
        B.rockVp[n] = B.v[n] * poro;
    }
}

__global__ void __launch_bounds__(256) krPcRockKernel(DevBulk B, const DevSat* __restrict__ sat, DevRock rock,
                                                      int krpc, int rockOn)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= B.nb || !cellOn(B, n)) return;
    krPcRockCell(B, sat, rock, krpc, rockOn, n);
}

hipError_t launchKrPcRock(const DevBulk& B, const DevSat* sat, const DevRock& rock, bool krpc, bool rockOn,
                          hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = (unsigned)((B.nb + 255) / 256);
    hipLaunchKernelGGL(krPcRockKernel, dim3(grid), dim3(256), 0, s, B, sat, rock, krpc ? 1 : 0, rockOn ? 1 : 0);
    return hipGetLastError();
}

// kr/Pc and rock of the overlapped flash's side cells.
__global__ void __launch_bounds__(256) krPcRockListKernel(DevBulk B, const DevSat* __restrict__ sat, DevRock rock,
                                                          CellLists L)
{
    const int64_t total = cellTotal(L);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        krPcRockCell(B, sat, rock, 1, 1, cellAt(L, i));
}

hipError_t launchKrPcRockList(const DevBulk& B, const DevSat* sat, const DevRock& rock, const CellLists& L,
                              hipStream_t s)
{
    hipLaunchKernelGGL(krPcRockListKernel, dim3(64), dim3(256), 0, s, B, sat, rock, L);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rows the overlapped form recomputes once its side streams have joined:
// what CalRes and the assembly computed on the main stream is final for a
// row whose cell and neighbours all finished there, and is recomputed for
// the side cells (bit 2) and their neighbours (bit 1); CalRes also for the
// perforated cells (bit 4: the wells' rates read the side cells too).  A
// row joins a list when its first bit of that list is set, so each row is
// listed once.  mark[] is all zero between calls (fixupClearKernel).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void markRow(uint32_t* mark, int32_t row, uint32_t bit, const FixupRows& R)
{
    const uint32_t old = atomicOr(&mark[row], bit);
    if (old == 0) R.resRows[atomicAdd(&R.counts[0], 1)] = row;
    if ((bit & 3u) && (old & 3u) == 0) R.asmRows[atomicAdd(&R.counts[1], 1)] = row;
}

__global__ void __launch_bounds__(256) fixupMarkKernel(FimDevice F, CellLists L, FixupRows R)
{
    const int64_t total = cellTotal(L);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = cellAt(L, i);
        markRow(R.mark, t, 2u, R);
        for (int k = F.IA[t] + 1; k < F.bulkEnd[t]; ++k) markRow(R.mark, F.JA[k], 1u, R);
    }
}

__global__ void __launch_bounds__(256) fixupMarkPerfKernel(WellDevice W, FixupRows R)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < W.npcell) markRow(R.mark, W.pcell[i], 4u, R);
}

__global__ void __launch_bounds__(256) fixupClearKernel(FixupRows R)
{
    const int64_t total = R.counts[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        R.mark[R.resRows[i]] = 0;
}

__global__ void fixupResetKernel(FixupRows R)
{
    if (threadIdx.x < 2) R.counts[threadIdx.x] = 0;
}

hipError_t launchFixupRows(const FimDevice& F, const WellDevice& W, const CellLists& L, const FixupRows& R,
                           hipStream_t s)
{
    hipLaunchKernelGGL(fixupResetKernel, dim3(1), dim3(64), 0, s, R);
    hipLaunchKernelGGL(fixupMarkKernel, dim3(64), dim3(256), 0, s, F, L, R);
    if (W.npcell > 0) hipLaunchKernelGGL(fixupMarkPerfKernel, dim3((W.npcell + 255) / 256), dim3(256), 0, s, W, R);
    hipLaunchKernelGGL(fixupClearKernel, dim3(64), dim3(256), 0, s, R);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CalRes (src/OCPFluidMethod.cpp:1201-1389), one lane per cell.  A cell meets
// its connections in the order of its BSR row (lower neighbours ascending,
// then upper ones), which is the order the reference's connection loop adds
// them to that cell, so each cell's residual is summed in the reference's
// order.  The connection's upwind data is written by the bId side only.
// ---------------------------------------------------------------------------
constexpr int kResBlock = 256;

template <int NC>
__device__ __forceinline__ void calResCell(const DevBulk& B, const FimDevice& F, const WellDevice& W, double dt,
                                           int64_t n)
{
    constexpr int nc = NC + 1;
    constexpr int len = nc + 1;
    const int64_t ld = B.ld;
    {
        double r[len];
        r[0] = B.rockVp[n] - B.vf[n];
#pragma unroll
        for (int i = 0; i < nc; ++i) r[1 + i] = B.Ni[i * ld + n] - B.lNi[i * ld + n];
        const int k0 = F.IA[n] + 1, k1 = F.bulkEnd[n];
        for (int k = k0; k < k1; ++k) {
            const int64_t m = F.JA[k];
            const int64_t c = F.slotConn[k];
            const bool isB = n < m;
            const int64_t bId = isB ? n : m;
            const int64_t eId = isB ? m : n;
            const double Akd = kCONV1f * kCONV2f * F.area[c];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const bool exB = B.phaseExist[j * ld + bId];
                const bool exE = B.phaseExist[j * ld + eId];
                double rho;
                if (exB && exE)
                    rho = (B.rho[j * ld + bId] + B.rho[j * ld + eId]) / 2;
                else if (exB)
                    rho = B.rho[j * ld + bId];
                else if (exE)
                    rho = B.rho[j * ld + eId];
                else {
                    if (isB) {
                        F.upblock[c * 3 + j] = (int32_t)bId;
                        F.upRho[c * 3 + j] = 0;
                    }
                    continue;
                }
                const double dP = (B.Pj[j * ld + bId] - kGravityFactor * rho * B.depth[bId]) -
                                  (B.Pj[j * ld + eId] - kGravityFactor * rho * B.depth[eId]);
                const int64_t uId = (dP < 0) ? eId : bId;
                const bool exU = B.phaseExist[j * ld + uId];
                if (isB) {
                    F.upRho[c * 3 + j] = rho;
                    F.upblock[c * 3 + j] = (int32_t)uId;
                    F.upVel[c * 3 + j] =
                        exU ? Akd * B.kr[j * ld + uId] / B.mu[j * ld + uId] * dP : 0.0;
                }
                if (!exU) continue;
                const double tmp = dt * Akd * B.xi[j * ld + uId] * B.kr[j * ld + uId] / B.mu[j * ld + uId] * dP;
#pragma unroll
                for (int i = 0; i < nc; ++i) {
                    const double dNi = tmp * B.xij[(j * nc + i) * ld + uId];
                    if (isB)
                        r[1 + i] += dNi;
                    else
                        r[1 + i] -= dNi;
                }
            }
        }
        // well sources, wells then perforations in order (:1284-1293)
        if (W.nw > 0) {
            for (int q = W.cellPerfPtr[n]; q < W.cellPerfPtr[n + 1]; ++q) {
                const double* qi = W.perfQi + (int64_t)W.cellPerfIdx[q] * nc;
#pragma unroll
                for (int i = 0; i < nc; ++i) r[1 + i] += qi[i] * dt;
            }
        }
#pragma unroll
        for (int e = 0; e < len; ++e) F.res[n * len + e] = -r[e];
    }
}

template <int NC>
__global__ void __launch_bounds__(kResBlock) calResKernel(DevBulk B, FimDevice F, WellDevice W, double dt)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < B.nb && cellOn(B, n)) calResCell<NC>(B, F, W, dt, n);
}

// CalRes of the rows in rows[0, *count) (the overlapped form's recomputation).
template <int NC>
__global__ void __launch_bounds__(kResBlock) calResListKernel(DevBulk B, FimDevice F, WellDevice W, double dt,
                                                              const int32_t* __restrict__ rows,
                                                              const int32_t* __restrict__ count)
{
    const int64_t total = *count;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        calResCell<NC>(B, F, W, dt, rows[i]);
}

// CalRes's relative-residual maxima (src/OCPFluidMethod.cpp:1362-1387) from
// the stored right-hand side (the negated residual: |-r / v| = |r / v|).
template <int NC>
__global__ void __launch_bounds__(kResBlock) resNormKernel(DevBulk B, FimDevice F)
{
    constexpr int len = NC + 2;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double vmax = -1.0, nmax = -1.0;
    int64_t vidx = 0, nidx = 0;
    if (n < B.nb && cellOn(B, n)) {
        const double rvp = B.rockVp[n], Nt = B.Nt[n];
#pragma unroll
        for (int e = 0; e < len; ++e) {
            const double t = fabs(F.res[n * len + e] / rvp);
            if (vmax < t) {
                vmax = t;
                vidx = n;
            }
        }
#pragma unroll
        for (int e = 1; e < len; ++e) {
            const double t = fabs(F.res[n * len + e] / Nt);
            if (nmax < t) {
                nmax = t;
                nidx = n;
            }
        }
    }
    // block maxima, ties to the lowest cell (the reference scans cells upward
    // and keeps the first strict maximum)
    __shared__ double sv[kResBlock], sn[kResBlock];
    __shared__ int64_t iv[kResBlock], in_[kResBlock];
    sv[threadIdx.x] = vmax;
    sn[threadIdx.x] = nmax;
    iv[threadIdx.x] = vidx;
    in_[threadIdx.x] = nidx;
    __syncthreads();
    for (int s = kResBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int o = threadIdx.x + s;
            if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && iv[o] < iv[threadIdx.x])) {
                sv[threadIdx.x] = sv[o];
                iv[threadIdx.x] = iv[o];
            }
            if (sn[o] > sn[threadIdx.x] || (sn[o] == sn[threadIdx.x] && in_[o] < in_[threadIdx.x])) {
                sn[threadIdx.x] = sn[o];
                in_[threadIdx.x] = in_[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        F.normPart[2 * blockIdx.x + 0] = sv[0];
        F.normPart[2 * blockIdx.x + 1] = sn[0];
        F.normIdx[2 * blockIdx.x + 0] = iv[0];
        F.normIdx[2 * blockIdx.x + 1] = in_[0];
    }
}

// ---------------------------------------------------------------------------
// AssembleMatBulksNew (src/OCPFluidMethod.cpp:1569-1839), one wavefront per
// BSR row.  For each connection the reference forms two blocks, one per side
// (B: dFdXpB + dFdXsB * dSec[bId], E: dFdXpE + dFdXsE * dSec[eId]), adds each
// to its own cell's diagonal and writes it, signed, as the other cell's
// off-diagonal.  Here the row of cell n forms only its own side's block for
// each of its connections: it accumulates the diagonal in registers (in the
// reference's connection order) and writes the off-diagonal into the
// neighbour's row at the slot the reference would have filled.  Every block
// is computed once, nothing is accumulated across waves, no atomics.
//
// Arithmetic.  Every entry of the own side's dFdXp / dFdXs is a sum over
// the phases j of x_ij(upwind of j) times a coefficient that does not depend
// on the component i (the flux derivatives are transIJ = x_ij xi_j transJ_j
// times per-phase factors), plus, in the x columns, a term on the
// diagonal kk == i.  So per connection the wavefront computes
//   C[j][col]  the coefficient of column col (pressure, the own cell's
//              saturations, its x_kj) for phase j          (lanes (j, col))
//   G[j][cc] = sum_col C[j][col] dSec_own[col][cc] (+ the pressure part)
//                                                          (lanes (j, cc))
//   block[i][cc] = dt (sum_j x_ij G[j][cc] + sum_j d_j dSec_own[x(j, i)][cc])
//                                                          (lanes (i, cc))
// instead of the nc x (1 + ncol_own) derivative entries and their product
// with dSec one by one; the sums are the reference's, reassociated (the
// blocks agree with the reference's to rounding).  kr/mu, xi/mu, muP/mu and
// mux/mu are formed once per cell by the record packing, so the assembly
// divides nothing.
//
// Memory plan: a neighbour contributes only its pressures, depth, phase
// flags and, where it is upwind, its xi, kr/mu and x_ij: a record of
// AsmRec::LEN doubles, read with one load per lane and prefetched into
// registers while the previous connection is computed.  The row's own cell
// also contributes its derivatives (AsmOwn, read once per row).  Rows are
// handed to workgroups so that each XCD walks a contiguous range of cells,
// which keeps the neighbour records (cells n +- 1, n +- nx, ...) in that
// XCD's L2.
// ---------------------------------------------------------------------------
constexpr int kAsmBlock = 64;
constexpr int kXcds = 8;

template <int NC>
struct AsmDims {
    static constexpr int nc = NC + 1;
    static constexpr int ncol = nc + 1;
    static constexpr int bsize = ncol * ncol;
    static constexpr int maxRows = 3 + 2 * NC;          // compact secondary rows
    static constexpr int perLane = (bsize + kAsmBlock - 1) / kAsmBlock;
    // columns per phase of the coefficient table: the own cell's secondary
    // columns, then the pressure coefficient at CS - 1
    static constexpr int CS = maxRows + 1 <= 16 ? 16 : 32;
    static constexpr int cPasses = (3 * CS + kAsmBlock - 1) / kAsmBlock;
    static_assert(3 * ncol <= kAsmBlock, "G needs a lane per (phase, column)");
};

// The record a row reads for each neighbour (and for itself): phase j,
// component i.
template <int NC>
struct AsmRec {
    static constexpr int nc = NC + 1;
    static constexpr int PJ = 0, XI = 3, LAM = 6;   // LAM: kr / mu
    static constexpr int XIJ = 9;                   // xij[j*nc+i]
    static constexpr int DEPTH = XIJ + 3 * nc, EX = DEPTH + 1, LEN = EX + 3;
    static constexpr int perLane = (LEN + kAsmBlock - 1) / kAsmBlock;
};

// The own-row data of a row's assembly: compact dSec_dPri rows, dPcj_dS,
// the accumulation terms, the secondary-column bookkeeping (counts stored
// as doubles, exactly), then the own cell's derivatives (DER block).
template <int NC>
struct AsmOwn {
    using D = AsmDims<NC>;
    static constexpr int nc = NC + 1;
    static constexpr int DSEC = 0;                    // [r * ncol + c], r < rows
    static constexpr int DPC = D::maxRows * D::ncol;  // dPcj_dS[9]
    static constexpr int VFI = DPC + 9;               // vfi[nc]
    static constexpr int VFP = VFI + nc, V = VFP + 1, POROP = V + 1;
    static constexpr int SDEX = POROP + 1;            // pSderExist[3]
    static constexpr int PVN = SDEX + 3;              // pVnumCom[3]
    static constexpr int ROWS = PVN + 3;              // bRowSizedSdP
    static constexpr int DER = ROWS + 1;
    static constexpr int XIP = DER, RHOP = XIP + 3, MUPM = RHOP + 3;  // MUPM: muP / mu
    static constexpr int XIM = MUPM + 3;                              // xi / mu
    static constexpr int DKR = XIM + 3;                               // dKr_dS[j*3+k]
    static constexpr int RHOX = DKR + 9, XIX = RHOX + 3 * nc;         // [j*nc+k]
    static constexpr int MUXM = XIX + 3 * nc;                         // mux / mu
    static constexpr int LEN = MUXM + 3 * nc;
    static constexpr int perLane = (LEN + kAsmBlock - 1) / kAsmBlock;
};

// Both records are packed per cell (AoS, rows padded to whole 128-byte
// lines), so a wavefront reads a neighbour's record as one contiguous span
// instead of one 8-byte value from each of ~40 arrays.
template <int NC>
constexpr int recStride() { return (AsmRec<NC>::LEN + 15) / 16 * 16; }
template <int NC>
constexpr int ownStride() { return (AsmOwn<NC>::LEN + 15) / 16 * 16; }

__device__ __forceinline__ double ratioOrZero(double a, double b) { return b != 0.0 ? a / b : 0.0; }

// Record entry q of cell n, straight from the SoA state.
template <int NC>
__device__ __forceinline__ double recValue(const DevBulk& B, int q, int64_t n)
{
    using R = AsmRec<NC>;
    const int64_t ld = B.ld;
    if (q < R::XI) return B.Pj[(int64_t)(q - R::PJ) * ld + n];
    if (q < R::LAM) return B.xi[(int64_t)(q - R::XI) * ld + n];
    if (q < R::XIJ) {
        const int64_t o = (int64_t)(q - R::LAM) * ld + n;
        return ratioOrZero(B.kr[o], B.mu[o]);
    }
    if (q < R::DEPTH) return B.xij[(int64_t)(q - R::XIJ) * ld + n];
    if (q == R::DEPTH) return B.depth[n];
    if (q < R::LEN) return B.phaseExist[(int64_t)(q - R::EX) * ld + n] ? 1.0 : 0.0;
    return 0.0;
}

// Own-row entry q of cell n (q >= DSEC + rows * ncol inside the dSec part is
// never read and not stored).
template <int NC>
__device__ __forceinline__ double ownValue(const DevBulk& B, int q, int64_t n)
{
    using O = AsmOwn<NC>;
    constexpr int nc = O::nc;
    const int64_t ld = B.ld;
    if (q < O::DPC) return B.dSec[(int64_t)q * ld + n];
    if (q < O::VFI) return B.dPcj_dS[(int64_t)(q - O::DPC) * ld + n];
    if (q < O::VFP) return B.vfi[(int64_t)(q - O::VFI) * ld + n];
    if (q == O::VFP) return B.vfP[n];
    if (q == O::V) return B.v[n];
    if (q == O::POROP) return B.poroP[n];
    if (q < O::PVN) return B.pSderExist[(int64_t)(q - O::SDEX) * ld + n] ? 1.0 : 0.0;
    if (q < O::ROWS) return (double)B.pVnumCom[(int64_t)(q - O::PVN) * ld + n];
    if (q == O::ROWS) return (double)B.bRowSizedSdP[n];
    if (q < O::RHOP) return B.xiP[(int64_t)(q - O::XIP) * ld + n];
    if (q < O::MUPM) return B.rhoP[(int64_t)(q - O::RHOP) * ld + n];
    if (q < O::XIM) {
        const int64_t o = (int64_t)(q - O::MUPM) * ld + n;
        return ratioOrZero(B.muP[o], B.mu[o]);
    }
    if (q < O::DKR) {
        const int64_t o = (int64_t)(q - O::XIM) * ld + n;
        return ratioOrZero(B.xi[o], B.mu[o]);
    }
    if (q < O::RHOX) return B.dKr_dS[(int64_t)(q - O::DKR) * ld + n];
    if (q < O::XIX) return B.rhox[(int64_t)(q - O::RHOX) * ld + n];
    if (q < O::MUXM) return B.xix[(int64_t)(q - O::XIX) * ld + n];
    if (q < O::LEN) {
        const int jk = q - O::MUXM;
        return ratioOrZero(B.mux[(int64_t)jk * ld + n], B.mu[(int64_t)(jk / nc) * ld + n]);
    }
    return 0.0;
}

// Pack both records of packTile<NC>() consecutive cells: read field by field
// (each field's kPackTile values are contiguous), transpose in LDS, form the
// ratios, write the cells' records as contiguous spans.  Three passes over
// one LDS tile: the neighbour record (raw kr in the LAM slots, mu after the
// record), the own record's first part, its derivative block (raw muP, xi,
// mux in the ratio slots, mu after the block).
// Cells per block: 32 for the 6-component records; 16 for larger ones
// (their tile in LDS and the loads in flight per thread grow with NC).
template <int NC>
constexpr int packTile() { return NC <= 6 ? 32 : 16; }
constexpr int kPackBlock = 256;
// every tile ends inside the padded leading dimension (a multiple of 64),
// and the flag and count passes take one pass over the block
static_assert(64 % packTile<6>() == 0 && 64 % packTile<10>() == 0, "a tile must not cross the padding of ld");
static_assert(7 * packTile<6>() <= kPackBlock && 7 * packTile<10>() <= kPackBlock, "flag/count pass exceeds the block");

// Loads Q raw double fields (src[q] already advanced to the tile's first
// cell) of the tile's cells into t[c * L + q]: all loads first, then the LDS
// stores (generic loads may alias LDS: interleaved, each load would be
// waited for before the next store).  Entries past the tile read the
// padding up to ld and are never written back.
template <int Q, int TILE, int L>
__device__ __forceinline__ void packLoad(const double* const* src, double* t)
{
    constexpr int K = (Q * TILE + kPackBlock - 1) / kPackBlock;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = threadIdx.x + k * kPackBlock;
        v[k] = i < Q * TILE ? src[i / TILE][i % TILE] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = threadIdx.x + k * kPackBlock;
        if (i < Q * TILE) t[(i % TILE) * L + i / TILE] = v[k];
    }
}

template <int NC>
__device__ __forceinline__ void packTileBody(const DevBulk& B, const FimDevice& F, int64_t tileIdx)
{
    using R = AsmRec<NC>;
    using O = AsmOwn<NC>;
    constexpr int nc = NC + 1;
    constexpr int ncol = AsmDims<NC>::ncol;
    constexpr int RS = recStride<NC>(), OS = ownStride<NC>();
    constexpr int kPackTile = packTile<NC>();
    // raw fields per pass: the record's doubles + mu; the own part before
    // the flags; the derivative block + mu
    constexpr int QR = R::EX + 3, QO = O::SDEX, QD = O::LEN - O::DER + 3;
    constexpr int QMAX = QR > QO ? (QR > QD ? QR : QD) : (QO > QD ? QO : QD);
    constexpr int L = (QMAX > O::DER ? QMAX : O::DER) + 1;  // odd LDS row: no bank conflicts
    __shared__ double t[kPackTile * L];
    __shared__ int rows[kPackTile];
    __shared__ int on[kPackTile];
    __shared__ int anyOn;
    __shared__ const double* src[QMAX];
    const int64_t n0 = tileIdx * kPackTile;
    const int tile = (int)min<int64_t>(kPackTile, B.nb - n0);
    const int64_t ld = B.ld;
    if (threadIdx.x == 0) anyOn = 0;
    __syncthreads();
    if (threadIdx.x < kPackTile) {
        on[threadIdx.x] = threadIdx.x < tile && cellOn(B, n0 + threadIdx.x);
        if (on[threadIdx.x]) atomicOr(&anyOn, 1);
    }
    __syncthreads();
    if (!anyOn) return;  // the whole block leaves: a pool's idle realisations
    if (threadIdx.x < kPackTile)
        rows[threadIdx.x] = threadIdx.x < tile ? min(max(B.bRowSizedSdP[n0 + threadIdx.x], 0),
                                                     AsmDims<NC>::maxRows) : 0;
    // ---- the neighbour record ----
    for (int q = threadIdx.x; q < QR; q += kPackBlock) {
        const double* p;
        if (q < R::XI) p = B.Pj + (int64_t)(q - R::PJ) * ld;
        else if (q < R::LAM) p = B.xi + (int64_t)(q - R::XI) * ld;
        else if (q < R::XIJ) p = B.kr + (int64_t)(q - R::LAM) * ld;
        else if (q < R::DEPTH) p = B.xij + (int64_t)(q - R::XIJ) * ld;
        else if (q == R::DEPTH) p = B.depth;
        else p = B.mu + (int64_t)(q - R::EX) * ld;  // raw mu at EX..EX+2, the flags go in after
        src[q] = p + n0;
    }
    __syncthreads();
    packLoad<QR, kPackTile, L>(src, t);
    __syncthreads();
    // kr -> kr / mu, then the flags over the raw mu
    if (threadIdx.x < 3 * kPackTile) {
        const int j = threadIdx.x / kPackTile, c = threadIdx.x % kPackTile;
        double* row = t + c * L;
        row[R::LAM + j] = ratioOrZero(row[R::LAM + j], row[R::EX + j]);
    }
    __syncthreads();
    if (threadIdx.x < 3 * kPackTile) {
        const int q = R::EX + threadIdx.x / kPackTile, c = threadIdx.x % kPackTile;
        t[c * L + q] = B.phaseExist[(int64_t)(q - R::EX) * ld + n0 + c] ? 1.0 : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (RS * kPackTile + kPackBlock - 1) / kPackBlock; ++k) {
        const int i = threadIdx.x + k * kPackBlock;
        const int c = i / RS, q = i % RS;
        if (i < RS * kPackTile && c < tile && q < R::LEN && on[c]) F.rec[(n0 + c) * RS + q] = t[c * L + q];
    }
    __syncthreads();
    // ---- the own record, first part: the double fields (unused dSec rows
    // are read and not written back), then the seven counts ----
    for (int q = threadIdx.x; q < QO; q += kPackBlock) {
        const double* p;
        if (q < O::DPC) p = B.dSec + (int64_t)q * ld;
        else if (q < O::VFI) p = B.dPcj_dS + (int64_t)(q - O::DPC) * ld;
        else if (q < O::VFP) p = B.vfi + (int64_t)(q - O::VFI) * ld;
        else if (q == O::VFP) p = B.vfP;
        else if (q == O::V) p = B.v;
        else p = B.poroP;
        src[q] = p + n0;
    }
    __syncthreads();
    packLoad<QO, kPackTile, L>(src, t);
    if (threadIdx.x < (O::DER - O::SDEX) * kPackTile) {
        const int q = O::SDEX + threadIdx.x / kPackTile, c = threadIdx.x % kPackTile;
        if (c < tile) t[c * L + q] = ownValue<NC>(B, q, n0 + c);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (O::DER * kPackTile + kPackBlock - 1) / kPackBlock; ++k) {
        const int i = threadIdx.x + k * kPackBlock;
        const int c = i / O::DER, q = i % O::DER;
        if (i < O::DER * kPackTile && c < tile && (q >= O::DPC || q < rows[c] * ncol) && on[c])
            F.own[(n0 + c) * OS + q] = t[c * L + q];
    }
    __syncthreads();
    // ---- the own record's derivative block (tile index q - DER) ----
    for (int q = threadIdx.x; q < QD; q += kPackBlock) {
        const int o = O::DER + q;
        const double* p;
        if (o < O::RHOP) p = B.xiP + (int64_t)(o - O::XIP) * ld;
        else if (o < O::MUPM) p = B.rhoP + (int64_t)(o - O::RHOP) * ld;
        else if (o < O::XIM) p = B.muP + (int64_t)(o - O::MUPM) * ld;
        else if (o < O::DKR) p = B.xi + (int64_t)(o - O::XIM) * ld;
        else if (o < O::RHOX) p = B.dKr_dS + (int64_t)(o - O::DKR) * ld;
        else if (o < O::XIX) p = B.rhox + (int64_t)(o - O::RHOX) * ld;
        else if (o < O::MUXM) p = B.xix + (int64_t)(o - O::XIX) * ld;
        else if (o < O::LEN) p = B.mux + (int64_t)(o - O::MUXM) * ld;
        else p = B.mu + (int64_t)(o - O::LEN) * ld;  // raw mu after the block
        src[q] = p + n0;
    }
    __syncthreads();
    packLoad<QD, kPackTile, L>(src, t);
    __syncthreads();
    {
        // the ratios: muP / mu, xi / mu (3 each), mux / mu (3 nc)
        constexpr int NR = 6 + 3 * nc;
        constexpr int MU = O::LEN - O::DER;
        for (int i = threadIdx.x; i < NR * kPackTile; i += kPackBlock) {
            const int r = i / kPackTile, c = i % kPackTile;
            double* row = t + c * L;
            int q, j;
            if (r < 6) {
                q = (r < 3 ? O::MUPM : O::XIM) - O::DER + r % 3;
                j = r % 3;
            } else {
                q = O::MUXM - O::DER + (r - 6);
                j = (r - 6) / nc;
            }
            row[q] = ratioOrZero(row[q], row[MU + j]);
        }
    }
    __syncthreads();
    constexpr int ND = O::LEN - O::DER;
#pragma unroll
    for (int k = 0; k < (ND * kPackTile + kPackBlock - 1) / kPackBlock; ++k) {
        const int i = threadIdx.x + k * kPackBlock;
        const int c = i / ND, q = i % ND;
        if (i < ND * kPackTile && c < tile && on[c]) F.own[(n0 + c) * OS + O::DER + q] = t[c * L + q];
    }
}

template <int NC>
__global__ void __launch_bounds__(kPackBlock) packRecordsKernel(DevBulk B, FimDevice F)
{
    packTileBody<NC>(B, F, blockIdx.x);
}

// CalRes and the record packing in one launch (they read the same state and
// neither needs the other): each group of 1 + tiles blocks is one CalRes
// block and the packing of the same cells, so the two kinds of work share
// the chip instead of running one after the other.
template <int NC>
constexpr int packTilesPerResBlock() { return kResBlock / packTile<NC>(); }

template <int NC>
__global__ void __launch_bounds__(kResBlock) calResPackKernel(DevBulk B, FimDevice F, WellDevice W, double dt)
{
    constexpr int PT = packTilesPerResBlock<NC>();
    const int64_t group = blockIdx.x / (1 + PT);
    const int r = (int)(blockIdx.x % (1 + PT));
    if (r == 0) {
        const int64_t n = group * kResBlock + threadIdx.x;
        if (n < B.nb && cellOn(B, n)) calResCell<NC>(B, F, W, dt, n);
    } else {
        const int64_t tileIdx = group * PT + (r - 1);
        if (tileIdx * packTile<NC>() < B.nb) packTileBody<NC>(B, F, tileIdx);
    }
}

// Both records of the overlapped flash's side cells: one thread per entry.
template <int NC>
__global__ void __launch_bounds__(256) packRecordsListKernel(DevBulk B, FimDevice F, CellLists L)
{
    using R = AsmRec<NC>;
    using O = AsmOwn<NC>;
    constexpr int ncol = AsmDims<NC>::ncol;
    constexpr int RS = recStride<NC>(), OS = ownStride<NC>();
    constexpr int W = R::LEN + O::LEN;
    const int64_t total = cellTotal(L) * W;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = cellAt(L, i / W);
        const int q = (int)(i % W);
        if (q < R::LEN) {
            F.rec[n * RS + q] = recValue<NC>(B, q, n);
        } else {
            const int qo = q - R::LEN;
            const int rows = min(max(B.bRowSizedSdP[n], 0), AsmDims<NC>::maxRows);
            if (qo >= O::DPC || qo < rows * ncol) F.own[n * OS + qo] = ownValue<NC>(B, qo, n);
        }
    }
}

// A value that is the same on every lane, moved to scalar registers.
__device__ __forceinline__ double uniformD(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v != 0.0 for a value that is the same on every lane, on its bits in
// scalar registers (-0.0 is zero, a NaN is not, as for the comparison).
__device__ __forceinline__ bool uniformNonzero(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return (lo | (hi & 0x7fffffffu)) != 0;
}

template <int NC>
struct AsmShared {
    using D = AsmDims<NC>;
    using R = AsmRec<NC>;
    using O = AsmOwn<NC>;
    double sOwn[O::LEN];                // own record
    double sRec[2][R::LEN];             // [0] own cell, [1] current neighbour
    double sC[3][D::CS];                // per phase: column coefficients, [CS - 1] the pressure one
    double sD[3];                       // per phase: the x columns' diagonal term
    double sG[3][D::ncol + 1];          // per phase: C * dSec_own ([ncol]: C * res_n, IsoT_FIMn)
};

// Row n of the bulk assembly, by one wavefront.
//
// FN: IsoT_FIMn::AssembleMatBulksNew (src/OCPFluidMethod.cpp:3198-3501).
// The secondary columns are the phase moles n_ij (x_ij = n_ij / nj): where
// the row's cell is upwind, every n_kj column of phase j gains
// -x_ij xi transJ dP / nj and the diagonal term becomes xi transJ dP / nj,
// both per-phase factors of the same C / d form.  The row also forms its
// cell's right-hand-side terms (only where the cell has two hydrocarbon
// phases): per connection -dFdXs res_n, i.e. -(sum_j x_ij (C[j] . res_n) +
// sum_j d_j res_n[x(j, i)]), summed in the connection loop's order after
// -resPc on the volume row, into F.rhsN.
template <int NC, bool FN = false>
__device__ __forceinline__ void assembleRow(const DevBulk& B, const FimDevice& F, double dt, int64_t n,
                                            AsmShared<NC>& sh)
{
    using D = AsmDims<NC>;
    using R = AsmRec<NC>;
    using O = AsmOwn<NC>;
    constexpr int nc = D::nc;
    constexpr int ncol = D::ncol;
    constexpr int bsize = D::bsize;
    constexpr int CS = D::CS;
    constexpr int RS = recStride<NC>(), OS = ownStride<NC>();
    const int lane = threadIdx.x;
    double* sOwn = sh.sOwn;
    double (*sRec)[R::LEN] = sh.sRec;
    // a previous row of this wavefront may still read the shared arrays
    __syncthreads();

    const double* __restrict__ own = F.own + n * OS;
    const int rowsOwn = min(max((int)own[O::ROWS], 0), D::maxRows);
#pragma unroll
    for (int t = 0; t < O::perLane; ++t) {
        const int q = lane + t * kAsmBlock;
        if (q < O::LEN && (q >= O::DPC || q < rowsOwn * ncol)) sOwn[q] = own[q];
    }
#pragma unroll
    for (int t = 0; t < R::perLane; ++t) {
        const int q = lane + t * kAsmBlock;
        if (q < R::LEN) sRec[0][q] = F.rec[n * RS + q];
    }

    // secondary columns of this row's cell: existing saturations (their
    // phases in sph), then the x_ij of each phase with derivatives (phase j's
    // from column xb[j])
    bool sdO[3];
    int pvO[3];
    int nsat = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        sdO[j] = own[O::SDEX + j] != 0.0;
        pvO[j] = min(max((int)own[O::PVN + j], 0), NC);
        nsat += sdO[j] ? 1 : 0;
    }
    int sph[3] = {0, 0, 0};
    {
        int s = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (sdO[j]) sph[s++] = j;
    }
    const int ncolO = min(nsat + pvO[0] + pvO[1] + pvO[2], D::maxRows);
    int xb[3];
    xb[0] = nsat;
    xb[1] = xb[0] + pvO[0];
    xb[2] = xb[1] + pvO[1];

    // this lane's column of dSec_own for G (lanes (j, cc)); with FN, lanes
    // (j, ncol) take res_n instead
    constexpr int GW = FN ? ncol + 1 : ncol;
    const int gcc = lane % GW;
    double dS[D::maxRows];
    if constexpr (FN) {
        const int64_t ld = B.ld;
#pragma unroll
        for (int k = 0; k < D::maxRows; ++k)
            dS[k] = k < rowsOwn ? (gcc < ncol ? own[O::DSEC + k * ncol + gcc] : B.resN[(int64_t)k * ld + n]) : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < D::maxRows; ++k) dS[k] = k < rowsOwn ? own[O::DSEC + k * ncol + gcc] : 0.0;
    }
    // IsoT_FIMn: the own cell's phase moles and whether it has two
    // hydrocarbon phases (then its right-hand-side terms are formed)
    double njO[3] = {1.0, 1.0, 1.0};
    bool twoPh = false;
    double rhsAcc = 0.0;
    // the right-hand-side lanes' n_ij corrections of the own cell
    // (res_n rows x(j, i), i = lane - 1), read once for all connections
    double resX[3] = {0.0, 0.0, 0.0};
    if constexpr (FN) {
        const int64_t ld = B.ld;
#pragma unroll
        for (int j = 0; j < 3; ++j) njO[j] = uniformD(B.nj[(int64_t)j * ld + n]);
        twoPh = __builtin_amdgcn_readfirstlane(B.phaseNum[n]) > 2;
        if (lane == 0) rhsAcc = 0.0 + -B.resPc[n];
        if (twoPh && lane >= 1 && lane < ncol) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (lane - 1 < pvO[j]) resX[j] = B.resN[(int64_t)(xb[j] + lane - 1) * ld + n];
        }
    }

    // diagonal starts as the accumulation block (src/OCPFluidMethod.cpp:1586-1597)
    double acc[D::perLane];
#pragma unroll
    for (int t = 0; t < D::perLane; ++t) {
        const int e = lane + t * kAsmBlock;
        const int r = e / ncol, cc = e % ncol;
        double v = 0;
        if (e < bsize) {
            if (r == 0 && cc == 0)
                v = own[O::V] * own[O::POROP] - own[O::VFP];
            else if (r == 0)
                v = -own[O::VFI + cc - 1];
            else if (r == cc)
                v = 1;
        }
        acc[t] = v;
    }

    const int k0 = F.IA[n] + 1, k1 = F.bulkEnd[n];
    // Slot indices and connection values are the same on every lane: they
    // live in scalar registers, and the branches on them are scalar.
    double pre[R::perLane];
    int64_t mNext = 0, cNext = 0;
    if (k0 < k1) {
        mNext = __builtin_amdgcn_readfirstlane(F.JA[k0]);
        cNext = __builtin_amdgcn_readfirstlane(F.slotConn[k0]);
#pragma unroll
        for (int t = 0; t < R::perLane; ++t) {
            const int q = lane + t * kAsmBlock;
            pre[t] = q < R::LEN ? F.rec[mNext * RS + q] : 0.0;
        }
    }
    for (int k = k0; k < k1; ++k) {
        const int64_t m = mNext, c = cNext;
        // One round trip for this connection's values and the next slot's
        // indices, issued before the next neighbour's record is prefetched:
        // loads complete in order, so waiting for these leaves that
        // prefetch in flight while this connection is computed.
        const double areaV = F.area[c];
        int32_t upV[3];
        double upRhoV[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            upV[j] = F.upblock[c * 3 + j];
            upRhoV[j] = F.upRho[c * 3 + j];
        }
        const int32_t mirV = F.slotMirror[k];
        const bool more = k + 1 < k1;
        int32_t mNv = 0, cNv = 0;
        if (more) {
            mNv = F.JA[k + 1];
            cNv = F.slotConn[k + 1];
        }
#pragma unroll
        for (int t = 0; t < R::perLane; ++t) {
            const int q = lane + t * kAsmBlock;
            if (q < R::LEN) sRec[1][q] = pre[t];
        }
        __syncthreads();
        mNext = __builtin_amdgcn_readfirstlane(mNv);
        cNext = __builtin_amdgcn_readfirstlane(cNv);
        if (more) {
#pragma unroll
            for (int t = 0; t < R::perLane; ++t) {
                const int q = lane + t * kAsmBlock;
                pre[t] = q < R::LEN ? F.rec[mNext * RS + q] : 0.0;
            }
        }
        const bool ownB = n < m;  // this row is the connection's bId side
        const double* rB = ownB ? sRec[0] : sRec[1];
        const double* rE = ownB ? sRec[1] : sRec[0];
        const double Akd = kCONV1f * kCONV2f * uniformD(areaV);
        const double dGamma = kGravityFactor * (uniformD(rB[R::DEPTH]) - uniformD(rE[R::DEPTH]));
        // per-phase flags as bit masks (scalar): the own cell has the phase,
        // the other cell has it, the own cell is upwind, the upwind cell has
        // it (the phase flows)
        unsigned ownExM = 0, othExM = 0, upOwnM = 0, liveM = 0;
        double upRho[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const bool exO = uniformNonzero(sRec[0][R::EX + j]);
            const bool exN = uniformNonzero(sRec[1][R::EX + j]);
            const bool upO = __builtin_amdgcn_readfirstlane(upV[j]) == (int32_t)n;
            ownExM |= exO ? 1u << j : 0u;
            othExM |= exN ? 1u << j : 0u;
            upOwnM |= upO ? 1u << j : 0u;
            liveM |= (upO ? exO : exN) ? 1u << j : 0u;
            upRho[j] = uniformD(upRhoV[j]);
        }
        const int64_t mirror = __builtin_amdgcn_readfirstlane(mirV);
        const double sgn = ownB ? 1.0 : -1.0;

        // ---- coefficients C[j][col], the pressure one at CS - 1, d_j ----
#pragma unroll
        for (int t = 0; t < D::cPasses; ++t) {
            const int q = lane + t * kAsmBlock;
            const int j = q / CS, col = q % CS;
            if (j < 3) {
                double cv = 0.0, dv = 0.0;
                if ((liveM >> j) & 1u) {
                    const bool upO = (upOwnM >> j) & 1u;
                    const bool othEx = (othExM >> j) & 1u;
                    const bool ownEx = (ownExM >> j) & 1u;
                    const double* rU = upO ? sRec[0] : sRec[1];
                    const double dP = rB[R::PJ + j] - rE[R::PJ + j] - (j == 0 ? upRho[0] : j == 1 ? upRho[1] : upRho[2]) * dGamma;
                    const double transJ = Akd * rU[R::LAM + j];
                    const double A = rU[R::XI + j] * transJ;
                    // IsoT_FIMn: (xi transJ dP) / n_j of the upwind own cell,
                    // shared by the pressure lane's d_j and the x columns
                    double qN = 0.0;
                    if constexpr (FN)
                        if (upO) qN = A * dP / (j == 0 ? njO[0] : j == 1 ? njO[1] : njO[2]);
                    if (col == CS - 1) {
                        // pressure: sgn A, the density term when the other
                        // cell lacks the phase, half of it when both have it,
                        // the own cell's xi / mu derivatives when it is upwind
                        double v = sgn * A;
                        if (upO || (ownEx && othEx)) {
                            const double R1 = A * (-sOwn[O::RHOP + j] * dGamma);
                            const double T = transJ * sOwn[O::XIP + j] * dP - A * sOwn[O::MUPM + j] * dP;
                            if (!othEx) {
                                v += R1 + T;
                            } else if (ownEx) {
                                v += R1 / 2;
                                if (upO) v += T;
                            }
                        }
                        cv = v;
                        if constexpr (FN)
                            dv = qN;
                        else
                            dv = upO ? sRec[0][R::XI + j] * transJ * dP : 0.0;
                    } else if (col < nsat) {
                        const int s1 = col == 0 ? sph[0] : col == 1 ? sph[1] : sph[2];
                        double v = sgn * A * sOwn[O::DPC + j * 3 + s1];
                        if (upO) v += Akd * sOwn[O::XIM + j] * sOwn[O::DKR + j * 3 + s1] * dP;
                        cv = v;
                    } else if (col < ncolO) {
                        // x column (jx, kk) of the own cell: only phase jx's row
                        int kk = col - nsat, jx = 0;
                        if (kk >= pvO[0]) {
                            kk -= pvO[0];
                            jx = 1;
                            if (kk >= pvO[1]) {
                                kk -= pvO[1];
                                jx = 2;
                            }
                        }
                        if (jx == j) {
                            const double rhox = sOwn[O::RHOX + j * nc + kk];
                            if (upO) {
                                const double rx = othEx ? rhox / 2 : rhox;
                                cv = -A * rx * dGamma + transJ * sOwn[O::XIX + j * nc + kk] * dP -
                                     A * sOwn[O::MUXM + j * nc + kk] * dP;
                                if constexpr (FN) cv -= qN;
                            } else {
                                cv = -A * (rhox / 2) * dGamma;
                            }
                        }
                    }
                }
                sh.sC[j][col] = cv;
                if (col == CS - 1) sh.sD[j] = dv;
            }
        }
        __syncthreads();
        // ---- G[j][cc] = [cc == 0] C[j][pressure] + sum_col C[j][col] dSec[col][cc] ----
        if (lane < 3 * GW) {
            const int j = lane / GW;
            double g = 0.0;
            if ((liveM >> j) & 1u) {
                // columns past the own cell's hold zero coefficients and
                // zero dSec: a fixed trip count keeps dS in registers
                g = gcc == 0 ? sh.sC[j][CS - 1] : 0.0;
#pragma unroll
                for (int kq = 0; kq < D::maxRows; ++kq) g = fma(sh.sC[j][kq], dS[kq], g);
            }
            sh.sG[j][gcc] = g;
        }
        __syncthreads();
        // ---- block = dt (sum_j x_ij G[j] + sum_j d_j dSec[x(j, i)]), then scatter ----
        // (IsoT_FIMn: the block's row 0 is zero, and its lanes 1..nc form
        // this connection's -dFdXs res_n term of right-hand-side row
        // i = lane - 1 with the same sums over the res_n column of G and the
        // own cell's res_n)
        double* dst = F.val + mirror * bsize;
#pragma unroll
        for (int t = 0; t < D::perLane; ++t) {
            const int e = lane + t * kAsmBlock;
            if (e < bsize) {
                const int r = e / ncol, cc = e % ncol;
                bool rhsLane = false;
                if constexpr (FN) rhsLane = t == 0 && r == 0 && twoPh && cc >= 1;
                double v = 0;
                if (r > 0 || rhsLane) {
                    const int i = r > 0 ? r - 1 : cc - 1;
                    const int gc = r > 0 ? cc : ncol;
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        if (!((liveM >> j) & 1u)) continue;
                        const double* rU = ((upOwnM >> j) & 1u) ? sRec[0] : sRec[1];
                        v = fma(rU[R::XIJ + j * nc + i], sh.sG[j][gc], v);
                        if (((upOwnM >> j) & 1u) && i < pvO[j]) {
                            double sec;
                            if constexpr (FN)
                                sec = rhsLane ? resX[j] : sOwn[O::DSEC + (xb[j] + i) * ncol + cc];
                            else
                                sec = sOwn[O::DSEC + (xb[j] + i) * ncol + cc];
                            v = fma(sh.sD[j], sec, v);
                        }
                    }
                }
                if (rhsLane) {
                    rhsAcc += -v;
                    v = 0;
                }
                v = dt * v;
                if (ownB) {
                    acc[t] += v;
                    dst[e] = -1.0 * v;
                } else {
                    acc[t] += -1.0 * v;
                    dst[e] = v;
                }
            }
        }
        // the next iteration overwrites sRec[1] and the tables only after
        // its first barrier, which every lane reaches after these reads
    }
    double* diag = F.val + (int64_t)F.IA[n] * bsize;
#pragma unroll
    for (int t = 0; t < D::perLane; ++t) {
        const int e = lane + t * kAsmBlock;
        if (e < bsize) diag[e] = acc[t];
    }
    if constexpr (FN)
        if (lane < ncol) F.rhsN[n * ncol + lane] = rhsAcc;
}

template <int NC, bool FN = false>
__global__ void __launch_bounds__(kAsmBlock) __attribute__((amdgpu_waves_per_eu(NC <= 6 ? 5 : 4))) assembleKernel(DevBulk B, FimDevice F, double dt)
{
    __shared__ AsmShared<NC> sh;
    // workgroup b runs on XCD b % kXcds; XCD x takes rows [x*chunk, (x+1)*chunk)
    const int64_t chunk = gridDim.x / kXcds;
    const int64_t n = (int64_t)(blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
    if (n >= B.nb || !cellOn(B, n)) return;
    assembleRow<NC, FN>(B, F, dt, n, sh);
}

// LinearSystem::AssembleRhsAccumulate (src/LinearSystem.cpp:70-76) of the
// IsoT_FIMn right-hand side: rhsN = (bulk and perforation terms) + res; the
// well rows hold no terms of their own.
__global__ void __launch_bounds__(256) rhsAccumulateKernel(double* __restrict__ rhsN, const double* __restrict__ res,
                                                           int64_t bulkLen, int64_t len)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) rhsN[k] = (k < bulkLen ? rhsN[k] : 0.0) + res[k];
}

// The rows in rows[0, *count) (the overlapped form's recomputation).
template <int NC, bool FN = false>
__global__ void __launch_bounds__(kAsmBlock) assembleListKernel(DevBulk B, FimDevice F, double dt,
                                                                const int32_t* __restrict__ rows,
                                                                const int32_t* __restrict__ count)
{
    __shared__ AsmShared<NC> sh;
    const int64_t total = *count;
    for (int64_t i = blockIdx.x; i < total; i += gridDim.x) assembleRow<NC, FN>(B, F, dt, rows[i], sh);
}

template <int NC>
static hipError_t launchCalResT(const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, bool norms,
                                hipStream_t s)
{
    const unsigned grid = (unsigned)((B.nb + kResBlock - 1) / kResBlock);
    hipLaunchKernelGGL(calResKernel<NC>, dim3(grid), dim3(kResBlock), 0, s, B, F, W, dt);
    if (norms) hipLaunchKernelGGL(resNormKernel<NC>, dim3(grid), dim3(kResBlock), 0, s, B, F);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchCalResRowsT(const DevBulk& B, FimDevice& F, const WellDevice& W, double dt,
                                    const int32_t* rows, const int32_t* count, hipStream_t s)
{
    const unsigned grid = (unsigned)std::min<int64_t>((B.nb + kResBlock - 1) / kResBlock, 1024);
    hipLaunchKernelGGL(calResListKernel<NC>, dim3(grid), dim3(kResBlock), 0, s, B, F, W, dt, rows, count);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchResNormT(const DevBulk& B, FimDevice& F, hipStream_t s)
{
    const unsigned grid = (unsigned)((B.nb + kResBlock - 1) / kResBlock);
    hipLaunchKernelGGL(resNormKernel<NC>, dim3(grid), dim3(kResBlock), 0, s, B, F);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchAssembleRowsT(const DevBulk& B, FimDevice& F, double dt, const int32_t* rows,
                                      const int32_t* count, hipStream_t s)
{
    const unsigned grid = (unsigned)std::min<int64_t>(B.nb, 4096);
    if (F.fimn)
        hipLaunchKernelGGL((assembleListKernel<NC, true>), dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt, rows,
                           count);
    else
        hipLaunchKernelGGL((assembleListKernel<NC>), dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt, rows, count);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchAssembleT(const DevBulk& B, FimDevice& F, double dt, hipStream_t s)
{
    const unsigned grid = (unsigned)(((B.nb + kXcds - 1) / kXcds) * kXcds);
    hipLaunchKernelGGL(assembleKernel<NC>, dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchAssembleNT(const DevBulk& B, FimDevice& F, double dt, hipStream_t s)
{
    if (!F.fimn || !F.rhsN || !B.nj || !B.resN || !B.resPc) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(((B.nb + kXcds - 1) / kXcds) * kXcds);
    hipLaunchKernelGGL((assembleKernel<NC, true>), dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt);
    return hipGetLastError();
}

hipError_t launchRhsAccumulate(FimDevice& F, hipStream_t s)
{
    if (!F.rhsN) return hipErrorInvalidValue;
    const int64_t len = F.nrow * F.ncol;
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(rhsAccumulateKernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, F.rhsN, F.res,
                       F.nb * F.ncol, len);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// IsoT_FIM::GetSolution (src/OCPFluidMethod.cpp:2735-2857), one lane per
// cell, in the reference's operation order (DaAxpby, src/DenseMat.cpp:
// 203-219, for dSec_dPri u).  The signed first maxima NRdPmax / NRdNmax
// (largest magnitude, lowest index on ties, as the reference's sequential
// scan finds them) reduce per block here and over blocks on the host.
// ---------------------------------------------------------------------------
constexpr int kNewtonBlock = 256;

// First maximum by magnitude: (key, value, index); key 0 never wins (the
// reference starts from 0 and replaces on a strictly larger magnitude).
struct FirstMax {
    double key, val;
    int64_t idx;
};
__device__ __forceinline__ void fmTake(FirstMax& a, const FirstMax& b)
{
    if (b.key > a.key || (b.key == a.key && b.key > 0 && b.idx < a.idx)) a = b;
}

__device__ __forceinline__ void fmBlockReduce(FirstMax* sh, FirstMax v, double* part, int64_t* partIdx, int slot)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = kNewtonBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) fmTake(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 2 * slot + 0] = sh[0].val;
        part[4 * blockIdx.x + 2 * slot + 1] = sh[0].key;
        partIdx[2 * blockIdx.x + slot] = sh[0].idx;
    }
    __syncthreads();
}

template <int NC>
__global__ void __launch_bounds__(kNewtonBlock) getSolutionKernel(DevBulk B, NewtonDevice N,
                                                                  const double* __restrict__ u, double dSmaxlim)
{
    constexpr int nc = NC + 1;
    constexpr int col = nc + 1;
    constexpr int maxRows = 3 + 2 * NC;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ld = B.ld;
    FirstMax bp{0.0, 0.0, INT64_MAX}, bn{0.0, 0.0, INT64_MAX};
    if (n < B.nb) {
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            N.dSNR[j * ld + n] = B.S[j * ld + n];
            N.pexOld[j * ld + n] = B.phaseExist[j * ld + n];
        }
        const int rows = min(max(B.bRowSizedSdP[n], 0), maxRows);
        double dtmp[maxRows];
#pragma unroll
        for (int r = 0; r < maxRows; ++r) {
            double y = 0;
            if (r < rows) {
                y = 1.0 * y;
#pragma unroll
                for (int c = 0; c < col; ++c) y += 1.0 * B.dSec[(int64_t)(r * col + c) * ld + n] * u[n * col + c];
            }
            dtmp[r] = y;
        }
        double chopmin = 1;
        int js = 0;
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            if (!B.pSderExist[j * ld + n]) continue;
            const double S = B.S[j * ld + n];
            double d = 0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (r == js) d = dtmp[r];
            double choptmp = 1;
            if (fabs(d) > dSmaxlim) {
                choptmp = dSmaxlim / fabs(d);
            } else if (S + d < 0.0) {
                choptmp = 0.9 * S / fabs(d);
            }
            chopmin = (choptmp < chopmin) ? choptmp : chopmin;
            js++;
        }
        js = 0;
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            if (!B.pSderExist[j * ld + n]) {
                N.dSNRP[j * ld + n] = 0;
                continue;
            }
            double d = 0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (r == js) d = dtmp[r];
            const double dS = chopmin * d;
            N.dSNRP[j * ld + n] = dS;
            B.S[j * ld + n] += dS;
            js++;
        }
        if (B.phaseNum[n] >= 3) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    double d = 0;
#pragma unroll
                    for (int r = 0; r < maxRows; ++r)
                        if (r == js) d = dtmp[r];
                    B.xij[(int64_t)(j * nc + i) * ld + n] += chopmin * d;
                    js++;
                }
        }
        const double dP = u[n * col];
        bp = FirstMax{fabs(dP), dP, n};
        if (!(bp.key > 0)) bp = FirstMax{0.0, 0.0, INT64_MAX};
        B.P[n] += dP;
        const double Nt = B.Nt[n];
#pragma unroll
        for (int i = 0; i < nc; ++i) {
            const double dN = u[n * col + 1 + i] * chopmin;
            if (bn.key < fabs(dN) / Nt) bn = FirstMax{fabs(dN) / Nt, dN / Nt, n * nc + i};
            B.Ni[(int64_t)i * ld + n] += dN;
        }
    }
    __shared__ FirstMax sh[kNewtonBlock];
    fmBlockReduce(sh, bp, N.part, N.partIdx, 0);
    fmBlockReduce(sh, bn, N.part, N.partIdx, 1);
}

// IsoT_FIMn::GetSolution (src/OCPFluidMethod.cpp:3847-3990), one lane per
// cell in the reference's operation order: the secondary changes res_n +
// dSec_dPri u, the saturation chop (the limit dSmaxlim, the critical
// saturations scm, non-negativity), the phase moles nj and compositions
// x_ij of each phase with saturation derivatives, P and Ni.  Saturations
// are not updated (the next flash sets them); dSNRP keeps the predicted
// changes.  The reference takes the count of n_ij rows of phase j from
// bk.pVnumCom[j], cell 0's entry, for every cell (:3933-3936): kept.
template <int NC>
__global__ void __launch_bounds__(kNewtonBlock) getSolutionNKernel(DevBulk B, NewtonDevice N,
                                                                   const double* __restrict__ u, double dSmaxlim,
                                                                   double scm0, double scm1, double scm2)
{
    constexpr int nc = NC + 1;
    constexpr int col = nc + 1;
    constexpr int maxRows = 3 + 2 * NC;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ld = B.ld;
    FirstMax bp{0.0, 0.0, INT64_MAX}, bn{0.0, 0.0, INT64_MAX};
    if (n < B.nb) {
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            N.dSNR[j * ld + n] = B.S[j * ld + n];
            N.pexOld[j * ld + n] = B.phaseExist[j * ld + n];
        }
        const double dP = u[n * col];
        B.P[n] += dP;
        bp = FirstMax{fabs(dP), dP, n};
        if (!(bp.key > 0)) bp = FirstMax{0.0, 0.0, INT64_MAX};
        const int rows = min(max(B.bRowSizedSdP[n], 0), maxRows);
        double dtmp[maxRows];
#pragma unroll
        for (int r = 0; r < maxRows; ++r) {
            double y = 0;
            if (r < rows) {
                y = B.resN[(int64_t)r * ld + n];
                y = 1.0 * y;
#pragma unroll
                for (int c = 0; c < col; ++c) y += 1.0 * B.dSec[(int64_t)(r * col + c) * ld + n] * u[n * col + c];
            }
            dtmp[r] = y;
        }
        auto at = [&](int r) {
            double d = 0;
#pragma unroll
            for (int q = 0; q < maxRows; ++q)
                if (q == r) d = dtmp[q];
            return d;
        };
        const double scm[3] = {scm0, scm1, scm2};
        double dSmax = 0;
        int js = 0;
#pragma unroll
        for (int j = 0; j < kNP; ++j)
            if (B.pSderExist[j * ld + n]) {
                const double a = fabs(at(js));
                dSmax = (a < dSmax) ? dSmax : a;  // std::max(fabs(dtmp[js]), dSmax)
                js++;
            }
        double chop = 1;
        if (dSmax > dSmaxlim) {
            const double t = dSmaxlim / dSmax;
            chop = (t < chop) ? t : chop;
        }
        js = 0;
#pragma unroll
        for (int j = 0; j < kNP; ++j)
            if (B.pSderExist[j * ld + n]) {
                const double S = B.S[j * ld + n];
                const double d = at(js);
                if (fabs(S - scm[j]) > 1E-8 && (S - scm[j]) / (chop * d) < 0) {
                    const double t = -((S - scm[j]) / (chop * d));
                    chop *= (t < 1.0) ? t : 1.0;  // std::min(1.0, t)
                }
                if (S + chop * d < 0) {
                    const double t = -S / (chop * d);
                    chop *= (t < 1.0) ? t : 1.0;
                }
                js++;
            }
        js = 0;
        int jx = B.phaseNum[n];
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            if (!B.pSderExist[j * ld + n]) continue;
            N.dSNRP[j * ld + n] = chop * at(js);
            js++;
            const double nj = B.nj[j * ld + n];
            const int pv = min(max(B.pVnumCom[j * ld + 0], 0), NC);
            double t[nc];
#pragma unroll
            for (int i = 0; i < nc; ++i) {
                t[i] = 0.0;
                t[i] += nj * B.xij[(int64_t)(j * nc + i) * ld + n];
            }
#pragma unroll
            for (int i = 0; i < NC; ++i)
                if (i < pv) t[i] += chop * at(jx + i);
            jx += pv;
            double s1 = 0;
#pragma unroll
            for (int i = 0; i < nc; ++i) s1 += fabs(t[i]);
            B.nj[j * ld + n] = s1;
#pragma unroll
            for (int i = 0; i < nc; ++i) B.xij[(int64_t)(j * nc + i) * ld + n] = t[i] / s1;
        }
        const double Nt = B.Nt[n];
#pragma unroll
        for (int i = 0; i < nc; ++i) {
            const double dN = u[n * col + 1 + i] * chop;
            if (bn.key < fabs(dN) / Nt) bn = FirstMax{fabs(dN) / Nt, dN / Nt, n * nc + i};
            B.Ni[(int64_t)i * ld + n] += dN;
        }
    }
    __shared__ FirstMax sh[kNewtonBlock];
    fmBlockReduce(sh, bp, N.part, N.partIdx, 0);
    fmBlockReduce(sh, bn, N.part, N.partIdx, 1);
}

// Well::SetBHP(BHP() + u) of the open wells (src/OCPFluidMethod.cpp:2846-2854).
__global__ void wellBhpUpdateKernel(WellDevice W, const double* __restrict__ u, int64_t nb, int col)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < W.nw && W.row[w] >= 0) W.bhp[w] += u[(nb + W.row[w]) * col];
}

// PassFlashValue's dSNR = S - dSNR and maxNRdSSP (src/OCPFluidMethod.cpp:
// 1135-1146) and Bulk::CalNRdSmax (src/Bulk.cpp:1235-1247).
__global__ void __launch_bounds__(kNewtonBlock) nrChangeKernel(DevBulk B, NewtonDevice N)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ld = B.ld;
    FirstMax bs{0.0, 0.0, INT64_MAX}, bd{0.0, 0.0, INT64_MAX};
    if (n < B.nb) {
#pragma unroll
        for (int j = 0; j < kNP; ++j) {
            const double d = B.S[j * ld + n] - N.dSNR[j * ld + n];
            N.dSNR[j * ld + n] = d;
            if (N.pexOld[j * ld + n]) {
                const double e = d - N.dSNRP[j * ld + n];
                if (bs.key < fabs(e)) bs = FirstMax{fabs(e), e, n};
            }
            if (bd.key < fabs(d)) bd = FirstMax{fabs(d), d, n * kNP + j};
        }
    }
    __shared__ FirstMax sh[kNewtonBlock];
    fmBlockReduce(sh, bd, N.part, N.partIdx, 0);
    fmBlockReduce(sh, bs, N.part, N.partIdx, 1);
}

template <int NC>
static hipError_t launchGetSolutionT(const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                                     const double* u, double dSmaxlim, hipStream_t s)
{
    hipLaunchKernelGGL(getSolutionKernel<NC>, dim3(N.blocks), dim3(kNewtonBlock), 0, s, B, N, u, dSmaxlim);
    if (W.nw > 0)
        hipLaunchKernelGGL(wellBhpUpdateKernel, dim3((W.nw + 63) / 64), dim3(64), 0, s, W, u, nb, NC + 2);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchGetSolutionNT(const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                                      const double* u, double dSmaxlim, const double* scm, hipStream_t s)
{
    if (!B.nj || !B.resN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(getSolutionNKernel<NC>, dim3(N.blocks), dim3(kNewtonBlock), 0, s, B, N, u, dSmaxlim, scm[0],
                       scm[1], scm[2]);
    if (W.nw > 0)
        hipLaunchKernelGGL(wellBhpUpdateKernel, dim3((W.nw + 63) / 64), dim3(64), 0, s, W, u, nb, NC + 2);
    return hipGetLastError();
}

hipError_t launchNRChange(const DevBulk& B, NewtonDevice& N, hipStream_t s)
{
    hipLaunchKernelGGL(nrChangeKernel, dim3(N.blocks), dim3(kNewtonBlock), 0, s, B, N);
    return hipGetLastError();
}

// Bulk::CheckNi and Bulk::CheckP (src/Bulk.cpp:1250-1308): bit 1 of *flag
// when a component's moles are negative, bit 0 when a pressure is.
__global__ void __launch_bounds__(kNewtonBlock) bulkCheckKernel(DevBulk B, int nc, int32_t* flag)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0;
    if (n < B.nb) {
        if (B.P[n] < 0) f |= 1;
        for (int i = 0; i < nc; ++i)
            if (B.Ni[(int64_t)i * B.ld + n] < 0) f |= 2;
    }
    if (f) atomicOr(flag, f);
}

// Bulk::CalMaxChange (src/Bulk.cpp:1335-1387) against the last time step:
// per block the maxima of |dP|, the relative |dNi|, |dS| and the relative
// volume error, into part[4 * block + {0, 1, 2, 3}].
__global__ void __launch_bounds__(kNewtonBlock) maxChangeKernel(DevBulk B, DevBulk L, int nc, double* part)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ld = B.ld;
    double m[4] = {0, 0, 0, 0};
    if (n < B.nb) {
        m[0] = fabs(B.P[n] - L.P[n]);
        for (int j = 0; j < kNP; ++j) m[2] = fmax(m[2], fabs(B.S[j * ld + n] - L.S[j * ld + n]));
        for (int i = 0; i < nc; ++i) {
            const double a = B.Ni[(int64_t)i * ld + n], b = L.Ni[(int64_t)i * ld + n];
            const double t = fabs(fmax(a, b));
            if (t > 1E-8) m[1] = fmax(m[1], fabs(a - b) / t);
        }
        m[3] = fabs(B.vf[n] - B.rockVp[n]) / B.rockVp[n];
    }
    __shared__ double sh[4][kNewtonBlock];
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = m[q];
    __syncthreads();
    for (int st = kNewtonBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) part[4 * (int64_t)blockIdx.x + q] = sh[q][0];
}

hipError_t launchBulkCheck(const DevBulk& B, int nc, int32_t* flag, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(bulkCheckKernel, dim3((unsigned)((B.nb + kNewtonBlock - 1) / kNewtonBlock)),
                       dim3(kNewtonBlock), 0, s, B, nc, flag);
    return hipGetLastError();
}

hipError_t launchMaxChange(const DevBulk& B, const DevBulk& L, int nc, double* part, int blocks, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(maxChangeKernel, dim3(blocks), dim3(kNewtonBlock), 0, s, B, L, nc, part);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Serving pool: the activity mask and the restore of the active cells' saved
// iteration inputs (a kernel instead of whole-array copies, so an iteration
// over a few realisations of a large pool copies only theirs).
__global__ void __launch_bounds__(256) maskFillKernel(uint8_t* mask, int64_t base, int64_t n, uint8_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[base + i] = v;
}

hipError_t launchMaskFill(uint8_t* mask, int64_t base, int64_t n, uint8_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(maskFillKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mask, base, n, v);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) poolRebaseKernel(int32_t* IA, int32_t* bulkEnd, int64_t nb, int32_t* JA,
                                                         int32_t* slotConn, int32_t* slotMirror, int64_t nnzb,
                                                         int32_t* bId, int32_t* eId, int64_t nconn, int32_t dCell,
                                                         int32_t dConn, int32_t dBlock)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) {
        IA[i] += dBlock;
        bulkEnd[i] += dBlock;
    }
    if (i < nnzb) {
        JA[i] += dCell;
        if (slotConn[i] >= 0) slotConn[i] += dConn;
        slotMirror[i] += dBlock;
    }
    if (i < nconn) {
        bId[i] += dCell;
        eId[i] += dCell;
    }
}

hipError_t launchPoolRebase(int32_t* IA, int32_t* bulkEnd, int64_t nb, int32_t* JA, int32_t* slotConn,
                            int32_t* slotMirror, int64_t nnzb, int32_t* bId, int32_t* eId, int64_t nconn,
                            int64_t dCell, int64_t dConn, int64_t dBlock, hipStream_t s)
{
    const int64_t n = std::max(nb, std::max(nnzb, nconn));
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(poolRebaseKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, IA, bulkEnd, nb, JA,
                       slotConn, slotMirror, nnzb, bId, eId, nconn, (int32_t)dCell, (int32_t)dConn, (int32_t)dBlock);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) maskedCopyKernel(const uint8_t* __restrict__ mask, int64_t nb, int64_t ld,
                                                         MaskedCopy d)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb || !mask[n]) return;
    for (int f = 0; f < d.nfields; ++f) {
        const int e = d.elem[f];
        for (int r = 0; r < d.rows[f]; ++r) {
            const int64_t o = ((int64_t)r * ld + n) * e;
            if (e == 8)
                *reinterpret_cast<double*>(d.dst[f] + o) = *reinterpret_cast<const double*>(d.src[f] + o);
            else if (e == 4)
                *reinterpret_cast<int32_t*>(d.dst[f] + o) = *reinterpret_cast<const int32_t*>(d.src[f] + o);
            else
                d.dst[f][o] = d.src[f][o];
        }
    }
}

hipError_t launchMaskedCopy(const uint8_t* mask, int64_t nb, int64_t ld, const MaskedCopy& d, hipStream_t s)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(maskedCopyKernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, mask, nb, ld, d);
    return hipGetLastError();
}

#define OCPG_FIM_NC(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

hipError_t launchGetSolution(int NC, const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                             const double* u, double dSmaxlim, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N_) \
    if (NC == N_) return launchGetSolutionT<N_>(B, N, W, nb, u, dSmaxlim, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchGetSolutionN(int NC, const DevBulk& B, NewtonDevice& N, const WellDevice& W, int64_t nb,
                              const double* u, double dSmaxlim, const double* scm, hipStream_t s)
{
#define X(N_) \
    if (NC == N_) return launchGetSolutionNT<N_>(B, N, W, nb, u, dSmaxlim, scm, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchCalRes(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, hipStream_t s,
                        bool norms)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchCalResT<N>(B, F, W, dt, norms, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchCalResRows(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt,
                            const int32_t* rows, const int32_t* count, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchCalResRowsT<N>(B, F, W, dt, rows, count, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchResNorm(int NC, const DevBulk& B, FimDevice& F, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchResNormT<N>(B, F, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchAssembleRows(int NC, const DevBulk& B, FimDevice& F, double dt, const int32_t* rows,
                              const int32_t* count, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchAssembleRowsT<N>(B, F, dt, rows, count, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

template <int NC>
static hipError_t launchPackT(const DevBulk& B, FimDevice& F, hipStream_t s)
{
    const unsigned grid = (unsigned)((B.nb + packTile<NC>() - 1) / packTile<NC>());
    hipLaunchKernelGGL(packRecordsKernel<NC>, dim3(grid), dim3(kPackBlock), 0, s, B, F);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchCalResPackT(const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, hipStream_t s)
{
    const int64_t groups = (B.nb + kResBlock - 1) / kResBlock;
    const unsigned grid = (unsigned)(groups * (1 + packTilesPerResBlock<NC>()));
    hipLaunchKernelGGL(calResPackKernel<NC>, dim3(grid), dim3(kResBlock), 0, s, B, F, W, dt);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchPackListT(const DevBulk& B, FimDevice& F, const CellLists& L, hipStream_t s)
{
    hipLaunchKernelGGL(packRecordsListKernel<NC>, dim3(256), dim3(256), 0, s, B, F, L);
    return hipGetLastError();
}

int64_t recordDoubles(int NC)
{
#define X(N) \
    if (NC == N) return recStride<N>();
    OCPG_FIM_NC(X)
#undef X
    return 0;
}

int64_t ownDoubles(int NC)
{
#define X(N) \
    if (NC == N) return ownStride<N>();
    OCPG_FIM_NC(X)
#undef X
    return 0;
}

hipError_t launchPackRecords(int NC, const DevBulk& B, FimDevice& F, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchPackT<N>(B, F, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchCalResPack(int NC, const DevBulk& B, FimDevice& F, const WellDevice& W, double dt, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchCalResPackT<N>(B, F, W, dt, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchPackRecordsList(int NC, const DevBulk& B, FimDevice& F, const CellLists& L, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchPackListT<N>(B, F, L, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchAssembleN(int NC, const DevBulk& B, FimDevice& F, double dt, hipStream_t s)
{
#define X(N) \
    if (NC == N) return launchAssembleNT<N>(B, F, dt, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchAssemble(int NC, const DevBulk& B, FimDevice& F, double dt, hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
#define X(N) \
    if (NC == N) return launchAssembleT<N>(B, F, dt, s);
    OCPG_FIM_NC(X)
#undef X
    return hipErrorInvalidValue;
}

void fimFree(FimDevice& f)
{
    void* ptrs[] = {f.bId, f.eId, f.area, f.IA, f.JA, f.slotConn, f.slotMirror, f.bulkEnd, f.upblock,
                    f.upRho, f.upVel, f.res, f.val, f.normPart, f.normIdx, f.rec, f.own, f.rhsN};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    f = FimDevice{};
}

size_t fimBytes(const FimDevice& f) { return f.bytes; }

}  // namespace ocpg
