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
    if (n >= B.nb) return;
    krPcRockCell(B, sat, rock, krpc, rockOn, n);
}

// kr/Pc and rock of the cells in two lists (the overlapped flash's side cells).
__global__ void __launch_bounds__(256) krPcRockListKernel(DevBulk B, const DevSat* __restrict__ sat, DevRock rock,
                                                          const int32_t* __restrict__ L0,
                                                          const int32_t* __restrict__ L1,
                                                          const int32_t* __restrict__ c0,
                                                          const int32_t* __restrict__ c1)
{
    const int64_t n0 = *c0, total = n0 + *c1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        krPcRockCell(B, sat, rock, 1, 1, i < n0 ? L0[i] : L1[i - n0]);
}

hipError_t launchKrPcRockList(const DevBulk& B, const DevSat* sat, const DevRock& rock, const int32_t* const* lists,
                              const int32_t* const* counts, hipStream_t s)
{
    hipLaunchKernelGGL(krPcRockListKernel, dim3(64), dim3(256), 0, s, B, sat, rock, lists[0], lists[1], counts[0],
                       counts[1]);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rows the overlapped form recomputes once its side streams have joined:
// what CalRes and the assembly computed on the main stream is final for a
// row whose cell and neighbours all finished there, and is recomputed for
// the side cells (bit 2) and their neighbours (bit 1); CalRes also for the
// perforated cells (bit 4: the wells' rates read the side cells too).
// mark[] is all zero between calls: the compaction clears what it reads.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fixupMarkKernel(FimDevice F, const int32_t* __restrict__ L0,
                                                       const int32_t* __restrict__ L1,
                                                       const int32_t* __restrict__ c0,
                                                       const int32_t* __restrict__ c1, uint32_t* mark,
                                                       int32_t* counts)
{
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[threadIdx.x] = 0;
    const int64_t n0 = *c0, total = n0 + *c1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = i < n0 ? L0[i] : L1[i - n0];
        atomicOr(&mark[t], 2u);
        for (int k = F.IA[t] + 1; k < F.bulkEnd[t]; ++k) atomicOr(&mark[F.JA[k]], 1u);
    }
}

__global__ void __launch_bounds__(256) fixupMarkPerfKernel(WellDevice W, uint32_t* mark)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < W.npcell) atomicOr(&mark[W.pcell[i]], 4u);
}

__device__ __forceinline__ void appendRow(int32_t* list, int32_t* count, bool pred, int32_t row)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    if (pred) base += __popcll(lane == 0 ? 0ull : (m & ((1ull << lane) - 1ull)));
    if (pred) list[base] = row;
}

__global__ void __launch_bounds__(256) fixupCompactKernel(int64_t nb, uint32_t* mark, int32_t* resRows,
                                                          int32_t* asmRows, int32_t* counts)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = 0;
    if (n < nb) {
        m = mark[n];
        if (m) mark[n] = 0;
    }
    appendRow(resRows, counts + 0, m != 0, (int32_t)n);
    appendRow(asmRows, counts + 1, (m & 3u) != 0, (int32_t)n);
}

hipError_t launchFixupRows(const FimDevice& F, const WellDevice& W, int64_t nb, const int32_t* const* lists,
                           const int32_t* const* counts, const FixupRows& R, hipStream_t s)
{
    hipLaunchKernelGGL(fixupMarkKernel, dim3(64), dim3(256), 0, s, F, lists[0], lists[1], counts[0], counts[1],
                       R.mark, R.counts);
    if (W.npcell > 0)
        hipLaunchKernelGGL(fixupMarkPerfKernel, dim3((W.npcell + 255) / 256), dim3(256), 0, s, W, R.mark);
    hipLaunchKernelGGL(fixupCompactKernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, nb, R.mark,
                       R.resRows, R.asmRows, R.counts);
    return hipGetLastError();
}

hipError_t launchKrPcRock(const DevBulk& B, const DevSat* sat, const DevRock& rock, bool krpc, bool rockOn,
                          hipStream_t s)
{
    if (B.nb == 0) return hipSuccess;
    const unsigned grid = (unsigned)((B.nb + 255) / 256);
    hipLaunchKernelGGL(krPcRockKernel, dim3(grid), dim3(256), 0, s, B, sat, rock, krpc ? 1 : 0, rockOn ? 1 : 0);
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
    if (n < B.nb) calResCell<NC>(B, F, W, dt, n);
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
    if (n < B.nb) {
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
// Memory plan: the cell values a connection side needs (its upwind phase
// properties, pressures, depth, phase flags) form a record of AsmRec::LEN
// doubles.  The wavefront gathers a record with one load per lane (each lane
// owns a fixed record entry), keeps the own cell's record and the current
// neighbour's in LDS, and prefetches the next neighbour's record into
// registers while it works on the current connection.  Rows are handed to
// workgroups so that each XCD walks a contiguous range of cells, which keeps
// the neighbour records (cells n +- 1, n +- nx, ...) in that XCD's L2.
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
    static constexpr int dfStride = maxRows + 1;        // [i][0] = dFdXp, [i][1+k] = dFdXs
};

// Entry offsets of the per-cell record (phase j, component i).
template <int NC>
struct AsmRec {
    static constexpr int nc = NC + 1;
    static constexpr int PJ = 0, XI = 3, KR = 6, MU = 9, MUP = 12, XIP = 15, RHOP = 18;
    static constexpr int DKR = 21;  // dKr_dS[j*3+k]
    static constexpr int XIJ = 30;  // xij[j*nc+i]
    static constexpr int RHOX = XIJ + 3 * nc, XIX = RHOX + 3 * nc, MUX = XIX + 3 * nc;
    static constexpr int DEPTH = MUX + 3 * nc, EX = DEPTH + 1, LEN = EX + 3;
    static constexpr int perLane = (LEN + kAsmBlock - 1) / kAsmBlock;
};

// Where record entry q lives: a double field advanced to its entry, or (for
// the phase flags) a byte field.
struct RecSrc {
    const double* d;
    const uint8_t* b;
};

template <int NC>
__device__ __forceinline__ RecSrc recSource(const DevBulk& B, int q)
{
    using R = AsmRec<NC>;
    const int64_t ld = B.ld;
    RecSrc s{nullptr, nullptr};
    if (q < R::XI) s.d = B.Pj + (int64_t)(q - R::PJ) * ld;
    else if (q < R::KR) s.d = B.xi + (int64_t)(q - R::XI) * ld;
    else if (q < R::MU) s.d = B.kr + (int64_t)(q - R::KR) * ld;
    else if (q < R::MUP) s.d = B.mu + (int64_t)(q - R::MU) * ld;
    else if (q < R::XIP) s.d = B.muP + (int64_t)(q - R::MUP) * ld;
    else if (q < R::RHOP) s.d = B.xiP + (int64_t)(q - R::XIP) * ld;
    else if (q < R::DKR) s.d = B.rhoP + (int64_t)(q - R::RHOP) * ld;
    else if (q < R::XIJ) s.d = B.dKr_dS + (int64_t)(q - R::DKR) * ld;
    else if (q < R::RHOX) s.d = B.xij + (int64_t)(q - R::XIJ) * ld;
    else if (q < R::XIX) s.d = B.rhox + (int64_t)(q - R::RHOX) * ld;
    else if (q < R::MUX) s.d = B.xix + (int64_t)(q - R::XIX) * ld;
    else if (q < R::DEPTH) s.d = B.mux + (int64_t)(q - R::MUX) * ld;
    else if (q < R::EX) s.d = B.depth;
    else if (q < R::LEN) s.b = B.phaseExist + (int64_t)(q - R::EX) * ld;
    return s;
}

__device__ __forceinline__ double loadRec(const RecSrc& s, int64_t cell)
{
    if (s.d) return s.d[cell];
    if (s.b) return s.b[cell] ? 1.0 : 0.0;
    return 0.0;
}

template <int NC>
struct AsmShared {
    using D = AsmDims<NC>;
    using R = AsmRec<NC>;
    double sSec[D::maxRows * D::ncol];  // own dSec_dPri (compact rows x ncol)
    double sDF[D::nc * D::dfStride];    // own side's dFdXp | dFdXs
    double sRec[2][R::LEN];             // [0] own cell, [1] current neighbour
    double sDpc[9];                     // own dPcj_dS
    double sPh[2][3];                   // per phase of the current connection: dP, transJ
};

// Row n of the bulk assembly, by one wavefront.
template <int NC>
__device__ __forceinline__ void assembleRow(const DevBulk& B, const FimDevice& F, double dt, int64_t n,
                                            AsmShared<NC>& sh)
{
    using D = AsmDims<NC>;
    using R = AsmRec<NC>;
    constexpr int nc = D::nc;
    constexpr int ncol = D::ncol;
    constexpr int bsize = D::bsize;
    constexpr int W = D::dfStride;
    const int64_t ld = B.ld;
    const int lane = threadIdx.x;
    double* sSec = sh.sSec;
    double* sDF = sh.sDF;
    double (*sRec)[R::LEN] = sh.sRec;
    double* sDpc = sh.sDpc;
    double (*sPh)[3] = sh.sPh;
    // a previous row of this wavefront may still read the shared arrays
    __syncthreads();

    RecSrc src[R::perLane];
#pragma unroll
    for (int t = 0; t < R::perLane; ++t) src[t] = recSource<NC>(B, lane + t * kAsmBlock);

#pragma unroll
    for (int t = 0; t < R::perLane; ++t) {
        const int q = lane + t * kAsmBlock;
        if (q < R::LEN) sRec[0][q] = loadRec(src[t], n);
    }
    const int rowsOwn = min(max(B.bRowSizedSdP[n], 0), D::maxRows);
    for (int q = lane; q < rowsOwn * ncol; q += kAsmBlock) sSec[q] = B.dSec[(int64_t)q * ld + n];
    if (lane < 9) sDpc[lane] = B.dPcj_dS[(int64_t)lane * ld + n];

    // secondary columns of this row's cell: existing saturations, then the
    // x_ij of each phase with derivatives
    bool sdO[3];
    int pvO[3];
    int nsat = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        sdO[j] = B.pSderExist[j * ld + n];
        pvO[j] = min(max(B.pVnumCom[j * ld + n], 0), NC);
        nsat += sdO[j] ? 1 : 0;
    }
    const int ncolO = min(nsat + pvO[0] + pvO[1] + pvO[2], D::maxRows);

    // diagonal starts as the accumulation block (src/OCPFluidMethod.cpp:1586-1597)
    double acc[D::perLane];
#pragma unroll
    for (int t = 0; t < D::perLane; ++t) {
        const int e = lane + t * kAsmBlock;
        const int r = e / ncol, cc = e % ncol;
        double v = 0;
        if (e < bsize) {
            if (r == 0 && cc == 0)
                v = B.v[n] * B.poroP[n] - B.vfP[n];
            else if (r == 0)
                v = -B.vfi[(cc - 1) * ld + n];
            else if (r == cc)
                v = 1;
        }
        acc[t] = v;
    }

    const int k0 = F.IA[n] + 1, k1 = F.bulkEnd[n];
    double pre[R::perLane];
    if (k0 < k1) {
        const int64_t m = F.JA[k0];
#pragma unroll
        for (int t = 0; t < R::perLane; ++t) pre[t] = loadRec(src[t], m);
    }
    for (int k = k0; k < k1; ++k) {
        const int64_t m = F.JA[k];
        const int64_t c = F.slotConn[k];
#pragma unroll
        for (int t = 0; t < R::perLane; ++t) {
            const int q = lane + t * kAsmBlock;
            if (q < R::LEN) sRec[1][q] = pre[t];
        }
        __syncthreads();
        if (k + 1 < k1) {
            const int64_t m2 = F.JA[k + 1];
#pragma unroll
            for (int t = 0; t < R::perLane; ++t) pre[t] = loadRec(src[t], m2);
        }
        const bool ownB = n < m;  // this row is the connection's bId side
        const double* rB = ownB ? sRec[0] : sRec[1];
        const double* rE = ownB ? sRec[1] : sRec[0];
        const double* rO = sRec[0];
        const double Akd = kCONV1f * kCONV2f * F.area[c];
        const double dGamma = kGravityFactor * (rB[R::DEPTH] - rE[R::DEPTH]);
        bool exB[3], exE[3], upOwn[3];
        double upRho[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            exB[j] = rB[R::EX + j] != 0.0;
            exE[j] = rE[R::EX + j] != 0.0;
            upOwn[j] = F.upblock[c * 3 + j] == (int32_t)n;
            upRho[j] = F.upRho[c * 3 + j];
        }

        // ---- per phase: whether the upwind cell has it, dP and transJ ----
        // (lanes 0..2; every entry below reads them)
        bool live[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const bool upB = ownB ? upOwn[j] : !upOwn[j];
            live[j] = upB ? exB[j] : exE[j];
        }
        if (lane < 3) {
            const int j = lane;
            const double* rU = upOwn[j] ? sRec[0] : sRec[1];
            sPh[0][j] = rB[R::PJ + j] - rE[R::PJ + j] - upRho[j] * dGamma;
            sPh[1][j] = Akd * rU[R::KR + j] / rU[R::MU + j];
        }
        __syncthreads();

        // ---- own side's dFdXp column and dFdXs, entry by entry ----
        // Pressure and saturation columns: every phase contributes, in the
        // reference's phase order.
        const int nPS = 1 + nsat;
        for (int q = lane; q < nc * nPS; q += kAsmBlock) {
            const int i = q / nPS;
            const int col = q % nPS - 1;  // -1: pressure column of dFdXp
            int j1S = -1;
            {
                int s = 0;
#pragma unroll
                for (int j1 = 0; j1 < 3; ++j1)
                    if (sdO[j1]) {
                        if (s == col) j1S = j1;
                        s++;
                    }
            }
            double v = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (!live[j]) continue;
                const bool upB = ownB ? upOwn[j] : !upOwn[j];
                const double* rU = upOwn[j] ? sRec[0] : sRec[1];
                const double dP = sPh[0][j];
                const double transJ = sPh[1][j];
                const double xi = rU[R::XI + j];
                const double mu = rU[R::MU + j];
                const double xij = rU[R::XIJ + j * nc + i];
                const double transIJ = xij * xi * transJ;
                if (col < 0) {
                    const double muP = rU[R::MUP + j];
                    const double xiP = rU[R::XIP + j];
                    const double rhoP = rU[R::RHOP + j];
                    double tmp = xij * transJ * xiP * dP;
                    tmp += -transIJ * muP / mu * dP;
                    if (ownB) {
                        v += transIJ;
                        if (!exE[j]) {
                            tmp += transIJ * (-rhoP * dGamma);
                            v += tmp;
                        } else if (!exB[j]) {
                        } else {
                            v += transIJ * (-rO[R::RHOP + j] * dGamma) / 2;
                            if (upB) v += tmp;
                        }
                    } else {
                        v -= transIJ;
                        if (!exE[j]) {
                        } else if (!exB[j]) {
                            tmp += transIJ * (-rhoP * dGamma);
                            v += tmp;
                        } else {
                            v += transIJ * (-rO[R::RHOP + j] * dGamma) / 2;
                            if (!upB) v += tmp;
                        }
                    }
                } else {
                    // saturation column of phase j1S
                    const double dpc = sDpc[j * 3 + j1S];
                    if (ownB) {
                        v += transIJ * dpc;
                        if (upB) v += Akd * xij * xi / mu * rU[R::DKR + j * 3 + j1S] * dP;
                    } else {
                        v -= transIJ * dpc;
                        if (!upB) v += Akd * xij * xi / mu * rU[R::DKR + j * 3 + j1S] * dP;
                    }
                }
            }
            sDF[i * W + col + 1] = v;
        }
        // x_{kX, jX} columns: only phase jX's pass touches them.
        const int nX = max(ncolO - nsat, 0);
        for (int q = lane; q < nc * nX; q += kAsmBlock) {
            const int i = q / nX;
            int kk = q % nX, j = 0;
            if (kk >= pvO[0]) {
                kk -= pvO[0];
                j = 1;
                if (kk >= pvO[1]) {
                    kk -= pvO[1];
                    j = 2;
                }
            }
            double v = 0;
            if (live[j]) {
                const bool upB = ownB ? upOwn[j] : !upOwn[j];
                const double* rU = upOwn[j] ? sRec[0] : sRec[1];
                const double dP = sPh[0][j];
                const double transJ = sPh[1][j];
                const double xi = rU[R::XI + j];
                const double mu = rU[R::MU + j];
                const double xij = rU[R::XIJ + j * nc + i];
                const double transIJ = xij * xi * transJ;
                const bool ownUp = ownB ? upB : !upB;
                const bool otherEx = ownB ? exE[j] : exB[j];
                if (ownUp) {
                    const double rhox = otherEx ? rO[R::RHOX + j * nc + kk] / 2 : rU[R::RHOX + j * nc + kk];
                    const double xix = rU[R::XIX + j * nc + kk];
                    const double mux = rU[R::MUX + j * nc + kk];
                    double tmp = -transIJ * rhox * dGamma;
                    tmp += xij * transJ * xix * dP;
                    tmp += -transIJ * mux / mu * dP;
                    v += tmp;
                    if (kk == i) v += xi * transJ * dP;
                } else if (otherEx) {
                    v += -transIJ * rO[R::RHOX + j * nc + kk] / 2 * dGamma;
                }
            }
            sDF[i * W + 1 + nsat + q % nX] = v;
        }
        __syncthreads();

        // ---- block = (dFdXp + dFdXs * dSec_own) * dt, then scatter ----
        const int64_t mirror = F.slotMirror[k];
        double* dst = F.val + mirror * bsize;
#pragma unroll
        for (int t = 0; t < D::perLane; ++t) {
            const int e = lane + t * kAsmBlock;
            if (e < bsize) {
                const int r = e / ncol, cc = e % ncol;
                double v = 0;
                if (r > 0) {
                    const double* row = &sDF[(r - 1) * W];
                    v = (cc == 0) ? row[0] : 0.0;
                    for (int kq = 0; kq < ncolO; ++kq) v = v + row[1 + kq] * sSec[kq * ncol + cc];
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
        // the next iteration overwrites sRec[1] (read only before the barrier
        // above) and sDF (after its own barrier)
    }
    double* diag = F.val + (int64_t)F.IA[n] * bsize;
#pragma unroll
    for (int t = 0; t < D::perLane; ++t) {
        const int e = lane + t * kAsmBlock;
        if (e < bsize) diag[e] = acc[t];
    }
}

template <int NC>
__global__ void __launch_bounds__(kAsmBlock) assembleKernel(DevBulk B, FimDevice F, double dt)
{
    __shared__ AsmShared<NC> sh;
    // workgroup b runs on XCD b % kXcds; XCD x takes rows [x*chunk, (x+1)*chunk)
    const int64_t chunk = gridDim.x / kXcds;
    const int64_t n = (int64_t)(blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
    if (n >= B.nb) return;
    assembleRow<NC>(B, F, dt, n, sh);
}

// The rows in rows[0, *count) (the overlapped form's recomputation).
template <int NC>
__global__ void __launch_bounds__(kAsmBlock) assembleListKernel(DevBulk B, FimDevice F, double dt,
                                                                const int32_t* __restrict__ rows,
                                                                const int32_t* __restrict__ count)
{
    __shared__ AsmShared<NC> sh;
    const int64_t total = *count;
    for (int64_t i = blockIdx.x; i < total; i += gridDim.x) assembleRow<NC>(B, F, dt, rows[i], sh);
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
    hipLaunchKernelGGL(assembleListKernel<NC>, dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt, rows, count);
    return hipGetLastError();
}

template <int NC>
static hipError_t launchAssembleT(const DevBulk& B, FimDevice& F, double dt, hipStream_t s)
{
    const unsigned grid = (unsigned)(((B.nb + kXcds - 1) / kXcds) * kXcds);
    hipLaunchKernelGGL(assembleKernel<NC>, dim3(grid), dim3(kAsmBlock), 0, s, B, F, dt);
    return hipGetLastError();
}

#define OCPG_FIM_NC(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

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
                    f.upRho, f.upVel, f.res, f.val, f.normPart, f.normIdx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    f = FimDevice{};
}

size_t fimBytes(const FimDevice& f) { return f.bytes; }

}  // namespace ocpg
