// Device-side data layout of the flash / assembly library.
//
// Cell arrays are structure-of-arrays in HBM: field X of cell n, entry k, is
// X[k * ld + n] with ld = the padded cell count.  Consecutive lanes of a
// wavefront handle consecutive cells, so every per-cell load and store of a
// wavefront is one coalesced 512-byte (FP64) or 64-byte (byte flags) access.
// The entry index k follows the reference's in-cell order (include/Bulk.hpp:
// 98-102): Ni[i], S[j], xij[j*nc+i], dKr_dS[j*np+k], dSec_dPri[r*ncol+c].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocpg {

constexpr int kNP = 3;
constexpr int kMaxNC = 16;
constexpr int kMaxTab = 64;

// EoS constants as the kernels read them (uniform across a launch: every
// access compiles to scalar loads).
struct DevEos {
    int nc;
    int pvtwRows;
    double Tc[kMaxNC], Pc[kMaxNC], Vc[kMaxNC], MWC[kMaxNC], Acf[kMaxNC];
    double OmegaA[kMaxNC], OmegaB[kMaxNC], Vshift[kMaxNC], Vcvis[kMaxNC];
    double oneMinusBIC[kMaxNC * kMaxNC];  // (1 - BIC[i*NC+k]), row-major NC x NC
    double LBCcoef[5];
    double muAux1I[kMaxNC];               // 5.4402 * Tc^(1/6) / Pc^(2/3)
    double sqrtMWi[kMaxNC];
    double pvtw[kMaxTab][5];
    double stdRhoW;
    // controls
    int ssmStaMaxIt, nrStaMaxIt, ssmSpMaxIt, nrSpMaxIt, rrMaxIt;
    double ssmStaTol2, ssmStaEYt, ssmStaKtol, ssmStaDYtol, nrStaTol;
    double ssmSpTol2, nrSpTol, rrTol;
};

struct DevSat {
    int swofRows, sgofRows;
    double swof[kMaxTab][4];
    double sgof[kMaxTab][4];
    double kroMax;
};

struct DevRock {
    double Pref, cp1, cp2;
};

// SoA cell state.  Every pointer addresses ld * (entries per cell) values.
struct DevBulk {
    int64_t nb;  // cells
    int64_t ld;  // leading dimension (>= nb)
    double *P, *T, *Ni, *lNi, *S, *xij;
    int32_t* phaseNum;
    uint8_t* phaseExist;
    double *Nt, *vf, *vfP, *vfi;
    double *rho, *xi, *mu, *rhoP, *xiP, *muP;
    double *rhox, *xix, *mux;
    uint8_t* pSderExist;
    int32_t *pVnumCom, *bRowSizedSdP;
    double* dSec;  // maxLendSdP entries per cell
    uint8_t* skipFlag;
    double *skipMinEig, *skipP, *skipT, *skipZi;
    double *kr, *Pc, *Pj, *dKr_dS, *dPcj_dS;
    double *v, *depth, *poroInit, *poro, *poroP, *rockVp;
    int32_t *ftype, *iters, *nonconv;
    // IsoT_FIMn only (NULL otherwise): phase moles nj (np), the flash's
    // secondary-variable correction res_n (np * (nc + 1) rows, the first
    // bRowSizedSdP used) and its volume-balance term resPc
    double *nj, *resN, *resPc;
    // serving pool: per cell 1 when its realisation takes part in the
    // iteration (NULL: every cell does)
    const uint8_t* mask;
};

__device__ __forceinline__ bool cellOn(const DevBulk& B, int64_t n) { return !B.mask || B.mask[n]; }

// Scratch of the staged flash (cells in SoA with the bulk's ld; lists of
// cell indices, their lengths in count[0, kFlashLists) and their work-queue
// cursors in count[kFlashLists, 2 * kFlashLists)).
constexpr int kFlashLists = 17;

// The staged flash's cell lists.
enum {
    LIST_STAB = 0,     // ftype 0: stability test
    LIST_SPLIT = 1,    // split from the start
    LIST_NP1 = 2,      // finished with one hydrocarbon phase
    LIST_NP2 = 3,      // finished with two
    LIST_C1 = 4,       // split continues after the first slice
    LIST_C2 = 5,       // ... after the second
    LIST_NR = 6,       // successive substitution done: Newton stage
    LIST_STABG = 7,    // stability test trying a learned guess first
    LIST_SPLITB = 8,   // side stream B: unstable cells of its stability test
    LIST_TSSM = 9,     // side stream C: split slices' leftovers
    LIST_TNR = 10,     // side stream D: cells whose first Newton split failed
    LIST_NP1B = 11,    // side stream B's cells with one hydrocarbon phase
    LIST_NP2B = 12,    // ... with two
    LIST_NP1C = 13,    // side stream C's
    LIST_NP2C = 14,
    LIST_NP1D = 15,    // side stream D's
    LIST_NP2D = 16,
};
static_assert(LIST_NP2D < kFlashLists, "FlashWork holds too few lists");

// The side streams' finished cells (device lists and the device addresses of
// their lengths), for the rows recomputed after the side streams join.
constexpr int kSideLists = 6;
struct CellLists {
    const int32_t* list[kSideLists];
    const int32_t* count[kSideLists];
};

// Cell i of the concatenated lists (i < cellTotal).
__device__ __forceinline__ int64_t cellTotal(const CellLists& L)
{
    int64_t t = 0;
#pragma unroll
    for (int q = 0; q < kSideLists; ++q) t += *L.count[q];
    return t;
}
__device__ __forceinline__ int32_t cellAt(const CellLists& L, int64_t i)
{
#pragma unroll
    for (int q = 0; q < kSideLists; ++q) {
        const int64_t c = *L.count[q];
        if (i < c) return L.list[q][i];
        i -= c;
    }
    return 0;
}

struct FlashWork {
    double* zi;      // NC per cell: feed mole fractions
    double* Nh;      // hydrocarbon moles
    double* Yt;      // sum of the stability trial phase (PhaseSplit input)
    double* x;       // 2*NC: split compositions
    double* n;       // 2*NC: split moles (fractions of Nh)
    double* nu;      // 2: phase fractions
    double* Ks;      // NC: K-values of the split's successive substitution
    double* Se;      // its last residual
    int32_t* ssmIt;  // its iterations so far
    int32_t* rrIt;   // Rachford-Rice iterations so far
    int32_t* nrIt;   // the split's Newton iterations so far (a cell whose restarts are deferred)
    int8_t* ftype;
    int8_t* lNP;
    uint8_t* flags;
    int32_t* list[kFlashLists];
    int32_t* count;  // list lengths, then work-queue cursors
    // learned initial guess (NULL: none): per cell, ln K_i (NC values) and a
    // stability logit at guess[n * guessLd + ...]
    const float* guess;
    int guessLd;
};

// Totals gathered by the flash kernels (one atomic per wavefront).
struct DevFlashCounters {
    unsigned long long cells;
    unsigned long long ftype[3];
    unsigned long long twoPhase;
    unsigned long long iters[5];
    unsigned long long nonconverged;
};

}  // namespace ocpg
