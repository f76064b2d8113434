// Per-cell Peng-Robinson flash with FIM derivatives, one GPU lane per cell.
//
// Computes what MixtureComp::FlashFIM (src/MixtureComp.cpp:494-573) and
// MixtureComp::InitFlashFIM (:261-353) compute for one cell, followed by
// IsoT_FIM::PassFlashValue (src/OCPFluidMethod.cpp:1122-1187): Michelsen
// stability (strict and relaxed SSM + Newton), the two-phase split (SSM with
// Rachford-Rice, then Newton on n_ij), the PR cubic, fugacities, LBC
// viscosity, phase identification, the implicit derivatives of the split,
// the compact d(S, x_ij)/d(P, N) block and the skip-stability memory.
// Everything is FP64 except the skip eigenvalue, which the reference takes
// from an FP32 LAPACK call.  Expressions keep the reference's operand order
// (the library is built without FP contraction) so that results track the
// reference's rounding.
//
// NC is a template parameter: every per-cell vector and matrix has a
// compile-time size and every loop over components unrolls, so the working
// set stays in VGPRs.
#pragma once

#include "dense_solve.hpp"
#include "device_state.hpp"
#include "flash_kernel.hpp"

namespace ocpg {

constexpr double kGasConst = 10.73159;     // include/OCPConst.hpp:37
constexpr double kPiTrunc = 3.141592653;   // include/OCPConst.hpp:39 (truncated)
constexpr double kCONV1 = 5.61458;         // include/OCPConst.hpp:59
constexpr double kCONV5 = 459.67;          // include/OCPConst.hpp:63
constexpr double kDelta1 = 2.41421356237;  // include/MixtureComp.hpp:350
constexpr double kDelta2 = -0.41421356237; // include/MixtureComp.hpp:351

__device__ __forceinline__ double signD(double d)
{
    return d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
}

// NTcubicroot, src/MixtureComp.cpp:5078-5102: Newton polish, keeps the best
// root seen within 10 steps.
__device__ __forceinline__ void ntCubicRoot(double& root, double a, double b, double c)
{
    double e = root * (root * (root + a) + b) + c;
    double optroot = root;
    double opte = fabs(e);
    int iter = 0;
    while (fabs(e) > 1E-8) {
        const double df = root * (3 * root + 2 * a) + b;
        root = root - e / df;
        iter++;
        if (iter > 10) break;
        e = root * (root * (root + a) + b) + c;
        if (fabs(e) <= opte) {
            opte = fabs(e);
            optroot = root;
        }
    }
    root = optroot;
}

// CubicRoot + SolEoS, src/MixtureComp.cpp:4999-5056 and :889-914.
__device__ __forceinline__ double solEoS(double aj, double bj)
{
    const double a = (kDelta1 + kDelta2 - 1) * bj - 1;
    const double b = (aj + kDelta1 * kDelta2 * bj * bj - (kDelta1 + kDelta2) * bj * (bj + 1));
    const double c = -(aj * bj + kDelta1 * kDelta2 * bj * bj * (bj + 1));
    const double Q = (a * a - 3 * b) / 9;
    const double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
    const double Q3 = Q * Q * Q;
    const double M = R * R - Q3;
    if (M <= 0) {
        const double theta = acos(R / sqrt(Q3));
        double z0 = -2 * sqrt(Q) * cos(theta / 3) - a / 3;
        double z1 = -2 * sqrt(Q) * cos((theta + 2 * kPiTrunc) / 3) - a / 3;
        double z2 = -2 * sqrt(Q) * cos((theta - 2 * kPiTrunc) / 3) - a / 3;
        ntCubicRoot(z0, a, b, c);
        ntCubicRoot(z1, a, b, c);
        ntCubicRoot(z2, a, b, c);
        // ascending order of the three roots
        double lo = fmin(fmin(z0, z1), z2);
        double hi = fmax(fmax(z0, z1), z2);
        const double zj1 = lo;
        const double zj2 = hi;
        const double dG = (zj2 - zj1) + log((zj1 - bj) / (zj2 - bj)) -
                          aj / (bj * (kDelta2 - kDelta1)) *
                              log((zj1 + kDelta1 * bj) * (zj2 + kDelta2 * bj) /
                                  ((zj1 + kDelta2 * bj) * (zj2 + kDelta1 * bj)));
        return dG > 0 ? zj1 : zj2;
    } else {
        const double tmp1 = -R + sqrt(M);
        const double tmp2 = R + sqrt(M);
        const double S = signD(tmp1) * pow(fabs(tmp1), 1.0 / 3);
        const double T = -signD(tmp2) * pow(fabs(tmp2), 1.0 / 3);
        double z = S + T - a / 3;
        ntCubicRoot(z, a, b, c);
        return z;
    }
}

// Common denominator/numerator of dZ/d(.) (Zx, Zn, Zp in the reference).
__device__ __forceinline__ double dZnum(double aj, double bj, double zj, double Ax, double Bx)
{
    return (bj - zj) * Ax +
           ((aj + kDelta1 * kDelta2 * (3 * bj * bj + 2 * bj)) +
            ((kDelta1 + kDelta2) * (2 * bj + 1) - 2 * kDelta1 * kDelta2 * bj) * zj -
            (kDelta1 + kDelta2 - 1) * zj * zj) *
               Bx;
}
__device__ __forceinline__ double dZden(double aj, double bj, double zj)
{
    return 3 * zj * zj + 2 * ((kDelta1 + kDelta2 - 1) * bj - 1) * zj +
           (aj + kDelta1 * kDelta2 * bj * bj - (kDelta1 + kDelta2) * bj * (bj + 1));
}

// Cap on PhaseSplit's SSM/NR restarts (none in the reference).
constexpr int kMaxSplitRestarts = 100;
// Successive-substitution iterations of a stability trial started from a
// learned K (an extra trial: when it does not prove instability within
// these, the reference's trials run as always).
constexpr int kGuidedTrialMaxIt = 20;

// Iteration counters, in the reference's order.
enum { IT_SSMSTA = 0, IT_NRSTA, IT_SSMSP, IT_NRSP, IT_RR };

template <int NC>
struct FlashCell {
    static constexpr int nc = NC + 1;          // with water
    static constexpr int ncol = nc + 1;        // block dimension
    static constexpr int NT = NC * (NC + 1) / 2;

    const DevEos& E;

    // input / basic state
    double P, T, Nh;
    double Ni[nc];
    double zi[NC];
    int lNP, ftype, NP;
    bool flagSkip;
    // EoS
    double Ai[NC], Bi[NC];
    // sqrt(Ai[i]*Ai[k]) for k <= i: NT values in LDS, this lane's entry t at
    // sqA[t * kFlashBlock] (the block's lanes interleaved, no bank conflicts)
    double* sqA;
    double Kw0[NC];  // Wilson K (Kw[0] of the reference)
    double Aj[2], Bj[2], Zj[2];
    double x[2][NC], phi[2][NC], fug[2][NC], n[2][NC];
    double nu[2];
    double Ks[NC], lKs[NC];
    double Yt;
    // stability scratch that survives between SSM and NR
    double Asta, Bsta, Zsta;
    // convergence bookkeeping (ePEC)
    bool ssmStaConflag, nrStaConflag, ssmSpConflag, nrSpConflag;
    // K-values proposed by the learned model (NULL: none), and whether a
    // trial started from them proved the feed unstable
    const double* Kguide = nullptr;
    bool guidedProof = false;
    int it[5];

    __device__ __forceinline__ explicit FlashCell(const DevEos& e, double* sqLds = nullptr) : E(e), sqA(sqLds) {}

    __device__ __forceinline__ static constexpr int tri(int i, int k) { return i * (i + 1) / 2 + k; }
    __device__ __forceinline__ double& sqE(int t) const { return sqA[t * kFlashBlock]; }
    __device__ __forceinline__ double sq(int i, int k) const { return i >= k ? sqE(tri(i, k)) : sqE(tri(k, i)); }
    __device__ __forceinline__ double omb(int i, int k) const { return E.oneMinusBIC[i * NC + k]; }

    // MixtureComp::CalAiBi, src/MixtureComp.cpp:916-937 (+ the sqrt(Ai*Ak)
    // products every mixing rule of the reference recomputes).
    __device__ __forceinline__ void calAiBi()
    {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double acf = E.Acf[i];
            double mwi;
            if (acf <= 0.49) {
                mwi = 0.37464 + 1.54226 * acf - 0.26992 * (acf * acf);
            } else {
                mwi = 0.379642 + 1.48503 * acf - 0.164423 * (acf * acf) + 0.016667 * pow(acf, 3.0);
            }
            const double Pri = P / E.Pc[i];
            const double Tri = T / E.Tc[i];
            const double g = (1 + mwi * (1 - sqrt(Tri)));
            Ai[i] = E.OmegaA[i] * Pri / (Tri * Tri) * (g * g);
            Bi[i] = E.OmegaB[i] * Pri / Tri;
        }
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int k = 0; k <= i; ++k) sqE(tri(i, k)) = sqrt(Ai[i] * Ai[k]);
    }

    // MixtureComp::CalAjBj, src/MixtureComp.cpp:955-969.
    __device__ __forceinline__ void calAjBj(double& AjT, double& BjT, const double* xj) const
    {
        AjT = 0;
        BjT = 0;
#pragma unroll
        for (int i1 = 0; i1 < NC; ++i1) {
            BjT += Bi[i1] * xj[i1];
            AjT += xj[i1] * xj[i1] * Ai[i1] * omb(i1, i1);
#pragma unroll
            for (int i2 = 0; i2 < i1; ++i2) AjT += 2 * xj[i1] * xj[i2] * sqE(tri(i1, i2)) * omb(i1, i2);
        }
    }

    // sum_k 2 (1 - BIC_ik) sqrt(Ai Ak) x_k, in the reference's order.
    __device__ __forceinline__ double aSum2(int i, const double* xj) const
    {
        double tmp = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) tmp += 2 * omb(i, k) * sq(i, k) * xj[k];
        return tmp;
    }
    // sum_m (1 - BIC_im) sqrt(Ai Am) x_m (CalFugNAll / CalFugPAll order).
    __device__ __forceinline__ double aSum1(int i, const double* xj) const
    {
        double tmp = 0;
#pragma unroll
        for (int m = 0; m < NC; ++m) tmp += omb(i, m) * sq(i, m) * xj[m];
        return tmp;
    }
    // sum_k x_k (1 - BIC_ik) sqrt(Ai Ak) (CalFugXSTA / CalXiPNX order).
    __device__ __forceinline__ double aSumX(int i, const double* xj) const
    {
        double tmp = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) tmp += xj[k] * omb(i, k) * sq(i, k);
        return tmp;
    }

    // MixtureComp::CalFugPhi, src/MixtureComp.cpp:993-1101: PR fugacity of a
    // trial phase; optionally the coefficients phi.  Sets Asta/Bsta/Zsta.
    template <bool WithPhi>
    __device__ __forceinline__ void calFugPhi(double* phiT, double* fugT, const double* xj)
    {
        double aj, bj;
        calAjBj(aj, bj, xj);
        const double zj = solEoS(aj, bj);
        const double lzb = log(zj - bj);
        const double lg = log((zj + kDelta1 * bj) / (zj + kDelta2 * bj));
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double tmp = aSum2(i, xj);
            const double ph = exp(Bi[i] / bj * (zj - 1) - lzb -
                                  aj / ((kDelta1 - kDelta2)) / bj * (tmp / aj - Bi[i] / bj) * lg);
            if (WithPhi) phiT[i] = ph;
            fugT[i] = ph * xj[i] * P;
        }
        Asta = aj;
        Bsta = bj;
        Zsta = zj;
    }

    // MixtureComp::CalFugPhiAll, src/MixtureComp.cpp:1103-1145.
    __device__ __forceinline__ void calFugPhiAll()
    {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= NP) break;
            calAjBj(Aj[j], Bj[j], x[j]);
            Zj[j] = solEoS(Aj[j], Bj[j]);
            const double aj = Aj[j], bj = Bj[j], zj = Zj[j];
            const double lzb = log(zj - bj);
            const double lg = log((zj + kDelta1 * bj) / (zj + kDelta2 * bj));
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double tmp = aSum2(i, x[j]);
                phi[j][i] = exp(Bi[i] / bj * (zj - 1) - lzb -
                                aj / ((kDelta1 - kDelta2)) / bj * (tmp / aj - Bi[i] / bj) * lg);
                fug[j][i] = phi[j][i] * x[j][i] * P;
            }
        }
    }

    // MixtureComp::CalKwilson, src/MixtureComp.cpp:1269-1277 (entry 0; the
    // other three trial vectors are formed when the strict test needs them).
    __device__ __forceinline__ void calKwilson()
    {
#pragma unroll
        for (int i = 0; i < NC; ++i)
            Kw0[i] = (E.Pc[i] / P) * exp(5.373 * (1 + E.Acf[i]) * (1 - E.Tc[i] / T));
    }
    __device__ __forceinline__ double kw(int k, int i) const
    {
        if (k == 0) return Kw0[i];
        if (k == 1) return 1 / Kw0[i];
        if (k == 2) return pow(Kw0[i], 1.0 / 3);
        return pow(1 / Kw0[i], 1.0 / 3);
    }

    // MixtureComp::StableNR, src/MixtureComp.cpp:1590-1636, with
    // CalFugXSTA (:1638-1714) and AssembleJmatSTA (:1716-1741) fused: only
    // the lower triangle that dsysv('U') reads is formed.
    __device__ __forceinline__ bool stableNR(const double* fugId, double* Y, double* fugSta)
    {
        double res[NC];
        double se2 = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            res[i] = log(fugId[i] / (fugSta[i] * Yt));
            se2 += res[i] * res[i];
        }
        double Se = sqrt(se2);
        const double Stol = E.nrStaTol;
        int iter = 0;
        while (Se > Stol) {
            SymPacked<NC> J;
            {
                const double aj = Asta, bj = Bsta, zj = Zsta;
                double Ax[NC], Zx[NC];
                const double den = dZden(aj, bj, zj);
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    Ax[i] = 2 * aSumX(i, Y);
                    Zx[i] = dZnum(aj, bj, zj, Ax[i], Bi[i]) / den;
                }
                const double G = (zj + kDelta1 * bj) / (zj + kDelta2 * bj);
                const double lG = log(G);
                const double m1Mm2 = kDelta1 - kDelta2;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double C = Y[i] * P / (zj - bj);
                    const double Ei = -aj / ((m1Mm2)*bj) * (Ax[i] / aj - Bi[i] / bj);
                    double fx[NC];
                    double tmp = 0;
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        const double aik = omb(i, k) * sq(i, k);
                        const double Cxk = ((zj - bj) * (i == k ? 1.0 : 0.0) - Y[i] * (Zx[k] - Bi[k])) * P /
                                           ((zj - bj) * (zj - bj));
                        const double Dxk = Bi[i] / bj * (Zx[k] - Bi[k] * (zj - 1) / bj);
                        double Exk = (2 * (aj / bj * Bi[k] * Bi[i] + bj * aik) - Ax[i] * Bi[k] - Ax[k] * Bi[i]) /
                                     (bj * bj);
                        Exk /= -(m1Mm2);
                        const double Gxk = (m1Mm2) / (zj + kDelta2 * bj) / (zj + kDelta2 * bj) *
                                           (zj * Bi[k] - Zx[k] * bj);
                        fx[k] = 1 / C * Cxk + Dxk + Exk * lG + Ei / G * Gxk;
                        tmp += Y[k] * fx[k];
                    }
#pragma unroll
                    for (int k = 0; k <= i; ++k) J.a[tri(i, k)] = (fx[k] - tmp + 1) / Yt;
                }
            }
            symIndefSolve<NC>(J, res);
#pragma unroll
            for (int i = 0; i < NC; ++i) Y[i] = Yt * Y[i];
#pragma unroll
            for (int i = 0; i < NC; ++i) Y[i] = Y[i] + 1.0 * res[i];
            Yt = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) Yt += Y[i];
            {
                const double s = 1 / Yt;
#pragma unroll
                for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
            }
            calFugPhi<false>(nullptr, fugSta, Y);
            se2 = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                res[i] = log(fugId[i] / (fugSta[i] * Yt));
                se2 += res[i] * res[i];
            }
            Se = sqrt(se2);
            iter++;
            if (iter > E.nrStaMaxIt) {
                it[IT_NRSTA] += iter;
                nrStaConflag = false;
                return false;
            }
        }
        it[IT_NRSTA] += iter;
        nrStaConflag = true;
        return true;
    }

    // One trial of StableSSM from the unnormalised trial composition Y:
    // successive substitution, then Newton (src/MixtureComp.cpp:1505-1585).
    // True when the trial proves the feed unstable.
    __device__ __forceinline__ bool stabilityTrial(double* Y, const double* di, int maxIt)
    {
        const double Stol = E.ssmStaTol2;
        const double eYt = E.ssmStaEYt;
        double phiSta[NC], fugSta[NC];
        Yt = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) Yt += Y[i];
        {
            const double s = 1 / Yt;
#pragma unroll
            for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
        }
        calFugPhi<true>(phiSta, fugSta, Y);
        double Se = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double l = log(fugSta[i] / fug[0][i] * Yt);
            Se += l * l;
        }
        int iter = 0;
        while (Se > Stol) {
            Yt = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                Y[i] = di[i] / phiSta[i];
                Yt += Y[i];
            }
            {
                const double s = 1 / Yt;
#pragma unroll
                for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
            }
            calFugPhi<true>(phiSta, fugSta, Y);
            Se = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double l = log(fugSta[i] / fug[0][i] * Yt);
                Se += l * l;
            }
            iter++;
            if (iter > maxIt) break;
        }
        it[IT_SSMSTA] += iter;
        const bool flag = stableNR(fug[0], Y, fugSta);
        return flag && Yt > 1 + eYt;
    }

    // MixtureComp::StableSSM (strict, 4 Wilson trials), src/MixtureComp.cpp:1492-1588.
    __device__ __forceinline__ bool stableSSM()
    {
        const double* xj = x[0];
        calFugPhi<true>(phi[0], fug[0], xj);
        double di[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) di[i] = phi[0][i] * xj[i];
        const double Stol = E.ssmStaTol2;
        const double eYt = E.ssmStaEYt;
        double Y[NC], phiSta[NC], fugSta[NC];
        for (int k = 0; k < 4; ++k) {
            Yt = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                Y[i] = xj[i] / kw(k, i);
                Yt += Y[i];
            }
            {
                const double s = 1 / Yt;
#pragma unroll
                for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
            }
            calFugPhi<true>(phiSta, fugSta, Y);
            double Se = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double l = log(fugSta[i] / fug[0][i] * Yt);
                Se += l * l;
            }
            int iter = 0;
            while (Se > Stol) {
                Yt = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    Y[i] = di[i] / phiSta[i];
                    Yt += Y[i];
                }
                {
                    const double s = 1 / Yt;
#pragma unroll
                    for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
                }
                calFugPhi<true>(phiSta, fugSta, Y);
                Se = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double l = log(fugSta[i] / fug[0][i] * Yt);
                    Se += l * l;
                }
                iter++;
                if (iter > E.ssmStaMaxIt) break;
            }
            it[IT_SSMSTA] += iter;
            const bool flag = stableNR(fug[0], Y, fugSta);
            if (flag && Yt > 1 + eYt) {
                ssmStaConflag = true;
                return false;
            }
        }
        return true;
    }

    // StableSSM with K-values a learned model proposes (Kguide): two extra
    // trials (x K and x / K) run first; a trial that proves instability
    // settles it, otherwise the reference's four trials decide as always.
    // A separate function from stableSSM so the plain flash keeps its own
    // code.
    __device__ __forceinline__ bool stableSSMGuided()
    {
        const double* xj = x[0];
        calFugPhi<true>(phi[0], fug[0], xj);
        double di[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) di[i] = phi[0][i] * xj[i];
        double Y[NC];
        if (Kguide) {
            for (int k = 0; k < 2; ++k) {
#pragma unroll
                for (int i = 0; i < NC; ++i) Y[i] = k == 0 ? xj[i] * Kguide[i] : xj[i] / Kguide[i];
                if (stabilityTrial(Y, di, kGuidedTrialMaxIt)) {
                    ssmStaConflag = true;
                    guidedProof = true;
                    return false;
                }
            }
        }
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int i = 0; i < NC; ++i) Y[i] = xj[i] / kw(k, i);
            if (stabilityTrial(Y, di, E.ssmStaMaxIt)) {
                ssmStaConflag = true;
                return false;
            }
        }
        return true;
    }

    // MixtureComp::StableSSM01 (relaxed, 2 trials, no Newton),
    // src/MixtureComp.cpp:1392-1490.
    __device__ __forceinline__ bool stableSSM01()
    {
        const double* xj = x[0];
        calFugPhi<true>(phi[0], fug[0], xj);
        const double Stol = E.ssmStaTol2;
        const double eYt = E.ssmStaEYt;
        const double Ktol = E.ssmStaKtol;
        const double dYtol = E.ssmStaDYtol;
        double Se = 0, Sk = 0;
        double Y[NC], di[NC], ks[NC], fugSta[NC];
        for (int k = 0; k < 2; ++k) {
#pragma unroll
            for (int i = 0; i < NC; ++i) ks[i] = kw(k, i);
            int iter = 0;
            bool flag = false;
            while (true) {
                Yt = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    Y[i] = xj[i] * ks[i];
                    Yt += Y[i];
                }
                {
                    const double s = 1 / Yt;
#pragma unroll
                    for (int i = 0; i < NC; ++i) Y[i] = s * Y[i];
                }
                if (iter > 0) {
                    double dY = 0;
#pragma unroll
                    for (int i = 0; i < NC; ++i) dY += (Y[i] - di[i]) * (Y[i] - di[i]);
                    if (dY < dYtol) {
                        flag = true;
                        break;
                    }
                }
                calFugPhi<false>(nullptr, fugSta, Y);
                Se = 0;
                Sk = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    di[i] = fug[0][i] / (fugSta[i] * Yt);
                    ks[i] *= di[i];
                    Se += (di[i] - 1) * (di[i] - 1);
                    const double lk = log(ks[i]);
                    Sk += lk * lk;
                }
                iter++;
                if (Se < Stol) {
                    flag = true;
                    break;
                }
                if (Sk < Ktol) {
                    flag = true;
                    break;
                }
                if (iter > E.ssmStaMaxIt) break;
#pragma unroll
                for (int i = 0; i < NC; ++i) di[i] = Y[i];
            }
            it[IT_SSMSTA] += iter;
            if (flag && Yt > 1 - 0.1 && Sk > 1) flagSkip = false;
            if (flag && Yt > 1 + eYt) {
                ssmStaConflag = true;
                // Ks keeps the last relaxed trial (never read by the split)
                return false;
            }
        }
        return true;
    }

    // MixtureComp::PhaseStable, src/MixtureComp.cpp:1355-1390 (NP = 1).
    template <bool Guided = false>
    __device__ __forceinline__ bool phaseStable()
    {
        ssmStaConflag = false;
        nrStaConflag = false;
        bool flag;
        int tmpNP = NP;
        if (lNP == 0) {
            if constexpr (Guided) flag = stableSSMGuided();
            else flag = stableSSM();
        } else {
            flag = stableSSM01();
            if (!flag) tmpNP++;
            if (tmpNP != lNP) {
                if constexpr (Guided) flag = stableSSMGuided();
                else flag = stableSSM();
                flagSkip = false;
            }
        }
        return flag;
    }

    // MixtureComp::RachfordRice2 + UpdateXRR, src/MixtureComp.cpp:1907-1959, 2030-2043.
    __device__ __forceinline__ void rachfordRice2()
    {
        double Kmin = Ks[0], Kmax = Ks[0];
#pragma unroll
        for (int i = 1; i < NC; ++i) {
            if (Ks[i] < Kmin) Kmin = Ks[i];
            if (Ks[i] > Kmax) Kmax = Ks[i];
        }
        const double numin = 1 / (1 - Kmax);
        const double numax = 1 / (1 - Kmin);
        double nu0 = 0.5 * (numin + numax);
        int iter = 0;
        const double tol = E.rrTol;
        const int maxIt = E.rrMaxIt;
        while (true) {
            double rj = 0, J = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double tmp = 1 + nu0 * (Ks[i] - 1);
                rj += zi[i] * (Ks[i] - 1) / tmp;
                J -= zi[i] * (Ks[i] - 1) * (Ks[i] - 1) / (tmp * tmp);
            }
            if (fabs(rj) < tol || iter > maxIt) break;
            const double dnuj = -rj / J;
            const double tmpnu = nu0 + dnuj;
            if (tmpnu < numax && tmpnu > numin) {
                nu0 = tmpnu;
            } else {
                nu0 = dnuj > 0 ? (nu0 + numax) / 2 : (nu0 + numin) / 2;
            }
            iter++;
        }
        it[IT_RR] += iter;
        nu[0] = nu0;
        nu[1] = 1 - nu0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            double tmp = 1;
            tmp += nu[0] * (Ks[i] - 1);
            x[1][i] = zi[i] / tmp;
            x[0][i] = Ks[i] * x[1][i];
        }
    }

    // MixtureComp::SplitSSM2, src/MixtureComp.cpp:1847-1903.
    __device__ __forceinline__ void splitSSM2(bool restart)
    {
        ssmSpConflag = true;
        double Se = 1;
        double Stol = E.ssmSpTol2;
        int maxIt = E.ssmSpMaxIt;
        if (!restart) {
            if (lNP == 2) {
#pragma unroll
                for (int i = 0; i < NC; ++i) Ks[i] = lKs[i];
            } else {
#pragma unroll
                for (int i = 0; i < NC; ++i) Ks[i] = Kw0[i];
            }
        } else {
            Stol *= 1E-1;
            maxIt *= 2;
        }
        if (Yt < 1.1) {
            Stol *= 1E-1;
            maxIt *= 2;
        }
        int iter = 0;
        while (Se > Stol) {
            rachfordRice2();
            calFugPhiAll();
            Se = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double r = fug[1][i] / fug[0][i] - 1;
                Se += r * r;
                Ks[i] = phi[1][i] / phi[0][i];
            }
            iter++;
            if (iter > maxIt) {
                ssmSpConflag = false;
                break;
            }
        }
        it[IT_SSMSP] += iter;
    }

    // SplitSSM2(false) run in slices: at most `budget` iterations per call
    // (budget < 0: no limit), resuming from (Ks, Se, iter).  The iterations
    // are those of splitSSM2, in the same order; returns true once the loop
    // has ended (converged, or past the iteration cap with ssmSpConflag
    // cleared).  The caller sets ssmSpConflag = true, Se = 1, iter = 0 and
    // the initial Ks before the first slice and adds iter to it[IT_SSMSP]
    // after the last.
    __device__ __forceinline__ bool splitSSM2Slice(double& Se, int& iter, int budget)
    {
        double Stol = E.ssmSpTol2;
        int maxIt = E.ssmSpMaxIt;
        if (Yt < 1.1) {
            Stol *= 1E-1;
            maxIt *= 2;
        }
        int steps = 0;
        while (Se > Stol) {
            if (steps == budget) return false;
            rachfordRice2();
            calFugPhiAll();
            Se = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double r = fug[1][i] / fug[0][i] - 1;
                Se += r * r;
                Ks[i] = phi[1][i] / phi[0][i];
            }
            iter++;
            steps++;
            if (iter > maxIt) {
                ssmSpConflag = false;
                break;
            }
        }
        return true;
    }

    // PhaseSplit after its first SplitSSM2 and first SplitNR: the restart
    // loop (which does not run when that Newton split converged) and
    // CheckSplit.  A restart depends only on Ks, the feed and the counters,
    // so a cell can leave after its first SplitNR and run this later.
    __device__ __forceinline__ void splitRestarts()
    {
        int restarts = 0;
        while (!nrSpConflag) {
            splitSSM2(true);
            splitNR();
            if (!checkSplit()) break;
            if (ssmSpConflag) break;
            if (++restarts >= kMaxSplitRestarts) break;
        }
        checkSplit();
    }

    // The rest of PhaseSplit after its first SplitSSM2: Newton, the restart
    // loop and CheckSplit.
    __device__ __forceinline__ void phaseSplitAfterSSM()
    {
        splitNR();
        splitRestarts();
    }

    // d ln f_ij / d n_kj of phase j (CalFugNAll, src/MixtureComp.cpp:2168-2250),
    // packed lower triangle (the matrix is symmetric), plus Zn.
    __device__ __forceinline__ void calFugN(int j, double* fugn, double* Znj) const
    {
        const double aj = Aj[j], bj = Bj[j], zj = Zj[j], nuj = nu[j];
        const double* xj = x[j];
        double An[NC], Bn[NC], tmpI[NC];
        const double den = dZden(aj, bj, zj);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            tmpI[i] = aSum1(i, xj);
            An[i] = 2 / nuj * (tmpI[i] - aj);
            Bn[i] = 1 / nuj * (Bi[i] - bj);
            Znj[i] = dZnum(aj, bj, zj, An[i], Bn[i]) / den;
        }
        const double G = (zj + kDelta1 * bj) / (zj + kDelta2 * bj);
        const double lG = log(G);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double C = xj[i] * P / (zj - bj);
            const double tmp = tmpI[i];
            const double Ei = -aj / ((kDelta1 - kDelta2) * bj) * (2 * tmp / aj - Bi[i] / bj);
#pragma unroll
            for (int k = 0; k <= i; ++k) {
                const double aik = omb(i, k) * sq(i, k);
                const double Cnk = P / (zj - bj) / (zj - bj) *
                                   ((zj - bj) / nuj * ((i == k ? 1.0 : 0.0) - xj[i]) - xj[i] * (Znj[k] - Bn[k]));
                const double Dnk = Bi[i] / bj * (Znj[k] - (Bi[k] - bj) * (zj - 1) / (nuj * bj));
                const double Gnk = (kDelta1 - kDelta2) / ((zj + kDelta2 * bj) * (zj + kDelta2 * bj)) *
                                   (Bn[k] * zj - Znj[k] * bj);
                double Enk = -1 / (kDelta1 - kDelta2) / (bj * bj) *
                             (2 * (bj * aik / nuj + Bn[k] * (Bi[i] * aj / bj - tmp)) - An[k] * Bi[i] -
                              aj * Bi[i] / nuj);
                Enk -= Ei / nuj;
                fugn[tri(i, k)] = 1 / C * Cnk + Dnk + Enk * lG + Ei / G * Gnk;
            }
        }
    }

    // MixtureComp::SplitNR (NP = 2), src/MixtureComp.cpp:2057-2156.
    __device__ __forceinline__ void splitNR()
    {
        nrSpConflag = false;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < NC; ++i) n[j][i] = nu[j] * x[j][i];
        double res[NC];
        double e2 = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            res[i] = log(fug[1][i] / fug[0][i]);
            e2 += res[i] * res[i];
        }
        double eNR = sqrt(e2);
        const double eNR0 = eNR;
        const double NRtol = E.nrSpTol;
        int iter = 0;
        while (eNR > NRtol) {
            double ln[2][NC];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < NC; ++i) ln[j][i] = n[j][i];
            SymPacked<NC> J;
            {
                double f0[NT], Zn[NC];
                calFugN(0, f0, Zn);
                calFugN(1, J.a, Zn);
#pragma unroll
                for (int t = 0; t < NT; ++t) J.a[t] = f0[t] + J.a[t];
            }
            symIndefSolve<NC>(J, res);
            // CalStepNRsp, src/MixtureComp.cpp:2301-2319
            double alpha = 1;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double tmp = n[0][i] + alpha * res[i];
                if (tmp < 0) alpha *= 0.9 * fabs(n[0][i] / res[i]);
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) n[1][i] = zi[i];
#pragma unroll
            for (int i = 0; i < NC; ++i) n[0][i] = n[0][i] + alpha * res[i];
#pragma unroll
            for (int i = 0; i < NC; ++i) n[1][i] = n[1][i] - n[0][i];
            nu[0] = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) nu[0] += n[0][i];
#pragma unroll
            for (int i = 0; i < NC; ++i) x[0][i] = fabs(n[0][i] / nu[0]);
            nu[1] = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) nu[1] += n[1][i];
#pragma unroll
            for (int i = 0; i < NC; ++i) x[1][i] = fabs(n[1][i] / nu[1]);
            calFugPhiAll();
            e2 = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                res[i] = log(fug[1][i] / fug[0][i]);
                e2 += res[i] * res[i];
            }
            eNR = sqrt(e2);
            iter++;
            if (eNR > eNR0 || iter > E.nrSpMaxIt) break;
            double en = 0;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                double s = 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    ln[j][i] = ln[j][i] - n[j][i];
                    s += ln[j][i] * ln[j][i];
                }
                en += sqrt(s);
            }
            if (en / (2 * NC) < 1E-8) {
                nrSpConflag = true;
                break;
            }
        }
        if (eNR < NRtol) nrSpConflag = true;
        it[IT_NRSP] += iter;
    }

    // MixtureComp::CheckSplit, src/MixtureComp.cpp:1784-1836 (the Gibbs-energy
    // comparison there only prints and is not evaluated).
    __device__ __forceinline__ bool checkSplit()
    {
        if (NP == 2) {
            double eX = 0;
            const double nuMax = nu[0] < nu[1] ? nu[1] : nu[0];
#pragma unroll
            for (int i = 0; i < NC; ++i) eX += (x[0][i] - x[1][i]) * (x[0][i] - x[1][i]);
            if (nuMax < 1 && nrSpConflag && isfinite(eX)) {
                // accepted
            } else if (!isfinite(eX) || 1 - nuMax < 1E-3) {
                NP = 1;
#pragma unroll
                for (int i = 0; i < NC; ++i) x[0][i] = zi[i];
                nu[0] = 1;
                calAjBj(Aj[0], Bj[0], x[0]);
                Zj[0] = solEoS(Aj[0], Bj[0]);
                ssmStaConflag = false;
                nrStaConflag = false;
                return false;
            }
        }
        return true;
    }

    // MixtureComp::PhaseSplit, src/MixtureComp.cpp:1743-1782.
    __device__ __forceinline__ void phaseSplit()
    {
        ssmSpConflag = false;
        nrSpConflag = false;
        splitSSM2(false);
        splitNR();
        // The reference loops until the Newton split converges, the split
        // collapses or a restarted SSM converges, with no cap; a cell where
        // none happens would spin forever.  kMaxSplitRestarts bounds it.
        int restarts = 0;
        while (!nrSpConflag) {
            splitSSM2(true);
            splitNR();
            if (!checkSplit()) break;
            if (ssmSpConflag) break;
            if (++restarts >= kMaxSplitRestarts) break;
        }
        checkSplit();
    }

    __device__ __forceinline__ void singlePhaseFeed()
    {
        NP = 1;
        nu[0] = 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) x[0][i] = zi[i];
        calAiBi();
        calAjBj(Aj[0], Bj[0], x[0]);
        Zj[0] = solEoS(Aj[0], Bj[0]);
    }

    // MixtureComp::PhaseEquilibrium, src/MixtureComp.cpp:1279-1353.
    // Returns true when the converged-flag bookkeeping would set ePEC = 1e8.
    __device__ __forceinline__ bool phaseEquilibrium()
    {
        if (ftype == 0) {
            flagSkip = true;
            singlePhaseFeed();
            calKwilson();
            if (!phaseStable()) {
                NP = 2;
                phaseSplit();
            }
            if (NP > 1) flagSkip = false;
            if (NP == 1) return !(nrStaConflag || ssmStaConflag);
            return !(nrSpConflag || ssmSpConflag);
        } else if (ftype == 1) {
            flagSkip = true;
            singlePhaseFeed();
            return false;
        } else {
            flagSkip = false;
            NP = 2;
            Yt = 1.01;
            calAiBi();
            calKwilson();
            phaseSplit();
            return !(nrSpConflag || ssmSpConflag);
        }
    }
};

}  // namespace ocpg
