// Wavefront-cooperative successive substitution of PhaseSplit for the few
// cells whose split needs hundreds of iterations (src/MixtureComp.cpp:
// 1847-1959, 2030-2043, 1103-1145).
//
// One lane per cell runs such a cell's iterations one dependent instruction
// at a time: a lone wavefront on a SIMD, ~10 us per iteration, and the whole
// Newton iteration waits for it.  Here the 64 lanes of a wavefront share
// one cell: lane (half, comp) = (threadIdx & 63) >> 5, threadIdx & 31 owns
// phase `half` and component `comp`.  The per-component work (the
// Rachford-Rice terms and their divisions, the compositions, the mixing-rule
// products, each component's fugacity sum, exp and the K update) runs across
// lanes; every sum is then formed on all lanes from the lanes' terms in the
// reference's order (readlane of lane k, k ascending), so each value is
// bit-identical to the one-lane code in flash_device.hpp, and all lanes hold
// the full cell state for the code that stays one-lane (SplitNR,
// CheckSplit, redundant on every lane).
#pragma once

#include "flash_device.hpp"

namespace ocpg {

// Lane src's value of v (src uniform: a compile-time constant here).
__device__ __forceinline__ double laneValue(double v, int src)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int NC>
struct CoopSplit {
    FlashCell<NC>& F;
    const int half;    // the phase this lane's fugacity work is for
    const int comp;    // the component this lane owns
    const bool valid;  // comp < NC
    double Kown, zown, Aown, Bown;
    double xown[2];
    double phiOwn, fugOwn;  // phase `half`, component `comp`

    __device__ __forceinline__ explicit CoopSplit(FlashCell<NC>& f)
        : F(f), half((int)((threadIdx.x & 63) >> 5)), comp((int)(threadIdx.x & 31)), valid((threadIdx.x & 31) < NC)
    {
        zown = pick(F.zi);
        Aown = pick(F.Ai);
        Bown = pick(F.Bi);
        Kown = pick(F.Ks);
        xown[0] = xown[1] = 0;
        phiOwn = fugOwn = 0;
    }

    __device__ __forceinline__ double pick(const double* a) const
    {
        double v = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (k == comp) v = a[k];
        return v;
    }

    // RachfordRice2 + UpdateXRR (FlashCell::rachfordRice2).
    __device__ __forceinline__ void rachfordRice2()
    {
        double Kmin = F.Ks[0], Kmax = F.Ks[0];
#pragma unroll
        for (int i = 1; i < NC; ++i) {
            if (F.Ks[i] < Kmin) Kmin = F.Ks[i];
            if (F.Ks[i] > Kmax) Kmax = F.Ks[i];
        }
        const double numin = 1 / (1 - Kmax);
        const double numax = 1 / (1 - Kmin);
        double nu0 = 0.5 * (numin + numax);
        int iter = 0;
        const double tol = F.E.rrTol;
        const int maxIt = F.E.rrMaxIt;
        while (true) {
            double ta = 0, tb = 0;
            if (valid) {
                const double tmp = 1 + nu0 * (Kown - 1);
                ta = zown * (Kown - 1) / tmp;
                tb = zown * (Kown - 1) * (Kown - 1) / (tmp * tmp);
            }
            double rj = 0, J = 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                rj += laneValue(ta, k);
                J -= laneValue(tb, k);
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
        F.it[IT_RR] += iter;
        F.nu[0] = nu0;
        F.nu[1] = 1 - nu0;
        double x0 = 0, x1 = 0;
        if (valid) {
            double tmp = 1;
            tmp += F.nu[0] * (Kown - 1);
            x1 = zown / tmp;
            x0 = Kown * x1;
        }
        xown[0] = x0;
        xown[1] = x1;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            F.x[1][k] = laneValue(x1, k);
            F.x[0][k] = laneValue(x0, k);
        }
    }

    // CalFugPhiAll with two phases (FlashCell::calFugPhiAll): Aj, Bj, Zj of
    // both phases on every lane, phi and fug of (half, comp) in this lane.
    __device__ __forceinline__ void calFugPhiAll()
    {
        double xj[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) xj[k] = half ? F.x[1][k] : F.x[0][k];
        const double xi1 = half ? xown[1] : xown[0];
        // CalAjBj terms of row i1 = comp
        double bt = 0, dt = 0, ot[NC];
#pragma unroll
        for (int i2 = 0; i2 < NC; ++i2) ot[i2] = 0;
        if (valid) {
            bt = Bown * xi1;
            dt = xi1 * xi1 * Aown * F.omb(comp, comp);
#pragma unroll
            for (int i2 = 0; i2 < NC - 1; ++i2)
                if (i2 < comp)
                    ot[i2] = 2 * xi1 * xj[i2] * F.sqE(FlashCell<NC>::tri(comp, i2)) * F.omb(comp, i2);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double A = 0, Bv = 0;
#pragma unroll
            for (int i1 = 0; i1 < NC; ++i1) {
                Bv += laneValue(bt, 32 * j + i1);
                A += laneValue(dt, 32 * j + i1);
#pragma unroll
                for (int i2 = 0; i2 < i1; ++i2) A += laneValue(ot[i2], 32 * j + i1);
            }
            F.Aj[j] = A;
            F.Bj[j] = Bv;
        }
        const double aj = half ? F.Aj[1] : F.Aj[0];
        const double bj = half ? F.Bj[1] : F.Bj[0];
        const double zj = solEoS(aj, bj);
        F.Zj[0] = laneValue(zj, 0);
        F.Zj[1] = laneValue(zj, 32);
        const double lzb = log(zj - bj);
        const double lg = log((zj + kDelta1 * bj) / (zj + kDelta2 * bj));
        double ph = 0, fg = 0;
        if (valid) {
            double tmp = 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) tmp += 2 * F.omb(comp, k) * F.sq(comp, k) * xj[k];
            ph = exp(Bown / bj * (zj - 1) - lzb - aj / ((kDelta1 - kDelta2)) / bj * (tmp / aj - Bown / bj) * lg);
            fg = ph * xi1 * F.P;
        }
        phiOwn = ph;
        fugOwn = fg;
    }

    // One SplitSSM2 iteration after the fugacities: the residual sum and the
    // new K (FlashCell::splitSSM2's loop body).
    __device__ __forceinline__ double updateK()
    {
        const double f0 = __shfl(fugOwn, comp, 64), f1 = __shfl(fugOwn, 32 + comp, 64);
        const double p0 = __shfl(phiOwn, comp, 64), p1 = __shfl(phiOwn, 32 + comp, 64);
        double rr = 0, Kn = 0;
        if (valid) {
            const double r = f1 / f0 - 1;
            rr = r * r;
            Kn = p1 / p0;
        }
        double Se = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) Se += laneValue(rr, k);
#pragma unroll
        for (int k = 0; k < NC; ++k) F.Ks[k] = laneValue(Kn, k);
        Kown = Kn;
        return Se;
    }

    // phi and fug of both phases on every lane.
    __device__ __forceinline__ void publishFug()
    {
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                F.phi[j][k] = laneValue(phiOwn, 32 * j + k);
                F.fug[j][k] = laneValue(fugOwn, 32 * j + k);
            }
    }

    // FlashCell::splitSSM2Slice with no budget: SplitSSM2(false) resumed
    // from (Ks, Se, iter) to its end; then phi/fug of the last compositions
    // are on every lane.
    __device__ __forceinline__ void splitSSM2Rest(double& Se, int& iter)
    {
        double Stol = F.E.ssmSpTol2;
        int maxIt = F.E.ssmSpMaxIt;
        if (F.Yt < 1.1) {
            Stol *= 1E-1;
            maxIt *= 2;
        }
        Kown = pick(F.Ks);
        bool ran = false;
        while (Se > Stol) {
            rachfordRice2();
            calFugPhiAll();
            Se = updateK();
            ran = true;
            iter++;
            if (iter > maxIt) {
                F.ssmSpConflag = false;
                break;
            }
        }
        if (ran) publishFug();
        else F.calFugPhiAll();
    }

    // FlashCell::splitSSM2(true).
    __device__ __forceinline__ void splitSSM2Restart()
    {
        F.ssmSpConflag = true;
        double Se = 1;
        double Stol = F.E.ssmSpTol2;
        int maxIt = F.E.ssmSpMaxIt;
        Stol *= 1E-1;
        maxIt *= 2;
        if (F.Yt < 1.1) {
            Stol *= 1E-1;
            maxIt *= 2;
        }
        Kown = pick(F.Ks);
        int iter = 0;
        bool ran = false;
        while (Se > Stol) {
            rachfordRice2();
            calFugPhiAll();
            Se = updateK();
            ran = true;
            iter++;
            if (iter > maxIt) {
                F.ssmSpConflag = false;
                break;
            }
        }
        if (ran) publishFug();
        F.it[IT_SSMSP] += iter;
    }

    // FlashCell::splitRestarts with the cooperative successive substitution.
    __device__ __forceinline__ void splitRestarts()
    {
        int restarts = 0;
        while (!F.nrSpConflag) {
            splitSSM2Restart();
            F.splitNR();
            if (!F.checkSplit()) break;
            if (F.ssmSpConflag) break;
            if (++restarts >= kMaxSplitRestarts) break;
        }
        F.checkSplit();
    }
};

}  // namespace ocpg
