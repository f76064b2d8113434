// CPU restatement of OpenCAEPoro's compositional flash with FIM derivatives.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP kernels are checked
// against (and the CPU baseline the benchmark times).  Nothing in the product
// path may include, link or call it.
//
// It restates MixtureComp (src/MixtureComp.cpp, include/MixtureComp.hpp) and
// SkipStaAnaly (src/AcceleratePVT.cpp) for the FIM method, keeping the
// reference's arithmetic order, and the LAPACK routines the reference calls
// (dsysv, dgesv, ssyevd) by their published algorithms.
#ifndef ORACLE_FLASH_HPP
#define ORACLE_FLASH_HPP

#include "../include/ocpgpu.h"
#include "real.hpp"

namespace oracle {

constexpr int MAXNC = OCPG_MAXNC;
constexpr int NPMAX = 2;  // hydrocarbon phases
constexpr int NPH = 3;    // with water

// Constants of include/OCPConst.hpp:37-69 (values as the reference spells them).
constexpr double GAS_CONSTANT = 10.73159;
constexpr double PI_TRUNC = 3.141592653;
constexpr double CONV1 = 5.61458;
constexpr double CONV2 = 1.12712E-3;
constexpr double CONV5 = 459.67;
constexpr double GRAVITY_FACTOR = 0.00694444;
constexpr double TINY = 1E-8;

// Per-cell skip-stability memory (include/AcceleratePVT.hpp:71-85).
struct SkipState {
    bool flag;
    double minEigen;
    double P;
    double T;
    double zi[MAXNC];
};

class Mixture {
public:
    explicit Mixture(const ocpg_eos& eos);

    // MixtureComp::FlashFIM, src/MixtureComp.cpp:494-573.
    void FlashFIM(double Pin, double Tin, const double* Niin, const double* Sjin,
                  int lastNP, const double* xijin, SkipState& skip);
    // MixtureComp::InitFlashFIM, src/MixtureComp.cpp:261-353.
    void InitFlashFIM(double Pin, double Tin, const double* Sjin, double Vpore,
                      const double* Ziin, SkipState& skip);
    // MixtureComp::FlashFIMn, src/MixtureComp.cpp:575-695 (the IsoT_FIMn
    // flash: secondary variables S and n_ij).  The reference reads its
    // njin / phaseExistin arguments only for a branch it never takes
    // (`inputNP == 1 || OCP_TRUE`, :592), so they are not arguments here.
    void FlashFIMn(double Pin, double Tin, const double* Niin, const double* Sjin,
                   int lastNP, const double* xijin, SkipState& skip);
    // MixtureComp::InitFlashFIMn, src/MixtureComp.cpp:355-442.
    void InitFlashFIMn(double Pin, double Tin, const double* Sjin, double Vpore,
                       const double* Ziin, SkipState& skip);

    // MixtureComp::CalProdWeight, src/MixtureComp.cpp:830-853, on the
    // surface flash MixtureComp::Flash (:167-193).  A hydrocarbon phase the
    // flash does not produce contributes no volume (the reference reads what
    // its shared flash object last left in vj for it).
    double ProdWeight(double Pin, double Tin, const double* Niin, const double* prodPhase);
    // MixtureComp::Flash (src/MixtureComp.cpp:167-193): per phase whether it
    // exists, its volume, molar and mass density (zeros for a phase the flash
    // does not produce); returns Nt.
    double FlashPhases(double Pin, double Tin, const double* Niin, bool* ex, double* vjo, double* xio,
                       double* rhoo);
    // MixtureComp::CalProdRate (:855-865), absent phases contributing zero.
    void ProdRate(double Pin, double Tin, const double* Niin, double* rate);
    // xi, rho and LBC mu of one hydrocarbon phase of composition x (HC
    // mole fractions) at P, T (degF): the EoS root SolEoS picks, CalMW,
    // CalVfXiRho, CalViscoLBC on that phase alone.  Test helper (the finite
    // differences that pin the phase-property derivatives).
    void PhaseProps(double Pin, double Tin, const double* x, double* out3);
    // MixtureComp::XiPhase / RhoPhase (:711-768); phase 2 is water.
    void PhaseDensity(double Pin, double Tin, const double* Ziin, int phase, double* xio, double* rhoo);

    int NC, numCom, numPhase;
    // Outputs (Mixture base members, include/Mixture.hpp:241-292).
    real P, T;
    real vf, Nt, vfP;
    real Ni[MAXNC + 1];
    bool phaseExist[NPH];
    real S[NPH], vj[NPH], nj[NPH];
    real xij[NPH * (MAXNC + 1)];
    real rho[NPH], xi[NPH], mu[NPH];
    real vfi[MAXNC + 1];
    real rhoP[NPH], xiP[NPH], muP[NPH];
    real rhox[NPH * (MAXNC + 1)], xix[NPH * (MAXNC + 1)], mux[NPH * (MAXNC + 1)];
    real dXsdXp[(MAXNC + 2) * (NPH + NPH * (MAXNC + 1))];
    bool pSderExist[NPH];
    int pVnumCom[NPH];
    int ftype;
    real ePEC;
    // IsoT_FIMn outputs (MixtureComp::res / resPc, src/MixtureComp.cpp:
    // 4495-4632): the secondary variables' correction in dXsdXp's row order,
    // and its volume-balance term.
    real resN[NPH * (MAXNC + 2)];
    real resPc;
    // CaldXsdXp03's system as it was handed to dgesv (column-major, sysDim
    // rows, NC + 3 right-hand sides; sysDim = 0 for one hydrocarbon phase):
    // kept for the tests that measure its conditioning
    static constexpr int NSYS = NPH + 2 * MAXNC;
    real sysM[NSYS * NSYS], sysB[NSYS * (MAXNC + 3)];
    int sysDim = 0;
    long iters[5];  // SSMSTA, NRSTA, SSMSP, NRSP, RR for the last call

private:
    // constants
    real Tc[MAXNC], Pc[MAXNC], Vc[MAXNC], MWC[MAXNC], Acf[MAXNC];
    real OmegaA[MAXNC], OmegaB[MAXNC], Vshift[MAXNC], Vcvis[MAXNC];
    real BIC[MAXNC * MAXNC];
    real LBCcoef[5];
    real muAux1I[MAXNC], sqrtMWi[MAXNC];
    real delta1, delta2, delta1P2, delta1M2, delta1T2;
    int pvtwRows;
    double pvtw[OCPG_MAXTAB][5];
    real std_RhoW;
    // controls
    int ssmStaMaxIt, nrStaMaxIt, ssmSpMaxIt, nrSpMaxIt, rrMaxIt;
    real ssmStaTol, ssmStaTol2, ssmStaEYt, ssmStaKtol, ssmStaDYtol, nrStaTol;
    real ssmSpTol, ssmSpTol2, nrSpTol, rrTol;
    bool ssmStaConflag, nrStaConflag, ssmSpConflag, nrSpConflag;
    real ssmStaRealTol, nrStaRealTol, ssmSpRealTol, nrSpRealTol;
    int ssmStaCurIt, nrStaCurIt, ssmSpCurIt, nrSpCurIt, rrCurIt;

    // working state
    int lNP, NP;
    real Nh;
    real zi[MAXNC];
    real Ai[MAXNC], Bi[MAXNC], Aj[NPMAX], Bj[NPMAX], Zj[NPMAX];
    real Ztmp[3];
    real vC[NPMAX], nu[NPMAX], xiC[NPMAX], rhoC[NPMAX], MW[NPMAX];
    int phaseLabel[NPMAX];
    real x[NPMAX][MAXNC], phi[NPMAX][MAXNC], fug[NPMAX][MAXNC], n[NPMAX][MAXNC],
        ln[NPMAX][MAXNC];
    real Kw[4][MAXNC], Ks[NPMAX - 1][MAXNC], lKs[MAXNC];
    real Asta, Bsta, Zsta;
    real phiSta[MAXNC], fugSta[MAXNC];
    real Y[MAXNC], Yt, di[MAXNC];
    real resSTA[MAXNC], JmatSTA[MAXNC * MAXNC];
    real fugX[MAXNC * MAXNC];
    real Ax[MAXNC], Bx[MAXNC], Zx[MAXNC];
    real resSP[MAXNC], JmatSP[MAXNC * MAXNC];
    real fugN[NPMAX][MAXNC * MAXNC];
    real An[MAXNC], Bn[MAXNC], Zn[NPMAX][MAXNC];
    real muC[NPMAX], muAux[NPMAX][5];
    real fugP[NPMAX][MAXNC], Zp[NPMAX];
    real JmatDer[MAXNC * MAXNC], rhsDer[MAXNC * (MAXNC + 1)];
    real vjp[NPH], vji[NPH][MAXNC + 1];
    // CaldXsdXp03's system: dimension NP + 1 + NP * NC, NC + 3 right-hand sides
    static constexpr int NDER = NPH + NPMAX * MAXNC;
    real JmatDerN[NDER * NDER], JmatTmpN[NDER * NDER], rhsDerN[NDER * (MAXNC + 3)];
    real xixC[NPMAX * MAXNC], xiPC[NPMAX];
    bool flagSkip;
    real phiN[MAXNC * MAXNC];

    void InitPTN(double Pin, double Tin, const double* Niin);
    void InitPTZ(double Pin, double Tin, const double* Ziin);
    void CalFtypeFIM(const double* Sjin, const SkipState& skip);
    void CalFlash();
    void PhaseEquilibrium();
    void CalAiBi();
    void CalAjBj(real& AjT, real& BjT, const real* xj) const;
    void SolEoS(real& ZjT, const real& AjT, const real& BjT);
    int CubicRoot(const real& a, const real& b, const real& c, bool NTflag);
    void CalFugPhi(real* phiT, real* fugT, const real* xj);
    void CalFugPhiFugOnly(real* fugT, const real* xj);
    void CalFugPhiAll();
    void CalMW();
    void CalVfXiRho();
    void CalSaturation();
    int FindMWmax();
    void x2n();
    void CalKwilson();
    bool PhaseStable();
    bool CountedStable();
    void CountedSplit();
    bool StableSSM01(int Id);
    bool StableSSM(int Id);
    bool StableNR(int Id);
    void CalFugXSTA();
    void AssembleJmatSTA();
    bool CheckSplit();
    void PhaseSplit();
    void SplitSSM2(bool flag);
    void RachfordRice2();
    void UpdateXRR();
    void SplitNR();
    void CalResSP();
    void CalFugNAll(bool Znflag);
    void AssembleJmatSP();
    real CalStepNRsp();
    void IdentifyPhase();
    void CopyPhase();
    void CalViscoLBC();
    void CalFugPAll(bool Zpflag);
    void CalVfiVfp_full02();
    void CaldXsdXpAPI02p();
    void CalXiPNX_partial();
    void CalRhoPX_partial();
    void CalMuPXLBC_partial();
    void WaterProperties(bool init, const double* Sjin, double Vpore);
    void SetDerivFlags();
    void CalPhiNSkip();
    void CalSkipForNextStep(SkipState& skip);
    // IsoT_FIMn derivatives, src/MixtureComp.cpp:2581-2757, 4495-4853.
    void CalVjpVfpVfn_partial();
    void CaldXsdXpAPI03();
    void CaldXsdXp03();
    void CalXiPn_partial();
    void CalRhoPn_partial();
    void CalMuPnLBC_partial();
    void CalVfiVfp_full03();
    void PVTWEval(real Pv, real* data, real* cdata) const;
};

// LAPACK restatements (published algorithms of reference LAPACK).
// dsytf2+dsytrs with UPLO='U' on a column-major n x n array, one rhs.
int SymIndefSolve(int n, real* a, real* b);
// dgetf2+dgetrs on a column-major n x n array with nrhs columns.
int LUSolveColMajor(int n, int nrhs, real* a, real* b);
// Smallest eigenvalue of a symmetric float matrix (what ssyevd's first
// eigenvalue is), computed by cyclic Jacobi in double and rounded to float.
float MinEigenSym(int n, const float* a);

}  // namespace oracle

#endif
