"""Fluid, saturation-function and rock descriptions used by the workloads.

The 6-component fluid is the SPE5 fluid of the reference's examples
(examples/spe5/EGOIL.in: C1, C3, C6, C10, C15, C20 with TCRIT/PCRIT/ZCRIT/
MW/ACF/BIC and the solver keywords), with the SWOF/SGOF/PVTW/ROCK tables of
examples/spe5/spe5.data:67-124.  The 10-component fluid is a textbook
C1..C6 + four C7+ pseudo-components set, with seeded random BIPs.
"""
import numpy as np

from . import abi

# examples/spe5/EGOIL.in
SPE5 = dict(
    names=["C1", "C3", "C6", "C10", "C15", "C20"],
    Tc=[343.0, 665.7, 913.4, 1111.8, 1270.0, 1380.0],
    Pc=[667.8, 616.3, 436.9, 304.0, 200.0, 162.0],
    Zc=[0.290, 0.277, 0.264, 0.257, 0.245, 0.235],
    MW=[16.04, 44.10, 86.18, 149.29, 206.00, 282.00],
    Acf=[0.013, 0.1524, 0.3007, 0.4885, 0.6500, 0.8500],
    # strict lower triangle, row by row (BIC keyword)
    BIC_lower=[0.0,
               0.0, 0.0,
               0.0, 0.0, 0.0,
               0.05, 0.005, 0.0, 0.0,
               0.05, 0.005, 0.0, 0.0, 0.0],
)

# Default LBC coefficients (src/ParamReservoir.cpp:933-938).
LBC_DEFAULT = [0.1023, 0.023364, 0.058533, -0.040758, 0.0093324]

# Solver keywords of EGOIL.in: RR / SSMSTA / NRSTA / SSMSP / NRSP.
EOS_CONTROLS = dict(rr_maxit=30, rr_tol=1e-12, ssmsta_maxit=100, ssmsta_tol=1e-12,
                    ssmsta_eyt=1e-8, nrsta_maxit=55, nrsta_tol=1e-12, ssmsp_maxit=100,
                    ssmsp_tol=1e-6, nrsp_maxit=55, nrsp_tol=1e-12)

# PVTW of spe5.data: Pref, Bw, Cw, muw, Cvw.
PVTW_SPE5 = [[14.7, 1.00, 3.3e-6, 0.7, 0.0]]
# RHOW_STD (include/OCPConst.hpp) times the default water gravity 1.0.
STD_RHOW = 62.3664

SWOF_SPE5 = [
    [0.2, 0, 1, 0], [0.2899, 0.0022, 0.6769, 0], [0.3778, 0.018, 0.4153, 0],
    [0.4667, 0.0607, 0.2178, 0], [0.5556, 0.1438, 0.0835, 0], [0.6444, 0.2809, 0.0123, 0],
    [0.7, 0.4089, 0, 0], [0.7333, 0.4855, 0, 0], [0.8222, 0.7709, 0, 0],
    [0.9111, 1, 0, 0], [1, 1, 0, 0]]
SGOF_SPE5 = [
    [0.0, 0.0, 1.0, 0.0], [0.05, 0.0, 0.88, 0.0], [0.0889, 0.001, 0.7023, 0.0],
    [0.1778, 0.01, 0.4705, 0.0], [0.2667, 0.03, 0.2963, 0.0], [0.3556, 0.05, 0.1715, 0.0],
    [0.4444, 0.1, 0.0878, 0.0], [0.5333, 0.2, 0.037, 0.0], [0.6222, 0.35, 0.011, 0.0],
    [0.65, 0.39, 0.0, 0.0], [0.7111, 0.56, 0.0, 0.0], [0.8, 1.0, 0.0, 0.0]]
ROCK_SPE5 = dict(Pref=3990.3, cp1=5e-6, cp2=0.0)

# 10 components: C1, C2, C3, nC4, nC5, C6 and C7+ split into four
# pseudo-components (C7-C12, C13-C19, C20-C29, C30+).
TEN = dict(
    names=["C1", "C2", "C3", "C4", "C5", "C6", "P1", "P2", "P3", "P4"],
    Tc=[343.0, 549.8, 665.7, 765.3, 845.4, 913.4, 1050.0, 1200.0, 1350.0, 1500.0],
    Pc=[667.8, 707.8, 616.3, 550.7, 488.6, 436.9, 390.0, 280.0, 200.0, 140.0],
    Zc=[0.290, 0.285, 0.277, 0.274, 0.262, 0.264, 0.260, 0.250, 0.240, 0.230],
    MW=[16.04, 30.07, 44.10, 58.12, 72.15, 86.18, 120.0, 200.0, 320.0, 480.0],
    Acf=[0.013, 0.0986, 0.1524, 0.2010, 0.2539, 0.3007, 0.40, 0.60, 0.85, 1.10],
)

# The 9-component SPE3 gas condensate (examples/spe3/EGOIL.in: CO2, N2, C1,
# C2, C3, C4-6 and three C7+ fractions, with their own OMEGAA / OMEGAB),
# PVTW and RTEMP of examples/spe3/spe3.data:140-143, 59-61, and the ZMFVD
# feed (:145-148, the same at every depth).
SPE3 = dict(
    names=["CO2", "N2", "C1", "C2", "C3", "C4-6", "C7+1", "C7+2", "C7+3"],
    Tc=[548.46, 227.16, 343.08, 549.774, 665.64, 806.54054, 838.11282, 1058.03863, 1291.89071],
    Pc=[1071.33111, 492.31265, 667.7817, 708.34238, 618.69739, 514.92549, 410.74956, 247.56341, 160.41589],
    Zc=[0.27408, 0.29115, 0.28473, 0.28463, 0.27748, 0.27640, 0.26120, 0.22706, 0.20137],
    MW=[44.01, 28.013, 16.043, 30.07, 44.097, 66.86942, 107.77943, 198.56203, 335.1979],
    Acf=[0.225, 0.04, 0.013, 0.0986, 0.1524, 0.21575, 0.3123, 0.5567, 0.91692],
    OmegaA=[0.4572355, 0.4572355, 0.5340210, 0.4572355, 0.4572355, 0.4572355, 0.6373344, 0.6373344,
            0.6373344],
    OmegaB=[0.0777961, 0.0777961, 0.0777961, 0.0777961, 0.0777961, 0.0777961, 0.0872878, 0.0872878,
            0.0872878],
    BIC_lower=[-0.02,
               0.1, 0.036,
               0.13, 0.05, 0.0,
               0.135, 0.08, 0.0, 0.0,
               0.1277, 0.1002, 0.09281, 0.0, 0.0,
               0.1, 0.1, 0.130663, 0.006, 0.006, 0.0,
               0.1, 0.1, 0.130663, 0.006, 0.006, 0.0, 0.0,
               0.1, 0.1, 0.130663, 0.006, 0.006, 0.0, 0.0, 0.0],
    T_F=200.0,
    feed=[0.0121, 0.0194, 0.6599, 0.0869, 0.0591, 0.0967, 0.04745, 0.01515, 0.0033],
    pvtw=[[3550.0, 1.0, 3e-6, 0.31, 0.0]],
)

# The 7-component CO2 + SPE5 fluid of the large SPE5 variant
# (examples/spe5large/SPE5C400x400X6ECL-CO2.DATA:79-193: CNAMES CO2, C1, C3,
# C6, C10, C15, C20; RTEMP 160; ZI; PVTW :250-252).
SPE5L = dict(
    names=["CO2", "C1", "C3", "C6", "C10", "C15", "C20"],
    Tc=[547.57, 343.0, 665.7, 913.4, 1111.8, 1270.0, 1380.0],
    Pc=[1070.0, 667.8, 616.3, 436.9, 304.0, 200.0, 162.0],
    Zc=[0.2742, 0.290, 0.277, 0.264, 0.257, 0.245, 0.235],
    MW=[44.01, 16.04, 44.10, 86.18, 149.29, 206.00, 282.00],
    Acf=[0.225, 0.013, 0.1524, 0.3007, 0.4885, 0.6500, 0.8500],
    BIC_lower=[0.092,
               0.124, 0.0,
               0.1, 0.0, 0.0,
               0.115, 0.0, 0.0, 0.0,
               0.115, 0.05, 0.005, 0.0, 0.0,
               0.115, 0.05, 0.005, 0.0, 0.0, 0.0],
    T_F=160.0,
    feed=[0.05, 0.45, 0.03, 0.07, 0.2, 0.15, 0.05],
    pvtw=[[1000.0, 1.0099, 3e-6, 0.70, 0.0]],
)


def bench_bic_lower_6():
    """Benchmark BIPs for the 6-component fluid: zero except C1 with the heavy
    fractions (C15, C20) at 0.05."""
    NC = 6
    low = []
    for i in range(1, NC):
        for j in range(i):
            low.append(0.05 if (j == 0 and i >= 4) else 0.0)
    return low


def random_bic_lower(NC, seed, hi=0.15):
    rng = np.random.default_rng(seed)
    return list(rng.uniform(0.0, hi, NC * (NC - 1) // 2))


def make_deck(comp, bic_lower, controls=None, pvtw=None, std_rhow=STD_RHOW, lbc=None):
    """Fill an abi.EosDeck from a component table (ZCRIT given, no VCRIT;
    OMEGAA / OMEGAB when the table has them)."""
    d = abi.EosDeck()
    NC = len(comp["Tc"])
    d.nc = NC
    for i in range(NC):
        d.Tc[i] = comp["Tc"][i]
        d.Pc[i] = comp["Pc"][i]
        d.MW[i] = comp["MW"][i]
        d.Acf[i] = comp["Acf"][i]
        d.Zc[i] = comp["Zc"][i]
    d.has_Zc = 1
    for key in ("OmegaA", "OmegaB"):
        if key in comp:
            for i in range(NC):
                getattr(d, key)[i] = comp[key][i]
            setattr(d, "has_" + key, 1)
    for k, v in enumerate(bic_lower):
        d.BIC_lower[k] = v
    for k, v in enumerate(lbc or LBC_DEFAULT):
        d.LBCcoef[k] = v
    for k, v in (controls or EOS_CONTROLS).items():
        setattr(d, k, v)
    rows = pvtw or PVTW_SPE5
    d.pvtw_rows = len(rows)
    for r, row in enumerate(rows):
        for c in range(5):
            d.pvtw[r][c] = row[c]
    d.std_rhow = std_rhow
    return d


def make_satfunc(swof=None, sgof=None):
    s = abi.SatFunc()
    swof = swof or SWOF_SPE5
    sgof = sgof or SGOF_SPE5
    s.swof_rows = len(swof)
    s.sgof_rows = len(sgof)
    for r, row in enumerate(swof):
        for c in range(4):
            s.swof[r][c] = row[c]
    for r, row in enumerate(sgof):
        for c in range(4):
            s.sgof[r][c] = row[c]
    return s


def make_rock(Pref=ROCK_SPE5["Pref"], cp1=ROCK_SPE5["cp1"], cp2=ROCK_SPE5["cp2"]):
    r = abi.Rock()
    r.Pref, r.cp1, r.cp2 = Pref, cp1, cp2
    return r


def spe5_deck():
    return make_deck(SPE5, SPE5["BIC_lower"])


def bench6_deck():
    return make_deck(SPE5, bench_bic_lower_6())


def ten_deck(seed=20241203):
    return make_deck(TEN, random_bic_lower(10, seed))


def spe3_deck():
    """The 9-component SPE3 fluid (its T is SPE3["T_F"])."""
    return make_deck(SPE3, SPE3["BIC_lower"], pvtw=SPE3["pvtw"])


def spe5large_deck():
    """The 7-component CO2 + SPE5 fluid (its T is SPE5L["T_F"])."""
    return make_deck(SPE5L, SPE5L["BIC_lower"], pvtw=SPE5L["pvtw"])
