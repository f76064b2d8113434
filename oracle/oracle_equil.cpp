// TEST INFRASTRUCTURE (the oracle): CPU restatement of the compositional
// hydrostatic equilibration Bulk::InitPTSw (src/Bulk.cpp:443-1064) with
// FlowUnit_ODGW01's saturation lookups (include/FlowUnit.hpp:257-270,
// src/FlowUnit.cpp:351-367) and the reference's OCPTable searches
// (src/OCPTable.cpp:46-200), whose row cursor persists between calls: on a
// column that is not strictly monotone (a zero capillary-pressure column)
// the answer depends on the calls before it, and this restatement keeps the
// cursor exactly as the reference does.
#include <cmath>
#include <vector>

#include "oracle_flash.hpp"

namespace oracle {
namespace {

constexpr double kGravityFactor = 0.00694444;  // include/OCPConst.hpp:56
constexpr double kTiny = 1e-8;                 // include/OCPConst.hpp:38

// OCPTable with its search cursor: data[col][row].
struct Table {
    std::vector<std::vector<double>> d;
    int bId = 0;
    int rows() const { return (int)d[0].size(); }
    int cols() const { return (int)d.size(); }

    // OCPTable::Eval (:126-149)
    double Eval(int j, double val, int dj)
    {
        const int n = rows();
        if (val >= d[j][bId]) {
            for (int i = bId + 1; i < n; i++)
                if (val < d[j][i]) {
                    bId = i - 1;
                    const double k = (d[dj][bId + 1] - d[dj][bId]) / (d[j][bId + 1] - d[j][bId]);
                    return d[dj][bId] + k * (val - d[j][bId]);
                }
            return d[dj].back();
        }
        for (int i = bId - 1; i >= 0; i--)
            if (val >= d[j][i]) {
                bId = i;
                const double k = (d[dj][bId + 1] - d[dj][bId]) / (d[j][bId + 1] - d[j][bId]);
                return d[dj][bId] + k * (val - d[j][bId]);
            }
        return d[dj].front();
    }

    // OCPTable::Eval_Inv (:178-200): column j decreasing
    double EvalInv(int j, double val, int dj)
    {
        const int n = rows();
        if (val > d[j][bId]) {
            for (int i = bId - 1; i >= 0; i--)
                if (val <= d[j][i]) {
                    bId = i;
                    const double k = (d[dj][bId + 1] - d[dj][bId]) / (d[j][bId + 1] - d[j][bId]);
                    return d[dj][bId] + k * (val - d[j][bId]);
                }
            return d[dj].front();
        }
        for (int i = bId + 1; i < n; i++)
            if (val >= d[j][i]) {
                bId = i;
                const double k = (d[dj][bId] - d[dj][bId - 1]) / (d[j][bId] - d[j][bId - 1]);
                return d[dj][bId - 1] + k * (val - d[j][bId - 1]);
            }
        return d[dj].back();
    }

    // OCPTable::Eval_All on column 0 (:46-86): every column at val
    void EvalAll(double val, double* out)
    {
        const int n = rows(), m = cols();
        if (val >= d[0][bId]) {
            for (int i = bId + 1; i < n; i++)
                if (val < d[0][i]) {
                    bId = i - 1;
                    for (int k = 0; k < m; k++) {
                        const double s = (d[k][bId + 1] - d[k][bId]) / (d[0][bId + 1] - d[0][bId]);
                        out[k] = d[k][bId] + s * (val - d[0][bId]);
                    }
                    return;
                }
            for (int k = 0; k < m; k++) out[k] = d[k].back();
            return;
        }
        for (int i = bId - 1; i >= 0; i--)
            if (val >= d[0][i]) {
                bId = i;
                for (int k = 0; k < m; k++) {
                    const double s = (d[k][bId + 1] - d[k][bId]) / (d[0][bId + 1] - d[0][bId]);
                    out[k] = d[k][bId] + s * (val - d[0][bId]);
                }
                return;
            }
        for (int k = 0; k < m; k++) out[k] = d[k].front();
    }

    // OCPTable::Eval_All0 (:88-124): columns 1.. at val of column 0
    void EvalAll0(double val, double* out)
    {
        double tmp[64];
        EvalAll(val, tmp);
        for (int k = 1; k < cols(); k++) out[k - 1] = tmp[k];
    }
};

Table fromRows(const double (*rows)[4], int n, int ncol)
{
    Table t;
    t.d.assign(ncol, std::vector<double>(n));
    for (int r = 0; r < n; r++)
        for (int c = 0; c < ncol; c++) t.d[c][r] = rows[r][c];
    t.bId = n / 2;  // OCPTable::Setup
    return t;
}

}  // namespace

// Bulk::InitPTSw for a compositional model: P (the oil pressure at each
// cell's depth), the water saturation, the feed z per cell (ZMFVD at its
// depth) and T (RTEMP).  equil: Dref, Pref, DOWC, PcOW, DGOC, PcGO.
// zmfvd: nz rows of (depth, z_1 .. z_NC).  Returns 0.
int InitPTSw(const ocpg_eos& eos, const ocpg_satfunc& sat, int64_t nb, const double* depth, const double* dz,
             const double* equil, const double* zmfvd, int nzrows, double rtemp, int tabrow, double* P, double* Sw,
             double* zout)
{
    const int NC = eos.nc;
    Mixture mix(eos);
    auto rho = [&](double p, double t, const double* z, int phase) {
        double xi, r;
        mix.PhaseDensity(p, t, z, phase == 2 ? 2 : 1, &xi, &r);
        return r;
    };
    enum { OIL = 0, GAS = 1, WATER = 2 };
    // FlowUnit_ODGW01's tables (OCPTable::Setup leaves the cursor mid-table)
    Table SWOF = fromRows(sat.swof, sat.swof_rows, 4);
    Table SGOF = fromRows(sat.sgof, sat.sgof_rows, 4);
    // Generate_SWPCWG: Pcgw = Pcow + Pcgo(1 - Sw)
    Table SWPCGW;
    SWPCGW.d.assign(2, std::vector<double>(sat.swof_rows));
    for (int i = 0; i < sat.swof_rows; i++) {
        SWPCGW.d[0][i] = SWOF.d[0][i];
        SWPCGW.d[1][i] = SWOF.d[3][i] + SGOF.Eval(0, 1 - SWOF.d[0][i], 3);
    }
    SWPCGW.bId = sat.swof_rows / 2;  // OCPTable::SetRowCol
    Table ZT;
    ZT.d.assign(NC + 1, std::vector<double>(nzrows));
    for (int r = 0; r < nzrows; r++)
        for (int c = 0; c <= NC; c++) ZT.d[c][r] = zmfvd[r * (NC + 1) + c];
    ZT.bId = nzrows / 2;
    auto GetPcowBySw = [&](double s) { return SWOF.Eval(0, s, 3); };
    auto GetSwByPcow = [&](double pc) { return SWOF.EvalInv(3, pc, 0); };
    auto GetPcgoBySg = [&](double s) { return SGOF.Eval(0, s, 3); };
    auto GetSgByPcgo = [&](double pc) { return SGOF.Eval(3, pc, 0); };
    auto GetSwByPcgw = [&](double pc) { return SWPCGW.EvalInv(1, pc, 0); };

    const double Dref = equil[0], Pref = equil[1], DOWC = equil[2], PcOW = equil[3], DOGC = equil[4],
                 PcGO = equil[5];
    double Zmin = 1e8, Zmax = 0;
    for (int64_t n = 0; n < nb; n++) {
        const double t1 = depth[n] - dz[n] / 2, t2 = depth[n] + dz[n] / 2;
        Zmin = Zmin < t1 ? Zmin : t1;
        Zmax = Zmax > t2 ? Zmax : t2;
    }
    const double tabdz = (Zmax - Zmin) / (tabrow - 1);
    Table DP;
    DP.d.assign(4, std::vector<double>(tabrow, 0.0));
    std::vector<double>&Zt = DP.d[0], &Po = DP.d[1], &Pg = DP.d[2], &Pw = DP.d[3];
    Zt[0] = Zmin;
    for (int i = 1; i < tabrow; i++) Zt[i] = Zt[i - 1] + tabdz;
    int beginId = 0;
    if (Dref <= Zt[0]) {
        beginId = 0;
    } else if (Dref >= Zt[tabrow - 1]) {
        beginId = tabrow - 1;
    } else {
        int i = 0;
        while (i < tabrow && !(Zt[i] > Dref)) i++;
        beginId = i - 1;
    }
    std::vector<double> z(NC, 0.0);
    const double T = rtemp;
    const int mynum = 10;
    // the table column of phase ph from its value at beginId, up and down
    auto column = [&](std::vector<double>& col, int ph) {
        for (int id = beginId; id > 0; id--) {
            ZT.EvalAll0(Zt[id], z.data());
            const double g = kGravityFactor * rho(col[id], T, z.data(), ph);
            col[id - 1] = col[id] - g * (Zt[id] - Zt[id - 1]);
        }
        for (int id = beginId; id < tabrow - 1; id++) {
            ZT.EvalAll0(Zt[id], z.data());
            const double g = kGravityFactor * rho(col[id], T, z.data(), ph);
            col[id + 1] = col[id] + g * (Zt[id + 1] - Zt[id]);
        }
    };
    // phase ph's pressure at Dref from pRef: down (or up) to the contact in
    // phase ref's gradient, the capillary step, back in phase ph's.  In one
    // leg (gas zone, oil towards the water contact) the reference evaluates
    // the first half's density at pRef throughout (refFrozen).
    auto viaContact = [&](double pRef, int ref, int ph, double contact, double jump, bool refFrozen) {
        double p = pRef;
        const double mydz = (contact - Dref) / mynum;
        double myz = Dref;
        for (int i = 0; i < mynum; i++) {
            ZT.EvalAll0(myz, z.data());
            const double g = kGravityFactor * rho(refFrozen ? pRef : p, T, z.data(), ref);
            p += g * mydz;
            myz += mydz;
        }
        p += jump;
        for (int i = 0; i < mynum; i++) {
            ZT.EvalAll0(myz, z.data());
            const double g = kGravityFactor * rho(p, T, z.data(), ph);
            p -= g * mydz;
            myz -= mydz;
        }
        return p;
    };
    auto start = [&](std::vector<double>& col, double pAtDref, int ph) {
        ZT.EvalAll0(Dref, z.data());
        const double g = kGravityFactor * rho(pAtDref, T, z.data(), ph);
        col[beginId] = pAtDref + g * (Zt[beginId] - Dref);
        column(col, ph);
    };
    if (Dref < DOGC) {
        // gas zone: gas, then oil through the gas-oil contact, then water
        start(Pg, Pref, GAS);
        const double Poref = viaContact(Pref, GAS, OIL, DOGC, -PcGO, false);
        start(Po, Poref, OIL);
        const double Pwref = viaContact(Poref, OIL, WATER, DOWC, -PcOW, true);
        start(Pw, Pwref, WATER);
    } else if (Dref > DOWC) {
        // water zone: water, oil through the oil-water contact, then gas
        start(Pw, Pref, WATER);
        const double Poref = viaContact(Pref, WATER, OIL, DOWC, PcOW, false);
        start(Po, Poref, OIL);
        const double Pgref = viaContact(Poref, OIL, GAS, DOGC, PcGO, false);
        start(Pg, Pgref, GAS);
    } else {
        // oil zone: oil, gas through the gas-oil contact, water through the
        // oil-water contact
        start(Po, Pref, OIL);
        const double Pgref = viaContact(Pref, OIL, GAS, DOGC, PcGO, false);
        start(Pg, Pgref, GAS);
        const double Pwref = viaContact(Pref, OIL, WATER, DOWC, -PcOW, false);
        start(Pw, Pwref, WATER);
    }
    DP.bId = 0;
    // capillary pressure between water and oil considered (the reference's
    // test compares the second value itself, not its magnitude, with TINY)
    const bool flagPcow = !(std::fabs(GetPcowBySw(0.0 - kTiny)) < kTiny && GetPcowBySw(1.0 + kTiny) < kTiny);
    const double swco = SWOF.d[0][0];
    double dat[4];
    for (int64_t n = 0; n < nb; n++) {
        ZT.EvalAll0(depth[n], z.data());
        for (int i = 0; i < NC; i++) zout[n * NC + i] = z[i];
        DP.EvalAll(depth[n], dat);
        double po = dat[1];
        const double pg = dat[2], pw = dat[3];
        double Pcgo = pg - po, Pcow = po - pw;
        double sw = GetSwByPcow(Pcow);
        double sg = GetSgByPcgo(Pcgo);
        if (sw + sg > 1) {
            sw = GetSwByPcgw(Pcow + Pcgo);
            sg = 1 - sw;
        }
        if (1 - sw < kTiny) {
            po = pw + GetPcowBySw(1.0);
        } else if (1 - sg < kTiny) {
            po = pg - GetPcgoBySg(1.0);
        } else if (1 - sw - sg < kTiny) {
            po = pg - GetPcgoBySg(sg);
        }
        P[n] = po;
        if (!flagPcow) {
            Sw[n] = swco;
            continue;
        }
        const int ncut = 10;
        double swSum = 0;
        for (int k = 0; k < ncut; k++) {
            const double dep = depth[n] + dz[n] / ncut * (k - (ncut - 1) / 2.0);
            DP.EvalAll(dep, dat);
            const double pcow = dat[1] - dat[3], pcgo = dat[2] - dat[1];
            double tsw = GetSwByPcow(pcow);
            const double tsg = GetSgByPcgo(pcgo);
            if (tsw + tsg > 1) tsw = GetSwByPcgw(pcow + pcgo);
            swSum += tsw;
        }
        Sw[n] = swSum / ncut;
    }
    (void)T;
    return 0;
}

}  // namespace oracle
