"""Hydrostatic equilibration of a compositional deck (Bulk::InitPTSw,
src/Bulk.cpp:443-1064): the initial oil pressure and water saturation of
every cell from EQUIL (datum depth and pressure, the water-oil and gas-oil
contacts and their capillary pressures), the ZMFVD composition at depth,
RTEMP and the SWOF / SGOF tables.  Host code, run once per deck; the phase
densities come from the library (ocpg_phase_density, MixtureComp::RhoPhase)
through the caller's ``density(P, T, z, phase)``.

The phases' pressures are tabulated at ``tabrow`` equally spaced depths
spanning the cells: each phase's column is integrated up and down from the
datum row with its density at the row's pressure, and a phase other than
the datum's starts from its pressure at the datum depth, found by going to
the contact in the datum phase's gradient, stepping by the contact's
capillary pressure and coming back in its own (ten steps each way).  A
cell's pressure is the oil pressure interpolated at its depth, switched to
the water (or gas) pressure where the capillary pressures say the cell is
all water (or gas); its water saturation is the connate one without a
water-oil capillary pressure, else the mean over ten sub-depths of the
saturation the capillary pressure inverts to.

Table lookups keep the reference's search cursor (OCPTable, src/OCPTable.cpp:
46-200): a lookup starts from the row the previous one on the same table
stopped at.  On a column that is not strictly monotone (a capillary-pressure
column of zeros) the answer then depends on the lookups before it, so the
lookups here run on the same tables in the same order as the reference's.
"""
import numpy as np

GRAVITY_FACTOR = 0.00694444  # include/OCPConst.hpp:56, psi ft2 / lb
TINY = 1e-8                  # include/OCPConst.hpp:38
OIL, GAS, WATER = 0, 1, 2


class _Table:
    """A table (columns of rows) with a persistent search cursor."""

    def __init__(self, cols, cursor=None):
        self.c = [np.asarray(c, dtype=np.float64) for c in cols]
        self.n = len(self.c[0])
        self.at = self.n // 2 if cursor is None else cursor

    def _line(self, j, dj, r0, r1, val):
        k = (self.c[dj][r1] - self.c[dj][r0]) / (self.c[j][r1] - self.c[j][r0])
        return self.c[dj][r0] + k * (val - self.c[j][r0])

    def value(self, j, val, dj):
        """Column dj at val of increasing column j; flat outside."""
        x = self.c[j]
        if val >= x[self.at]:
            for i in range(self.at + 1, self.n):
                if val < x[i]:
                    self.at = i - 1
                    return self._line(j, dj, self.at, self.at + 1, val)
            return self.c[dj][-1]
        for i in range(self.at - 1, -1, -1):
            if val >= x[i]:
                self.at = i
                return self._line(j, dj, i, i + 1, val)
        return self.c[dj][0]

    def inverse(self, j, val, dj):
        """Column dj at val of decreasing column j; flat outside."""
        x = self.c[j]
        if val > x[self.at]:
            for i in range(self.at - 1, -1, -1):
                if val <= x[i]:
                    self.at = i
                    return self._line(j, dj, i, i + 1, val)
            return self.c[dj][0]
        for i in range(self.at + 1, self.n):
            if val >= x[i]:
                self.at = i
                return self._line(j, dj, i - 1, i, val)
        return self.c[dj][-1]

    def row(self, val):
        """Every column at val of column 0 (increasing)."""
        x = self.c[0]
        if val >= x[self.at]:
            hit = next((i for i in range(self.at + 1, self.n) if val < x[i]), None)
            if hit is None:
                return [c[-1] for c in self.c]
            self.at = hit - 1
        else:
            hit = next((i for i in range(self.at - 1, -1, -1) if val >= x[i]), None)
            if hit is None:
                return [c[0] for c in self.c]
            self.at = hit
        return [self._line(0, k, self.at, self.at + 1, val) for k in range(len(self.c))]


def init_ptsw(depth, dz, equil, zmfvd, rtemp, swof, sgof, density, tabrow=50, table=False):
    """Returns (P, Sw, z): per-cell oil pressure, water saturation and feed
    (nb x NC); with table=True also the depth table (Z, Po, Pg, Pw).
    equil: dict with Dref, Pref, DOWC, PcOW, DGOC, PcGO; zmfvd: rows (depth,
    z_1 .. z_NC); swof / sgof: rows (S, kr, kro, Pc); density(P, T, z,
    phase) -> lb/ft3 (phase 0 oil, 1 gas, 2 water)."""
    depth = np.asarray(depth, dtype=np.float64)
    dz = np.asarray(dz, dtype=np.float64)
    nb = len(depth)
    Dref, Pref = equil["Dref"], equil["Pref"]
    DOWC, PcOW, DGOC, PcGO = equil["DOWC"], equil["PcOW"], equil["DGOC"], equil["PcGO"]
    T = rtemp
    zt = _Table(np.asarray(zmfvd, dtype=np.float64).T)
    sw_t = _Table(np.asarray(swof, dtype=np.float64).T)
    sg_t = _Table(np.asarray(sgof, dtype=np.float64).T)
    # water saturation against the gas-water capillary pressure
    # Pcow(Sw) + Pcgo(1 - Sw) (FlowUnit_ODGW01::Generate_SWPCWG)
    pcgw = [sw_t.c[3][i] + sg_t.value(0, 1 - sw_t.c[0][i], 3) for i in range(sw_t.n)]
    swgw_t = _Table([sw_t.c[0], pcgw])

    def feed(d):
        return np.array(zt.row(d)[1:])

    def gamma(p, d, phase):
        return GRAVITY_FACTOR * density(p, T, feed(d), phase)

    # the depth table
    zmin = min(float(np.min(depth - dz / 2)), 1e8)
    zmax = max(float(np.max(depth + dz / 2)), 0.0)
    step = (zmax - zmin) / (tabrow - 1)
    Z = np.empty(tabrow)
    Z[0] = zmin
    for i in range(1, tabrow):
        Z[i] = Z[i - 1] + step
    if Dref <= Z[0]:
        b = 0
    elif Dref >= Z[-1]:
        b = tabrow - 1
    else:
        b = int(np.argmax(Z > Dref)) - 1
    cols = {OIL: np.zeros(tabrow), GAS: np.zeros(tabrow), WATER: np.zeros(tabrow)}

    def fill(phase, p_datum):
        col = cols[phase]
        col[b] = p_datum + gamma(p_datum, Dref, phase) * (Z[b] - Dref)
        for i in range(b, 0, -1):
            col[i - 1] = col[i] - gamma(col[i], Z[i], phase) * (Z[i] - Z[i - 1])
        for i in range(b, tabrow - 1):
            col[i + 1] = col[i] + gamma(col[i], Z[i], phase) * (Z[i + 1] - Z[i])

    def across(p0, phase0, phase1, contact, jump, frozen=False):
        """Phase1's pressure at Dref from phase0's p0 there, through the
        contact (phase0's density taken at p0 throughout when frozen)."""
        h = (contact - Dref) / 10
        p, d = p0, Dref
        for _ in range(10):
            p += gamma(p0 if frozen else p, d, phase0) * h
            d += h
        p += jump
        for _ in range(10):
            p -= gamma(p, d, phase1) * h
            d -= h
        return p

    if Dref < DGOC:
        fill(GAS, Pref)
        po = across(Pref, GAS, OIL, DGOC, -PcGO)
        fill(OIL, po)
        fill(WATER, across(po, OIL, WATER, DOWC, -PcOW, frozen=True))
    elif Dref > DOWC:
        fill(WATER, Pref)
        po = across(Pref, WATER, OIL, DOWC, PcOW)
        fill(OIL, po)
        fill(GAS, across(po, OIL, GAS, DGOC, PcGO))
    else:
        fill(OIL, Pref)
        fill(GAS, across(Pref, OIL, GAS, DGOC, PcGO))
        fill(WATER, across(Pref, OIL, WATER, DOWC, -PcOW))
    dp = _Table([Z, cols[OIL], cols[GAS], cols[WATER]], cursor=0)

    # water-oil capillary pressure present (the reference compares the value
    # at Sw = 1 + TINY itself, not its magnitude, with TINY)
    no_pcow = abs(sw_t.value(0, 0.0 - TINY, 3)) < TINY and sw_t.value(0, 1.0 + TINY, 3) < TINY
    swco = sw_t.c[0][0]
    P = np.empty(nb)
    Sw = np.empty(nb)
    z = np.empty((nb, len(zt.c) - 1))
    for n in range(nb):
        z[n] = feed(depth[n])
        _, po, pg, pw = dp.row(depth[n])
        pcow, pcgo = po - pw, pg - po
        sw = sw_t.inverse(3, pcow, 0)
        sg = sg_t.value(3, pcgo, 0)
        if sw + sg > 1:
            sw = swgw_t.inverse(1, pcow + pcgo, 0)
            sg = 1 - sw
        if 1 - sw < TINY:
            po = pw + sw_t.value(0, 1.0, 3)
        elif 1 - sg < TINY:
            po = pg - sg_t.value(0, 1.0, 3)
        elif 1 - sw - sg < TINY:
            po = pg - sg_t.value(0, sg, 3)
        P[n] = po
        if no_pcow:
            Sw[n] = swco
            continue
        acc = 0.0
        for k in range(10):
            _, po_k, pg_k, pw_k = dp.row(depth[n] + dz[n] / 10 * (k - 4.5))
            s = sw_t.inverse(3, po_k - pw_k, 0)
            if s + sg_t.value(3, pg_k - po_k, 0) > 1:
                s = swgw_t.inverse(1, (po_k - pw_k) + (pg_k - po_k), 0)
            acc += s
        Sw[n] = acc / 10
    if table:
        return P, Sw, z, (Z, cols[OIL], cols[GAS], cols[WATER])
    return P, Sw, z
