"""IsoT_FIM's method set on the oracle (CPU restatement), for driving
opencaeporo_amd.fim_loop.FimLoop beside the device backend.  TEST
INFRASTRUCTURE: the oracle is the checker."""
import numpy as np

from opencaeporo_amd import abi
from oracle.oracle_py import Oracle, WellStepState


def critical_saturations(sat):
    """Bulk::satcm of FlowUnit_ODGW01 (src/FlowUnit.cpp:196-202): oil 0; gas
    and water the saturation of the last SGOF / SWOF row before kr reaches
    1e-8 (OCPTable::GetRowZero, src/OCPTable.cpp:35-44)."""
    def row_zero(tab, rows):
        for r in range(rows):
            if tab[r][1] >= 1e-8:
                return r - 1
        return rows - 1
    rg = row_zero(sat.sgof, sat.sgof_rows)
    rw = row_zero(sat.swof, sat.swof_rows)
    return np.array([0.0, sat.sgof[rg][0] if rg >= 0 else 0.0, sat.swof[rw][0] if rw >= 0 else 0.0])


class OracleBackend:
    """method "fimn": IsoT_FIMn's method set (FlashFIMn, its assembly and
    right-hand side, its GetSolution)."""

    def __init__(self, eos, sat, rock, grid, bulk, wells=None, setup=None, method="fim"):
        self.fimn = method == "fimn"
        self.scm = critical_saturations(sat)
        self.orc = Oracle()
        self.eos, self.sat, self.rock, self.grid = eos, sat, rock, grid
        self.b = bulk
        self.nc = bulk.nc
        self.ncol = self.nc + 1
        self.w = wells
        self.setup = setup
        self.step = WellStepState(wells) if wells is not None else None
        self.book = None
        self.last = bulk.copy()
        self.lbhp = wells.bhp.copy() if wells is not None else None
        if wells is not None:
            self.IA, self.JA = self.orc.bsr_pattern_wells(grid, wells)
        else:
            self.IA, self.JA = self.orc.bsr_pattern(grid)

    # AllWells::PrepareWell
    def prepare_wells(self):
        if self.w is not None:
            self.orc.prepare_wells(self.eos, self.w, self.b, self.step, setup=self.setup, what=0)

    def calres(self, dt):
        if self.w is not None:
            res, up, urho, _, n = self.orc.calres_wells(self.eos, self.w, self.grid, self.b, dt)
            # Well::CalFlux's perforation pressures (CheckP reads them)
            cnt = np.diff(self.w.perfStart)
            self.step.perfP[:self.w.nperf] = np.repeat(self.w.bhp, cnt) + self.w.perfDG
            self.step.perfTransj[:self.w.nperf * 3] = self.orc.well_flux(self.eos, self.w, self.b)[0]
        else:
            res, up, urho, _, n = self.orc.calres(self.nc, self.grid, self.b, dt)
        self.res, self.up, self.urho = res, up, urho
        return dict(V=n.maxRelRes_V, N=n.maxRelRes_N, well=n.maxWellRelRes_mol)

    def assemble(self, dt):
        if self.fimn:
            val, rhs = self.orc.assemble_fimn(self.eos, self.grid, self.b, dt, self.up, self.urho, self.IA, self.JA,
                                              self.res, wells=self.w)
            return self.IA, self.JA, val, rhs
        if self.w is not None:
            val = self.orc.assemble_wells(self.eos, self.w, self.grid, self.b, dt, self.up, self.urho, self.IA, self.JA)
        else:
            val = self.orc.assemble(self.nc, self.grid, self.b, dt, self.up, self.urho, self.IA, self.JA)
        return self.IA, self.JA, val, self.res.copy()

    def get_solution(self, u, dSmaxlim):
        if self.fimn:
            self.book, (dp, dn) = self.orc.get_solution_n(self.nc, self.b, u, dSmaxlim, self.scm)
        else:
            self.book, (dp, dn) = self.orc.get_solution(self.nc, self.b, u, dSmaxlim)
        if self.w is not None:
            nb = self.b.nb
            for w in range(self.w.nw):
                if self.w.open[w]:
                    self.w.bhp[w] += u[(nb + self.w.row_of(w)) * self.ncol]
        return dp, dn

    def check_bulk(self):
        # Bulk::CheckNi, then Bulk::CheckP
        if np.any(self.b.Ni < 0):
            return abi.BULK_NEGATIVE_COMPONENTS_MOLES
        if np.any(self.b.P < 0):
            return abi.BULK_NEGATIVE_PRESSURE
        return 0

    def update_property(self, dt):
        (self.orc.flash_fimn if self.fimn else self.orc.flash_fim)(self.eos, self.b)
        self.orc.calkrpc(self.sat, self.b)
        self.orc.calrock(self.rock, self.b)
        return self.calres(dt)

    def nr_change(self):
        _, out = self.orc.nr_change(self.b, self.book)
        return out["NRdSmax"]

    def check_wells(self):
        if self.w is None:
            return abi.WELL_SUCCESS
        return self.orc.check_wells(self.w, self.b, self.step)

    def reset_to_last_step(self):
        for k, v in self.last.arrays.items():
            self.b.arrays[k][:] = v
        if self.w is not None:
            self.step.lbhp = self.lbhp.copy()
            self.orc.prepare_wells(self.eos, self.w, self.b, self.step, setup=self.setup, what=1)

    def update_last_step(self):
        self.b.lNi[:] = self.b.Ni
        self.last = self.b.copy()
        if self.w is not None:
            self.lbhp = self.w.bhp.copy()

    def max_change(self):
        b, l = self.b, self.last
        top = np.abs(np.maximum(b.Ni, l.Ni))
        ok = top > 1e-8
        dN = np.abs(b.Ni - l.Ni)[ok] / top[ok]
        out = dict(dPmax=float(np.abs(b.P - l.P).max()), dSmax=float(np.abs(b.S - l.S).max()),
                   dNmax=float(dN.max()) if dN.size else 0.0,
                   eVmax=float((np.abs(b.vf - b.rockVp) / b.rockVp).max()), dBHPmax=0.0)
        if self.w is not None:
            op = self.w.open.astype(bool)
            out["dBHPmax"] = float(np.abs(self.w.bhp - self.lbhp)[op].max()) if op.any() else 0.0
        return out
