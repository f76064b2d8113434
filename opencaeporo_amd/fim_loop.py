"""Host time stepping around the device path: the reference's Newton loop
with its time-step control, driving any backend with IsoT_FIM's method set.

Restated from the reference:

* Solver::GoOneStepIsoT (src/Solver.cpp:109-130): Prepare, then
  AssembleMat / SolveLinearSystem / UpdateProperty / FinishNR until the
  Newton iteration converges, then FinishStep;
* IsoT_FIM::Prepare (src/OCPFluidMethod.cpp:852-858), UpdateProperty
  (:877-897), FinishNR (:939-970), FinishStep (:972-978),
  ResetToLastTimeStep (:2859-2913);
* OCPControl::Check (src/OCPControl.cpp:242-290), CalNextTimeStep
  (:292-336), InitTime (:183-195), with the TUNING defaults of
  ParamControl::InitTuning (src/ParamControl.cpp:31-96) for FIM.

The backend supplies the device work (``DeviceBackend`` wraps a library
context); the linear solve is the caller's (FASP in the reference; a sparse
direct solve here by default).
"""
from dataclasses import dataclass, field

import numpy as np

from opencaeporo_amd import abi

MIN_TIME_CURSTEP = 1e-6   # include/OCPConst.hpp:46
TINY = 1e-8               # include/OCPConst.hpp:38


@dataclass
class Control:
    """OCPControl's ControlTime / ControlPreTime / ControlNR with the FIM
    TUNING defaults (src/ParamControl.cpp:38-92)."""
    timeInit: float = 1.0
    timeMax: float = 365.0
    timeMin: float = 0.1
    maxIncreFac: float = 3.0
    minChopFac: float = 0.15
    cutFacNR: float = 0.3
    dPlim: float = 300.0
    dSlim: float = 0.2
    maxNRiter: int = 10
    NRtol: float = 1e-3
    NRdPmax: float = 200.0
    NRdSmax: float = 0.2
    NRdPmin: float = 1.0
    NRdSmin: float = 0.01


@dataclass
class StepRecord:
    time: float            # end of the step (days)
    dt: float              # the step's accepted length
    nr: int                # Newton iterations of the accepted attempt
    wasted: int            # Newton iterations of rejected attempts
    cuts: int              # rejected attempts
    norms: dict = field(default_factory=dict)


def solve_bsr(IA, JA, val, rhs, ncol):
    """The block-CSR system LinearSystem hands to FASP, solved directly."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    nrow = len(IA) - 1
    A = sp.bsr_matrix((val.reshape(-1, ncol, ncol), JA, IA), shape=(nrow * ncol, nrow * ncol))
    return spla.spsolve(A.tocsc(), rhs)


class DeviceBackend:
    """IsoT_FIM's method set on a library context (lib.Context) whose cells,
    connections, wells and well setup are already set."""

    def __init__(self, ctx, has_wells):
        self.ctx = ctx
        self.has_wells = has_wells
        self.ncol = ctx.ncol

    def prepare_wells(self):
        if self.has_wells:
            self.ctx.prepare_wells()

    def calres(self, dt):
        n = self.ctx.calres(dt)
        return dict(V=n.maxRelRes_V, N=n.maxRelRes_N, well=n.maxWellRelRes_mol)

    def assemble(self, dt):
        self.ctx.assemble(dt)
        return self.ctx.get_bsr()

    def get_solution(self, u, dSmaxlim):
        return self.ctx.get_solution(u, dSmaxlim)

    def check_bulk(self):
        return self.ctx.check_bulk()

    def update_property(self, dt):
        self.ctx.flash_fim(stats=False)
        self.ctx.calkrpc_rock()
        return self.calres(dt)

    def nr_change(self):
        return self.ctx.nr_change()["NRdSmax"]

    def check_wells(self):
        return self.ctx.check_wells() if self.has_wells else abi.WELL_SUCCESS

    def reset_to_last_step(self):
        self.ctx.reset_to_last_step()

    def update_last_step(self):
        self.ctx.update_last_step()

    def max_change(self):
        return self.ctx.max_change()


class FimLoop:
    """Solver::GoOneStepIsoT + OCPControl on a backend.  ``observer``, when
    given, is called as observer(event, loop) at "solution" (after
    GetSolution, before the flash), "update" (after UpdateProperty) and
    "system" (after AssembleMat, with loop.system set)."""

    def __init__(self, backend, ctrl=None, solver=solve_bsr, end_time=1e30, observer=None):
        self.b = backend
        self.c = ctrl or Control()
        self.solver = solver
        self.end_time = end_time
        self.observer = observer
        self.time = 0.0
        # OCPControl::InitTime on the first step
        self.dt = min(end_time, self.c.timeInit)
        self.iterNR = 0
        self.iterNR_total = 0
        self.wasted = 0
        self.records = []
        self.system = None
        self.res0_V = 0.0

    def _event(self, name):
        if self.observer is not None:
            self.observer(name, self)

    def _calres(self, reset_res0):
        norms = self.b.calres(self.dt)
        if reset_res0:
            self.res0_V = norms["V"]
        return norms

    def _reset(self):
        # IsoT_FIM::ResetToLastTimeStep: state, wells, iteration counters,
        # then CalRes with the initial norm reset
        self.b.reset_to_last_step()
        self.wasted += self.iterNR
        self.iterNR = 0
        self._calres(True)

    def go_one_step(self):
        c = self.c
        cuts = 0
        wasted0 = self.wasted
        # IsoT_FIM::Prepare
        self.b.prepare_wells()
        self._calres(True)
        while True:
            if self.dt < MIN_TIME_CURSTEP:
                raise RuntimeError("time step size is too small")
            self.system = self.b.assemble(self.dt)
            self._event("system")
            IA, JA, val, rhs = self.system
            u = self.solver(IA, JA, val, rhs, self.b.ncol)
            self.iterNR += 1
            NRdPmax, _ = self.b.get_solution(u, c.NRdSmax)
            self._event("solution")
            # IsoT_FIM::UpdateProperty
            if self.b.check_bulk() != 0:
                self.dt *= c.cutFacNR
                self._reset()
                cuts += 1
                continue
            norms = self.b.update_property(self.dt)
            self.norms = norms
            self._event("update")
            # IsoT_FIM::FinishNR
            NRdSmax = self.b.nr_change()
            converged = ((norms["V"] <= self.res0_V * c.NRtol or norms["V"] <= c.NRtol or norms["N"] <= c.NRtol)
                         and norms["well"] <= c.NRtol) or \
                        (abs(NRdPmax) <= c.NRdPmin and abs(NRdSmax) <= c.NRdSmin)
            if converged:
                ws = self.b.check_wells()
                if ws != abi.WELL_SUCCESS:
                    if ws == abi.WELL_NEGATIVE_PRESSURE:
                        self.dt *= c.cutFacNR
                    self._reset()
                    cuts += 1
                    continue
                break
            if self.iterNR >= c.maxNRiter:
                self.dt *= c.cutFacNR
                self._reset()
                cuts += 1
                continue
        # IsoT_FIM::FinishStep
        mc = self.b.max_change()
        self.b.update_last_step()
        rec = StepRecord(time=self.time + self.dt, dt=self.dt, nr=self.iterNR, wasted=self.wasted - wasted0,
                         cuts=cuts, norms=dict(norms))
        self._next_time_step(mc)
        self.iterNR_total += self.iterNR
        self.iterNR = 0
        self.records.append(rec)
        return rec

    def _next_time_step(self, mc):
        """OCPControl::CalNextTimeStep(rs, {"dP", "dS", "iter"})."""
        c = self.c
        self.time += self.dt
        factor = c.maxIncreFac
        dPmax = max(mc["dPmax"], mc["dBHPmax"])
        if dPmax > TINY:
            factor = min(factor, c.dPlim / dPmax)
        if mc["dSmax"] > TINY:
            factor = min(factor, c.dSlim / mc["dSmax"])
        if self.iterNR < 5:
            factor = min(factor, 2.0)
        elif self.iterNR > 10:
            factor = min(factor, 0.5)
        else:
            factor = min(factor, 1.5)
        factor = max(c.minChopFac, factor)
        self.dt *= factor
        self.dt = min(self.dt, c.timeMax)
        self.dt = max(self.dt, c.timeMin)
        self.dt = min(self.dt, self.end_time - self.time)

    def run(self, steps):
        for _ in range(steps):
            self.go_one_step()
        return self.records
