"""HIP per-step well preparation (AllWells::PrepareWell), well checks
(AllWells::CheckP), the rollback of the wells in ResetToLastTimeStep, the
bulk checks and CalMaxChange against the oracle through the C-ABI."""
import os

import numpy as np
import pytest

from opencaeporo_amd import abi, fluids, lib
from oracle.oracle_py import WellStepState
from tests import fim_case, gpu_util

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DT = 5.0


def well_set(case, seed=5):
    """A gas injector above its perforations on a rate control, a water
    injector below its perforations on BHP, an oil-rate producer and a BHP
    producer, on one column of the grid each."""
    b, NC = case["bulk"], case["NC"]
    nc = NC + 1
    grid = case["grid"]
    nxny = grid.nx * grid.ny
    col = lambda i: [i + k * nxny for k in range(grid.nz)]
    rng = np.random.default_rng(seed)
    gas = np.zeros(nc)
    gas[:3] = (0.77, 0.20, 0.03)
    water = np.zeros(nc)
    water[-1] = 1.0
    P = b.P
    specs, depth, temp, phase, init = [], [], [], [], []
    for kind, mode, cells, z, ph, dw in (
            (abi.WELL_INJ, abi.WELL_RATE, col(0), gas, 1, -30.0),
            (abi.WELL_INJ, abi.WELL_BHP, col(nxny - 1), water, 2, +40.0),
            (abi.WELL_PROD, abi.WELL_RATE, col(nxny // 2), None, 0, -15.0),
            (abi.WELL_PROD, abi.WELL_BHP, col(nxny // 3), None, 0, 0.0)):
        cells = np.asarray(cells, np.int32)
        inj = kind == abi.WELL_INJ
        bhp = float(P[cells].max() + 250.0) if inj else float(P[cells].min() - 250.0)
        s = dict(type=kind, mode=mode, perf=cells, bhp=bhp, ctrlBHP=bhp + (400.0 if inj else -400.0),
                 ctrlRate=300.0 if inj else 40.0, WI=rng.uniform(5.0, 40.0, len(cells)),
                 dG=np.zeros(len(cells)), prodPhaseWeight=[1.0, 0.0, 0.0])
        if inj:
            s["injZi"] = z
            s["perfXi"] = np.zeros(len(cells))
        specs.append(s)
        d0 = b.depth[cells]
        depth.append(d0.min() + dw if dw <= 0 else d0.max() + dw)
        temp.append(160.0)
        phase.append(ph if inj else 1)
        init.append(mode)
    wells = abi.Wells(nc, specs)
    setup = abi.WellSetup(depth, temp, phase, init)
    return wells, setup


def gpu_ctx(case, wells, setup):
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    b = case["bulk"]
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    ctx.set_wells(wells)
    ctx.set_well_setup(setup)
    return ctx


def compare_wells(ctx, wells, step, exact_dg=False):
    mode, bhp, dG, xi, mult = ctx.get_well_state()
    np.testing.assert_array_equal(mode, wells.mode)
    np.testing.assert_array_equal(bhp, wells.bhp)
    np.testing.assert_array_equal(mult, wells.perfMult)
    # the segment densities come from EoS roots (cbrt / acos / cos / log on
    # the GPU against glibc): the offsets agree to rounding
    scale = np.maximum(np.abs(wells.perfDG), 1.0)
    assert np.max(np.abs(dG - wells.perfDG) / scale) <= (0 if exact_dg else 1e-11)
    inj = np.concatenate([np.full(wells.perfStart[w + 1] - wells.perfStart[w], wells.type[w] == abi.WELL_INJ)
                          for w in range(wells.nw)])
    np.testing.assert_allclose(xi[inj], wells.perfXi[inj], rtol=1e-12)


@pytest.mark.parametrize("deck", ["spe5", "ten"])
def test_prepare_wells_matches_oracle(deck):
    gpu_util.say(f"[well prep] PrepareWell vs oracle ({deck})")
    d = fluids.spe5_deck() if deck == "spe5" else fluids.ten_deck()
    case = fim_case.build_case(6, 5, 4, seed=13, deck=d)
    wells, setup = well_set(case)
    b = case["bulk"]
    ctx = gpu_ctx(case, wells, setup)
    orc = case["orc"]
    step = WellStepState(wells)
    for call in range(2):
        # the second call starts each interval from the first call's offsets
        ctx.prepare_wells()
        orc.prepare_wells(case["eos"], wells, b, step, setup=setup)
        compare_wells(ctx, wells, step)
        # the wells' rates and rows with the prepared state
        norms = ctx.calres(DT)
        res_o, _, _, _, norms_o = orc.calres_wells(case["eos"], wells, case["grid"], b, DT)
        _, _, _, rhs = ctx.get_bsr()
        scale = np.abs(res_o).max()
        assert np.max(np.abs(rhs - res_o)) <= 1e-11 * scale
        np.testing.assert_allclose(norms.maxWellRelRes_mol, norms_o.maxWellRelRes_mol, rtol=1e-10)
    ctx.close()


def test_phase_density_matches_reference_goldens():
    gpu_util.say("[well prep] XiPhase / RhoPhase vs the reference")
    g = np.load(os.path.join(HERE, "golden", "wells_density.npz"))
    for tag, deck in (("spe5", fluids.spe5_deck()), ("ten", fluids.ten_deck())):
        eos = lib.eos_from_deck(deck)
        ctx = lib.Context(eos)
        for z, ph, P, T, xi, rho in list(zip(g[f"{tag}_dens_z"], g[f"{tag}_dens_phase"], g[f"{tag}_dens_P"],
                                             g[f"{tag}_dens_T"], g[f"{tag}_dens_xi"], g[f"{tag}_dens_rho"]))[::7]:
            gx, gr = ctx.phase_density(P, T, z, int(ph))
            assert abs(gx - xi) <= 1e-13 * xi and abs(gr - rho) <= 1e-13 * rho
        ctx.close()


def test_check_wells_matches_oracle():
    gpu_util.say("[well prep] CheckP / CheckCrossFlow vs oracle")
    case = fim_case.build_case(6, 5, 4, seed=17)
    b = case["bulk"]
    orc = case["orc"]
    wells, setup = well_set(case)
    # the rate injector above its BHP limit (switch), the BHP producer above
    # its cells' pressure (cross flow)
    wells.bhp[0] = wells.ctrlBHP[0] + 100.0
    cells3 = wells.perfCell[wells.perfStart[3]:wells.perfStart[4]]
    wells.bhp[3] = wells.ctrlBHP[3] = float(b.P[cells3].max()) + 30.0
    ctx = gpu_ctx(case, wells, setup)
    step = WellStepState(wells)
    ctx.calres(DT)   # CalFlux: the perforation pressures CheckP reads
    step.perfP[:wells.nperf] = wells.bhp[np.repeat(np.arange(wells.nw), np.diff(wells.perfStart))] + wells.perfDG
    for rnd in range(3):
        st_g = ctx.check_wells()
        st_o = orc.check_wells(wells, b, step)
        gpu_util.say(f"  round {rnd}: status gpu {st_g} oracle {st_o}")
        assert st_g == st_o
        mode, bhp, _, _, mult = ctx.get_well_state()
        np.testing.assert_array_equal(mode, wells.mode)
        np.testing.assert_array_equal(bhp, wells.bhp)
        np.testing.assert_array_equal(mult, wells.perfMult)
        ctx.calres(DT)
        step.perfP[:wells.nperf] = wells.bhp[np.repeat(np.arange(wells.nw), np.diff(wells.perfStart))] + \
            wells.perfDG
    ctx.close()


def test_reset_to_last_step_restores_wells():
    """update_last_step, a Newton update that moves the BHPs, reset: the
    state and the wells' BHP are the saved ones and the next CalRes is the
    oracle's on that state."""
    gpu_util.say("[well prep] ResetToLastTimeStep with wells")
    case = fim_case.build_case(6, 5, 4, seed=19)
    b = case["bulk"]
    orc = case["orc"]
    wells, setup = well_set(case)
    ctx = gpu_ctx(case, wells, setup)
    ctx.prepare_wells()
    step = WellStepState(wells)
    orc.prepare_wells(case["eos"], wells, b, step, setup=setup)
    ctx.calres(DT)
    ctx.update_last_step()
    saved = b.copy()
    ctx.download(saved)
    bhp0 = ctx.get_well_bhp().copy()
    rows = ctx.rows()
    ncol = case["NC"] + 2
    rng = np.random.default_rng(3)
    u = rng.uniform(-1.0, 1.0, rows * ncol) * 1e-3
    u[b.nb * ncol::ncol] = 25.0      # every open well's BHP moves by 25 psi
    ctx.get_solution(u, 0.2)
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    assert not np.array_equal(ctx.get_well_bhp(), bhp0)
    ctx.reset_to_last_step()
    ctx.calres(DT)
    after = b.copy()
    ctx.download(after)
    for f in ("P", "Ni", "S", "xij", "phaseNum", "skipFlag", "skipMinEig", "dSec_dPri", "kr", "Pj"):
        np.testing.assert_array_equal(getattr(after, f), getattr(saved, f), err_msg=f)
    # ResetBHP: bhp = lbhp, then a BHP control back at its target
    want = bhp0.copy()
    want[wells.mode == abi.WELL_BHP] = wells.ctrlBHP[wells.mode == abi.WELL_BHP]
    np.testing.assert_array_equal(ctx.get_well_bhp(), want)
    # CaldG on the restored state, then the residual
    step.lbhp = bhp0.copy()
    orc.prepare_wells(case["eos"], wells, saved, step, setup=setup, what=1)
    compare_wells(ctx, wells, step)
    res_o, _, _, _, _ = orc.calres_wells(case["eos"], wells, case["grid"], saved, DT)
    _, _, _, rhs = ctx.get_bsr()
    assert np.max(np.abs(rhs - res_o)) <= 1e-11 * np.abs(res_o).max()
    ctx.close()


def test_bulk_checks_and_max_change():
    gpu_util.say("[well prep] CheckNi / CheckP / CalMaxChange")
    case = fim_case.build_case(6, 5, 4, seed=23)
    b = case["bulk"]
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    assert ctx.check_bulk() == 0
    ctx.update_last_step()
    c = b.copy()
    rng = np.random.default_rng(1)
    c.P[:] = b.P + rng.uniform(-30, 30, b.nb)
    c.S[:] = b.S + rng.uniform(-0.01, 0.01, b.S.size)
    c.Ni[:] = b.Ni * (1 + rng.uniform(-0.05, 0.05, b.Ni.size))
    ctx.upload(c)
    m = ctx.max_change()
    assert m["dPmax"] == np.abs(c.P - b.P).max()
    assert m["dSmax"] == np.abs(c.S - b.S).max()
    # the upload of c also set the last-step moles to c.lNi (ocpg_bulk's lNi)
    top = np.abs(np.maximum(c.Ni, c.lNi))
    rel = np.where(top > 1e-8, np.abs(c.Ni - c.lNi) / np.where(top > 1e-8, top, 1), 0)
    assert m["dNmax"] == rel.max()
    assert m["eVmax"] == (np.abs(c.vf - c.rockVp) / c.rockVp).max()
    c.P[7] = -1.0
    ctx.upload(c)
    assert ctx.check_bulk() == abi.BULK_NEGATIVE_PRESSURE
    c.Ni[3 * (case["NC"] + 1) + 2] = -1e-6
    ctx.upload(c)
    assert ctx.check_bulk() == abi.BULK_NEGATIVE_COMPONENTS_MOLES
    ctx.close()
