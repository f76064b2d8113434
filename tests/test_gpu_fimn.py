"""HIP IsoT_FIMn (ocpg_set_method(OCPG_METHOD_FIMN)) against the oracle and
the reference's golden vectors through the C-ABI: the FlashFIMn /
InitFlashFIMn outputs (derivatives by the phase moles, nj, res_n, resPc),
AssembleMatBulksNew / AssembleMatWellsNew with the right-hand side's
correction terms, and GetSolution; then a few Newton steps of the device
path against the oracle's, system by system.

Tolerances as for IsoT_FIM: integer fields exact; floating fields within
1e-10 per cell in at least 90 % of the cells and everywhere within ten
times the cell's rounding sensitivity (tests/gpu_util.py); Jacobian blocks
within 1e-10 of their largest entry in 99 % of the blocks."""
import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from oracle.oracle_py import Oracle
from tests import fim_case, gpu_util, parity

pytestmark = pytest.mark.gpu

DT = 5.0
CASES = ["flashn_spe5.npz", "flashn_ten.npz", "flashn_spe3.npz", "flashn_spe5large.npz",
         "flashn_limits_bench6.npz"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_fimn_flash_matches_oracle_and_reference(orc, name, phase):
    gpu_util.say(f"[fimn flash parity] {name} {phase}")
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, phase)
    ctx = lib.Context(eos)
    g, stats = gpu_util.gpu_flash(ctx, b, method=abi.METHOD_FIMN)
    o = b.copy()
    orc.flash_fimn(eos, o)
    ref = gpu_util.snapshot_dict(o, "orc_", fimn=True)
    sens = parity.rounding_sensitivity(lambda x: orc.flash_fimn(eos, x), b)
    rep = parity.compare_state(g, ref, "orc_", gpu_util.GPU_TOL, frac=gpu_util.GPU_FRAC,
                               loose=gpu_util.GPU_LOOSE, cell_tol=gpu_util.cell_bounds(sens))
    gpu_util.say("  vs oracle:", {k: v for k, v in rep.items() if k != "cells_beyond_tol"})
    assert rep["iters_mismatch_cells"] <= 0.1 * b.nb, rep["iters_mismatch_cells"]
    rep2 = parity.compare_state(g, d, f"{phase}_out_", 1e-8, frac=0.97, loose=1e-4)
    gpu_util.say("  vs reference: within 1e-8", rep2["frac_within_tol"])
    assert stats.two_phase == int(np.sum(g.phaseNum == 3))
    ctx.close()


def test_fimn_initflash_matches_oracle(orc):
    gpu_util.say("[fimn] InitFlashFIMn")
    d = parity.load("initflashn_spe5.npz")
    eos = parity.eos_from_fixture(parity.load("flash_spe5.npz"))
    nb = len(d["init_in_P"])
    b = abi.Bulk(nb, int(d["deck_nc"]))
    b.P[:] = d["init_in_P"]
    b.T[:] = d["init_in_T"]
    b.S[:] = d["init_in_S"]
    b.rockVp[:] = d["init_rockVp"]
    ctx = lib.Context(eos)
    g, _ = gpu_util.gpu_flash(ctx, b, init_z=d["init_z"], method=abi.METHOD_FIMN)
    o = b.copy()
    orc.init_flash_fimn(eos, o, d["init_z"])
    ref = gpu_util.snapshot_dict(o, "orc_", fimn=True)
    ref["orc_Ni"] = o.Ni.copy()
    rep = parity.compare_state(g, ref, "orc_", gpu_util.GPU_TOL, frac=gpu_util.GPU_FRAC,
                               loose=gpu_util.GPU_LOOSE)
    gpu_util.say("  vs oracle:", {k: v for k, v in rep.items() if k != "cells_beyond_tol"})
    err = parity.scaled_err(g.Ni.reshape(nb, -1), o.Ni.reshape(nb, -1))
    assert err.max() <= 1e-10
    # against the reference as the flash tests hold it (a cell whose split
    # takes another iteration count has another correction res_n)
    rep2 = parity.compare_state(g, d, "init_out_", 1e-8, frac=0.97, loose=1e-4)
    gpu_util.say("  vs reference: within 1e-8", rep2["frac_within_tol"])
    ctx.close()


def block_err(a, b, ncol):
    A = a.reshape(-1, ncol * ncol)
    Bm = b.reshape(-1, ncol * ncol)
    s = np.abs(Bm).max(axis=1)
    s = np.where(s > 0, s, 1.0)
    return np.abs(A - Bm).max(axis=1) / s


def oracle_system(case, b, wells=None):
    orc, grid, nc = case["orc"], case["grid"], case["NC"] + 1
    if wells is None:
        res, up, urho, _, _ = orc.calres(nc, grid, b, DT)
        IA, JA = orc.bsr_pattern(grid)
    else:
        res, up, urho, _, _ = orc.calres_wells(case["eos"], wells, grid, b, DT)
        IA, JA = orc.bsr_pattern_wells(grid, wells)
    val, rhs = orc.assemble_fimn(case["eos"], grid, b, DT, up, urho, IA, JA, res, wells=wells)
    return dict(res=res, rhs=rhs, IA=IA, JA=JA, val=val)


def fimn_context(case, b, wells=None):
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_method(abi.METHOD_FIMN)
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    if wells is not None:
        ctx.set_wells(wells)
    return ctx


def rhs_err(got, ref, res, ncol):
    """Per row: |got - ref| relative to the row's residual scale (the rhs is
    the residual plus the flash correction's terms)."""
    g = got.reshape(-1, ncol)
    r = ref.reshape(-1, ncol)
    s = np.maximum(np.abs(res.reshape(-1, ncol)).max(axis=1), np.abs(r).max(axis=1))
    s = np.where(s > 0, s, 1.0)
    return np.abs(g - r).max(axis=1) / s


@pytest.mark.parametrize("with_wells,deck", [(False, "spe5"), (True, "spe5"), (True, "spe3"), (True, "spe5large")])
def test_fimn_assembly_on_same_state(with_wells, deck):
    """The oracle's FIMn state uploaded: CalRes + AssembleMat (blocks and the
    right-hand side) against the oracle's on the same state; the 6-component
    fluid and the 9- and 7-component ones of the spe3 / spe5large decks."""
    gpu_util.say(f"[fimn assembly same state] wells {with_wells}, {deck}")
    case = fim_case.build_case(8, 6, 5, seed=17, method="fimn", **fim_case.deck_case_args(deck))
    b = case["bulk"]
    wells = fim_case.add_wells(case, seed=5) if with_wells else None
    ref = oracle_system(case, b, wells)
    ctx = fimn_context(case, b, wells)
    ctx.calres(DT)
    ctx.assemble(DT)
    IA, JA, val, rhs = ctx.get_bsr()
    assert np.array_equal(IA, ref["IA"]) and np.array_equal(JA, ref["JA"])
    ncol = case["NC"] + 2
    err = block_err(val, ref["val"], ncol)
    rerr = rhs_err(rhs, ref["rhs"], ref["res"], ncol)
    two = np.sum(b.phaseNum == 3)
    gpu_util.say("  two-phase cells", int(two), " block max err", err.max(), " rhs max err", rerr.max(),
                 " rhs terms nonzero", float(np.abs(ref["rhs"] - ref["res"]).max()))
    assert two > 0 and np.abs(ref["rhs"] - ref["res"]).max() > 0
    assert err.max() <= 1e-12
    assert rerr.max() <= 1e-12
    ctx.close()


@pytest.mark.parametrize("with_wells", [False, True])
def test_fimn_pipeline_from_primary_variables(with_wells):
    """GPU FlashFIMn + kr/Pc + rock + CalRes + AssembleMat from the same P,
    N and last-step state as the oracle's."""
    gpu_util.say(f"[fimn pipeline] wells {with_wells}")
    case = fim_case.build_case(8, 6, 5, seed=23, method="fimn")
    b = case["bulk"]
    orc = case["orc"]
    wells = fim_case.add_wells(case, seed=6) if with_wells else None
    o = b.copy()
    orc.flash_fimn(case["eos"], o)
    orc.calkrpc(case["sat"], o)
    orc.calrock(case["rock"], o)
    ref = oracle_system(case, o, wells)
    ctx = fimn_context(case, b, wells)
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    ctx.calres(DT)
    ctx.assemble(DT)
    IA, JA, val, rhs = ctx.get_bsr()
    g = b.copy()
    ctx.download(g)
    assert np.array_equal(g.phaseNum, o.phaseNum)
    ncol = case["NC"] + 2
    err = block_err(val, ref["val"], ncol)
    rerr = rhs_err(rhs, ref["rhs"], ref["res"], ncol)
    gpu_util.say("  rhs max err", rerr.max(), " block max err", err.max(),
                 " blocks within 1e-10", np.mean(err <= 1e-10))
    assert rerr.max() <= 1e-10
    assert np.mean(err <= 1e-10) >= 0.99 and err.max() <= 1e-6
    ctx.close()


@pytest.mark.parametrize("scale,dSmax", [(1.0, 0.2), (60.0, 0.02)])
def test_fimn_get_solution_matches_oracle(scale, dSmax):
    from tests.test_oracle_newton import random_update
    gpu_util.say(f"[fimn] GetSolution, update scale {scale}, dSmax {dSmax}")
    case = fim_case.build_case(8, 6, 5, seed=41, method="fimn")
    b, nc, orc = case["bulk"], case["NC"] + 1, case["orc"]
    u = random_update(b, nc, seed=9, scale=scale).ravel()
    ctx = fimn_context(case, b)
    scm = ctx.critical_saturations()
    dPmax, dNmax = ctx.get_solution(u, dSmax)
    g = b.copy()
    ctx.download(g)
    o = b.copy()
    book, (oPmax, oNmax) = orc.get_solution_n(nc, o, u, dSmax, scm)
    for f in ("P", "Ni", "S"):
        np.testing.assert_array_equal(getattr(g, f), getattr(o, f), f)
    # nj (dasum) and x_ij = n / nj: the same operations in the same order
    ex = o.phaseExist != 0
    np.testing.assert_allclose(g.nj[ex], o.nj[ex], rtol=1e-14)
    np.testing.assert_allclose(g.xij, o.xij, rtol=1e-13, atol=1e-300)
    assert dPmax == oPmax and dNmax == oNmax
    ctx.flash_fim(stats=False)
    ch = ctx.nr_change()
    f = b.copy()
    ctx.download(f)
    _, och = orc.nr_change(f, book)
    assert ch == och
    ctx.close()


def test_fimn_method_is_set_before_cells_and_pool_refused():
    case = fim_case.build_case(3, 3, 2, seed=2)
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(case["bulk"].nb)
    assert ctx.lib.ocpg_set_method(ctx.h, abi.METHOD_FIMN) == abi.ESTATE
    ctx.close()
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_method(abi.METHOD_FIMN)
    assert ctx.lib.ocpg_get_method(ctx.h) == abi.METHOD_FIMN
    assert ctx.lib.ocpg_pool_init(ctx.h, 100, 6) == abi.EUNSUP
    ctx.close()


def test_fimn_config1_ten_steps_match_the_oracle():
    """The 147-cell two-well case for 10 time steps with IsoT_FIMn under the
    host Newton loop: at every Newton iteration the device's FlashFIMn
    outputs against the oracle's from the same primary variables, and the
    device's system (blocks and right-hand side) against the oracle's
    assembly of the device's own state; then the same loop on the oracle
    alone takes the same steps with the same Newton counts and rollback."""
    from opencaeporo_amd import fim_loop
    from tests import config1_case
    from tests.oracle_backend import OracleBackend
    from tests.test_gpu_config1 import _device_wells, residual_scale
    gpu_util.say("[fimn config 1] 7x7x3, gas injector + oil producer, 10 FIM steps")
    probe = config1_case.build(method="fimn")
    ctx0 = config1_case.device_context(probe)
    xi_dev = ctx0.phase_density(14.7, 60.0, probe["wells"].injZi[:probe["NC"] + 1], 1)[0]
    ctx0.close()
    case = config1_case.build(inj_factor=1000.0 * xi_dev, method="fimn")
    ref_case = config1_case.build(inj_factor=1000.0 * xi_dev, method="fimn")
    nc = case["NC"] + 1
    ncol = nc + 1
    ctx = config1_case.device_context(case)
    dev = fim_loop.DeviceBackend(ctx, has_wells=True)
    orc = Oracle()
    stats = dict(iters=0, worst_state=0.0, worst_rhs=0.0, worst_block=0.0)
    pre = {}

    def observer(event, loop):
        b = case["bulk"]
        if event == "solution":
            pre["b"] = b.copy()
            ctx.download(pre["b"])
        elif event == "update":
            ob = pre["b"]
            orc.flash_fimn(case["eos"], ob)
            gb = b.copy()
            ctx.download(gb)
            assert np.array_equal(gb.phaseNum, ob.phaseNum), "phase counts differ"
            assert np.array_equal(gb.phaseExist, ob.phaseExist), "phase labels differ"
            for f in ("S", "xij", "rho", "xi", "mu", "vf", "dSec_dPri", "vfi"):
                e = parity.scaled_err(getattr(gb, f), getattr(ob, f)).max()
                stats["worst_state"] = max(stats["worst_state"], e)
                assert e <= 1e-10, (f, e)
        elif event == "system":
            gb = b.copy()
            ctx.download(gb)
            w = _device_wells(ctx, case["wells"])
            res, up, urho, _, _ = orc.calres_wells(case["eos"], w, case["grid"], gb, loop.dt)
            IA, JA = orc.bsr_pattern_wells(case["grid"], w)
            val_o, rhs_o = orc.assemble_fimn(case["eos"], case["grid"], gb, loop.dt, up, urho, IA, JA, res,
                                             wells=w)
            gIA, gJA, gval, grhs = loop.system
            assert np.array_equal(gIA, IA) and np.array_equal(gJA, JA)
            err = block_err(gval, val_o, ncol)
            stats["worst_block"] = max(stats["worst_block"], float(err.max()))
            assert err.max() <= 1e-10
            e = (np.abs(grhs - rhs_o) / residual_scale(gb, w, ncol)).max()
            stats["worst_rhs"] = max(stats["worst_rhs"], e)
            assert e <= 1e-10
            stats["iters"] += 1

    loop = fim_loop.FimLoop(dev, observer=observer)
    recs = loop.run(10)
    gpu_util.say("  device: dt", [round(r.dt, 4) for r in recs], "NR", [r.nr for r in recs],
                 "cuts", [r.cuts for r in recs])
    gpu_util.say(f"  lockstep over {stats['iters']} Newton systems: state {stats['worst_state']:.2e}, "
                 f"rhs {stats['worst_rhs']:.2e}, blocks {stats['worst_block']:.2e}")
    ob = OracleBackend(ref_case["eos"], ref_case["sat"], ref_case["rock"], ref_case["grid"], ref_case["bulk"],
                       ref_case["wells"], ref_case["setup"], method="fimn")
    orecs = fim_loop.FimLoop(ob).run(10)
    gpu_util.say("  oracle: dt", [round(r.dt, 4) for r in orecs], "NR", [r.nr for r in orecs],
                 "cuts", [r.cuts for r in orecs])
    assert [r.nr for r in recs] == [r.nr for r in orecs]
    assert [r.cuts for r in recs] == [r.cuts for r in orecs]
    np.testing.assert_allclose([r.dt for r in recs], [r.dt for r in orecs], rtol=1e-8)
    fin = case["bulk"].copy()
    ctx.download(fin)
    ofin = ref_case["bulk"]
    assert np.array_equal(fin.phaseNum, ofin.phaseNum)
    np.testing.assert_allclose(fin.P, ofin.P, rtol=1e-8)
    np.testing.assert_allclose(fin.Ni, ofin.Ni, rtol=1e-7, atol=1e-7 * np.abs(ofin.Ni).max())
    ctx.close()
