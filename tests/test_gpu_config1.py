"""The 147-cell two-well case for 10 FIM time steps with the device path
under the host Newton loop (opencaeporo_amd.fim_loop), against the oracle:

* lockstep: at every Newton iteration the device's flash, kr/Pc, rock and
  residual are checked against the oracle's on the same primary variables
  (downloaded right after GetSolution), and the device's assembled system
  against the oracle's assembly of the device's own state;
* trajectories: the same loop on the oracle alone takes the same time
  steps with the same Newton iteration counts (including the rejected
  attempt and its rollback) and ends at the same state.
"""
import numpy as np
import pytest

from opencaeporo_amd import abi, fim_loop
from tests import config1_case, gpu_util, parity
from tests.oracle_backend import OracleBackend
from tests.test_gpu_fim import block_err

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10


def _oracle_system_on(case, be_w, b, dt):
    from oracle.oracle_py import Oracle
    orc = Oracle()
    res, up, urho, _, norms = orc.calres_wells(case["eos"], be_w, case["grid"], b, dt)
    IA, JA = orc.bsr_pattern_wells(case["grid"], be_w)
    val = orc.assemble_wells(case["eos"], be_w, case["grid"], b, dt, up, urho, IA, JA)
    return res, IA, JA, val, norms


def residual_scale(b, w, ncol):
    """Per residual entry the scale IsoT_FIM::CalRes measures it by (src/
    OCPFluidMethod.cpp:1362-1385): the pore volume for a volume row, the
    cell's moles for a component row; a well row by its target (rate) or
    BHP.  A converged residual is a small difference of terms of that size,
    so rounding-level state differences show on this scale, not on the
    residual's own."""
    nb = b.nb
    sc = np.empty((nb + w.open_count(), ncol))
    sc[:nb, 0] = b.rockVp
    sc[:nb, 1:] = b.Nt[:, None]
    for k, wi in enumerate([i for i in range(w.nw) if w.open[i]]):
        sc[nb + k, :] = max(abs(w.ctrlRate[wi]), abs(w.bhp[wi]), 1.0)
    return sc.ravel()


def _device_wells(ctx, wells):
    w = config1_case.copy_wells(wells)
    mode, bhp, dG, xi, mult = ctx.get_well_state()
    w.mode[:] = mode
    w.bhp[:] = bhp
    w.perfDG[:] = dG
    w.perfXi[:] = xi
    w.perfMult[:] = mult
    return w


def test_config1_ten_steps_match_the_oracle():
    gpu_util.say("[config 1] 7x7x3, gas injector + oil producer, 10 FIM steps")
    # the injector's surface factor from the device's XiPhase (the product's
    # SetupWellOpt), the same input for both loops
    probe = config1_case.build()
    ctx0 = config1_case.device_context(probe)
    xi_dev = ctx0.phase_density(14.7, 60.0, probe["wells"].injZi[:probe["NC"] + 1], 1)[0]
    ctx0.close()
    assert abs(1000.0 * xi_dev - probe["factor"]) <= 1e-13 * probe["factor"]
    case = config1_case.build(inj_factor=1000.0 * xi_dev)
    ref_case = config1_case.build(inj_factor=1000.0 * xi_dev)
    nc = case["NC"] + 1
    ncol = nc + 1
    ctx = config1_case.device_context(case)
    dev = fim_loop.DeviceBackend(ctx, has_wells=True)
    from oracle.oracle_py import Oracle
    orc = Oracle()
    stats = dict(iters=0, worst_state=0.0, worst_res=0.0, worst_block=0.0)
    pre = {}

    def observer(event, loop):
        b = case["bulk"]
        if event == "solution":
            pre["b"] = b.copy()
            ctx.download(pre["b"])
            pre["w"] = _device_wells(ctx, case["wells"])
        elif event == "update":
            # the oracle's property pass from the same primary variables
            ob = pre["b"]
            orc.flash_fim(case["eos"], ob)
            orc.calkrpc(case["sat"], ob)
            orc.calrock(case["rock"], ob)
            gb = b.copy()
            ctx.download(gb)
            assert np.array_equal(gb.phaseNum, ob.phaseNum), "phase counts differ"
            assert np.array_equal(gb.phaseExist, ob.phaseExist), "phase labels differ"
            for f in ("S", "xij", "rho", "xi", "mu", "kr", "Pj", "vf", "dSec_dPri"):
                e = parity.scaled_err(getattr(gb, f), getattr(ob, f)).max()
                stats["worst_state"] = max(stats["worst_state"], e)
                assert e <= STATE_TOL, (f, e)
            res_o, _, _, _, _ = _oracle_system_on(case, pre["w"], ob, loop.dt)
            _, _, _, rhs = ctx.get_bsr()
            e = (np.abs(rhs - res_o) / residual_scale(ob, pre["w"], ncol)).max()
            stats["worst_res"] = max(stats["worst_res"], e)
            assert e <= STATE_TOL
        elif event == "system":
            # the device's Jacobian against the oracle's assembly of the
            # device's own state (the same upwinding, rates and weights)
            gb = b.copy()
            ctx.download(gb)
            w = _device_wells(ctx, case["wells"])
            res_o, IA, JA, val_o, _ = _oracle_system_on(case, w, gb, loop.dt)
            gIA, gJA, gval, grhs = loop.system
            assert np.array_equal(gIA, IA) and np.array_equal(gJA, JA)
            err = block_err(gval, val_o, ncol)
            stats["worst_block"] = max(stats["worst_block"], float(err.max()))
            assert err.max() <= 1e-10
            stats["iters"] += 1

    loop = fim_loop.FimLoop(dev, observer=observer)
    recs = loop.run(10)
    gpu_util.say("  device: dt", [round(r.dt, 4) for r in recs], "NR", [r.nr for r in recs],
                 "cuts", [r.cuts for r in recs], "wasted", [r.wasted for r in recs])
    gpu_util.say(f"  lockstep over {stats['iters']} Newton systems: state {stats['worst_state']:.2e}, "
                 f"residual {stats['worst_res']:.2e}, blocks {stats['worst_block']:.2e}")
    # the same loop on the oracle alone
    ob = OracleBackend(ref_case["eos"], ref_case["sat"], ref_case["rock"], ref_case["grid"], ref_case["bulk"],
                       ref_case["wells"], ref_case["setup"])
    oloop = fim_loop.FimLoop(ob)
    orecs = oloop.run(10)
    gpu_util.say("  oracle: dt", [round(r.dt, 4) for r in orecs], "NR", [r.nr for r in orecs],
                 "cuts", [r.cuts for r in orecs])
    assert [r.nr for r in recs] == [r.nr for r in orecs]
    assert [r.cuts for r in recs] == [r.cuts for r in orecs]
    assert [r.wasted for r in recs] == [r.wasted for r in orecs]
    np.testing.assert_allclose([r.dt for r in recs], [r.dt for r in orecs], rtol=1e-8)
    assert sum(r.cuts for r in recs) >= 1          # a rejected attempt and its rollback
    fin = case["bulk"].copy()
    ctx.download(fin)
    ofin = ref_case["bulk"]
    assert np.array_equal(fin.phaseNum, ofin.phaseNum)
    np.testing.assert_allclose(fin.P, ofin.P, rtol=1e-8)
    np.testing.assert_allclose(fin.Ni, ofin.Ni, rtol=1e-7, atol=1e-7 * np.abs(ofin.Ni).max())
    np.testing.assert_allclose(ctx.get_well_bhp(), ob.w.bhp, rtol=1e-8)
    gpu_util.say(f"  Newton iterations {loop.iterNR_total} (+{loop.wasted} wasted) over {len(recs)} steps, "
                 f"{loop.time:.2f} days")
    ctx.close()
