"""Failure paths on the GPU, through the C-ABI: NaN cells end the flash
normally and are flagged while every other cell keeps its result; capped
iterations give the per-cell non-convergence flags and count the oracle
gives; a rolled-back step's next warm flash (restored skip memory) equals
the oracle's flash of the restored state."""
import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from oracle.oracle_py import Oracle
from tests import fim_case, gpu_util, parity

pytestmark = pytest.mark.gpu

FIELDS = ("phaseNum", "phaseExist", "S", "xij", "rho", "xi", "mu", "dSec_dPri", "skipFlag", "skipMinEig",
          "nonconv", "iters")


@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_nan_cells_are_flagged_and_contained(phase):
    gpu_util.say(f"[failure] NaN cells, {phase} pass")
    d = parity.load("flash_bench6.npz")
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, phase)
    ctx = lib.Context(eos)
    clean, st0 = gpu_util.gpu_flash(ctx, b)
    bad_in = b.copy()
    nc = b.nc
    bad_in.Ni.reshape(-1, nc)[5, 2] = np.nan
    bad_in.P[17] = np.nan
    bad, st1 = gpu_util.gpu_flash(ctx, bad_in)
    assert list(bad.nonconv[[5, 17]]) == [2, 2]
    assert np.all(clean.nonconv != 2)
    others = np.ones(b.nb, bool)
    others[[5, 17]] = False
    for f in FIELDS:
        a = getattr(bad, f).reshape(b.nb, -1)[others]
        c = getattr(clean, f).reshape(b.nb, -1)[others]
        assert np.array_equal(a, c), f
    # the oracle flags the same cells
    o = bad_in.copy()
    Oracle().flash_fim(eos, o)
    assert np.array_equal(o.nonconv, bad.nonconv)
    ctx.close()


@pytest.mark.parametrize("name", ["flash_limits_bench6.npz", "flash_maxit0_bench6.npz"])
def test_capped_iterations_flags_and_count(name):
    gpu_util.say(f"[failure] iteration caps {name}")
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    ctx = lib.Context(eos)
    for phase in ("cold", "warm"):
        b = parity.bulk_from_fixture(d, phase)
        g, st = gpu_util.gpu_flash(ctx, b)
        o = b.copy()
        Oracle().flash_fim(eos, o)
        assert np.array_equal(g.nonconv, o.nonconv), phase
        assert np.array_equal(g.nonconv, d[f"{phase}_out_nonconv"]), phase
        assert st.nonconverged == int(np.sum(g.nonconv == 1)), (phase, st.nonconverged)
        gpu_util.say(f"  {phase}: {int(np.sum(g.nonconv == 1))} of {b.nb} cells not converged")
    ctx.close()


def test_rollback_then_warm_flash_matches_oracle():
    """IsoT_FIM::ResetToLastTimeStep restores the skip-stability memory with
    the rest of the state (src/OCPFluidMethod.cpp:2859-2913): after a rolled
    back Newton update, the next flash of the restored state (from the same
    primary variables) is the oracle's flash of that state."""
    gpu_util.say("[failure] rollback, then the next warm flash")
    case = fim_case.build_case(8, 6, 5, seed=23)
    b = case["bulk"]
    eos = case["eos"]
    ctx = lib.Context(eos, case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    ctx.update_last_step()
    saved = b.copy()
    ctx.download(saved)
    # a large Newton-like move of P and the moles, flashed (changes the phase
    # state and the skip memory), then the rollback
    moved = abi.Bulk(b.nb, b.NC, fields=["P", "Ni"])
    ctx.download(moved)
    rng = np.random.default_rng(5)
    moved.P[:] += rng.uniform(200.0, 400.0, b.nb)
    moved.Ni[:] *= 1 + rng.uniform(-0.05, 0.05, moved.Ni.size)
    ctx.upload(moved)
    ctx.flash_fim(stats=False)
    mid = b.copy()
    ctx.download(mid)
    assert not np.array_equal(mid.skipP, saved.skipP)
    ctx.reset_to_last_step()
    back = b.copy()
    ctx.download(back)
    for f in ("P", "Ni", "S", "xij", "phaseNum", "skipFlag", "skipMinEig", "skipP", "skipT", "skipZi"):
        np.testing.assert_array_equal(getattr(back, f), getattr(saved, f), err_msg=f)
    # the next flash from the restored state: device and oracle
    ctx.flash_fim(stats=False)
    g = b.copy()
    ctx.download(g)
    o = saved.copy()
    orc = Oracle()
    orc.flash_fim(eos, o)
    ref = gpu_util.snapshot_dict(o, "orc_")
    sens = parity.rounding_sensitivity(lambda x: orc.flash_fim(eos, x), saved)
    rep = parity.compare_state(g, ref, "orc_", gpu_util.GPU_TOL, frac=gpu_util.GPU_FRAC,
                               loose=gpu_util.GPU_LOOSE, cell_tol=gpu_util.cell_bounds(sens))
    gpu_util.say("  next flash vs oracle:", {k: v for k, v in rep.items() if k != "cells_beyond_tol"})
    ctx.close()
