"""HIP flash (ocpg_flash_fim / ocpg_init_flash_fim) against the oracle and
the reference's golden vectors, through the C-ABI."""
import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from oracle.oracle_py import Oracle
from tests import gpu_util, parity

pytestmark = pytest.mark.gpu

CASES = ["flash_spe5.npz", "flash_bench6.npz", "flash_ten.npz",
         # the 9- and 7-component fluids of the reference's spe3 and
         # spe5large decks (their own instantiations)
         "flash_spe3.npz", "flash_spe5large.npz",
         # iteration caps (10, 2, 8, 2) and all 0: the non-convergence paths,
         # with the per-cell flag compared exactly
         "flash_limits_bench6.npz", "flash_maxit0_bench6.npz"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_flash_matches_oracle_and_reference(orc, name, phase):
    gpu_util.say(f"[flash parity] {name} {phase}")
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, phase)
    ctx = lib.Context(eos)
    g, stats = gpu_util.gpu_flash(ctx, b)
    o = b.copy()
    orc.flash_fim(eos, o)
    ref = gpu_util.snapshot_dict(o, "orc_")
    sens = parity.rounding_sensitivity(lambda x: orc.flash_fim(eos, x), b)
    rep = parity.compare_state(g, ref, "orc_", gpu_util.GPU_TOL, frac=gpu_util.GPU_FRAC,
                               loose=gpu_util.GPU_LOOSE, cell_tol=gpu_util.cell_bounds(sens))
    gpu_util.say("  vs oracle:", {k: v for k, v in rep.items() if k != "cells_beyond_tol"})
    gpu_util.say("  cells beyond 1e-10 (each within 10x its rounding sensitivity):",
                 rep["cells_beyond_tol"][:8])
    # the GPU takes the oracle's path: a different iteration count is a
    # last-bit difference deciding a convergence test, so it stays rare
    # (a wrong term in a Newton Jacobian would shift most cells)
    assert rep["iters_mismatch_cells"] <= 0.1 * b.nb, rep["iters_mismatch_cells"]
    rep2 = parity.compare_state(g, d, f"{phase}_out_", 1e-8, frac=0.97, loose=1e-4)
    gpu_util.say("  vs reference: within 1e-8", rep2["frac_within_tol"])
    assert stats.cells == b.nb
    assert stats.two_phase == int(np.sum(g.phaseNum == 3))
    ctx.close()


@pytest.mark.parametrize("name", ["spe5", "spe5large"])
def test_initflash_matches_oracle_and_reference(orc, name):
    gpu_util.say(f"[initflash parity] {name}")
    d = parity.load(f"initflash_{name}.npz")
    eos = parity.eos_from_fixture(parity.load(f"flash_{name}.npz"))
    nb = len(d["init_in_P"])
    b = abi.Bulk(nb, int(d["deck_nc"]))
    b.P[:] = d["init_in_P"]
    b.T[:] = d["init_in_T"]
    b.S[:] = d["init_in_S"]
    b.rockVp[:] = d["init_rockVp"]
    ctx = lib.Context(eos)
    g, _ = gpu_util.gpu_flash(ctx, b, init_z=d["init_z"])
    o = b.copy()
    orc.init_flash_fim(eos, o, d["init_z"])
    ref = gpu_util.snapshot_dict(o, "orc_")
    sens = parity.rounding_sensitivity(lambda x: orc.init_flash_fim(eos, x, d["init_z"]), b)
    bound = gpu_util.cell_bounds(sens)
    parity.compare_state(g, ref, "orc_", gpu_util.GPU_TOL, frac=gpu_util.GPU_FRAC,
                         loose=gpu_util.GPU_LOOSE, cell_tol=bound)
    err = parity.scaled_err(g.Ni.reshape(nb, -1), o.Ni.reshape(nb, -1))
    assert err.max() <= gpu_util.GPU_LOOSE and np.mean(err <= gpu_util.GPU_TOL) >= gpu_util.GPU_FRAC
    assert np.all(err <= bound)
    ctx.close()


def test_flash_is_deterministic():
    """Two runs on the same inputs agree bit for bit."""
    d = parity.load("flash_spe5.npz")
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, "cold")
    ctx = lib.Context(eos)
    g1, _ = gpu_util.gpu_flash(ctx, b)
    g2, _ = gpu_util.gpu_flash(ctx, b)
    for f in parity.STATE_FIELDS:
        assert np.array_equal(getattr(g1, f), getattr(g2, f)), f
    ctx.close()
