"""HIP IsoT_FIM::GetSolution and the Newton-change bookkeeping against the
oracle through the C-ABI, from the same state and the same solution vector:
P, Ni, S, x_ij, the reported maxima and the wells' BHP bit for bit, with
and without the saturation chop; then, after the device flash, dSNR /
NRdSmax / maxNRdSSP against the oracle's bookkeeping on the same
saturations."""
import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from tests import fim_case, gpu_util
from tests.test_oracle_newton import random_update

pytestmark = pytest.mark.gpu


def device_context(case):
    b = case["bulk"]
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    return ctx


@pytest.mark.parametrize("scale,dSmax", [(1.0, 0.2), (60.0, 0.02)])
def test_get_solution_matches_oracle(scale, dSmax):
    gpu_util.say(f"[newton] GetSolution, update scale {scale}, dSmax {dSmax}")
    case = fim_case.build_case(8, 6, 5, seed=31)
    b, nc, orc = case["bulk"], case["NC"] + 1, case["orc"]
    u = random_update(b, nc, seed=9, scale=scale).ravel()
    ctx = device_context(case)
    dPmax, dNmax = ctx.get_solution(u, dSmax)
    g = b.copy()
    ctx.download(g)
    o = b.copy()
    book, (oPmax, oNmax) = orc.get_solution(nc, o, u, dSmax)
    for f in ("P", "Ni", "S", "xij"):
        np.testing.assert_array_equal(getattr(g, f), getattr(o, f), f)
    assert dPmax == oPmax and dNmax == oNmax
    # the flash from the updated state, then the bookkeeping
    ctx.flash_fim(stats=False)
    ch = ctx.nr_change()
    f = b.copy()
    ctx.download(f)
    _, och = orc.nr_change(f, book)
    gpu_util.say("  ", ch)
    assert ch == och
    ctx.close()


def test_get_solution_updates_well_bhp():
    gpu_util.say("[newton] GetSolution with wells")
    case = fim_case.build_case(8, 6, 5, seed=32)
    wells = fim_case.add_wells(case)
    b, nc = case["bulk"], case["NC"] + 1
    ctx = device_context(case)
    ctx.set_wells(wells)
    rows = ctx.rows()
    u = np.zeros(rows * (nc + 1))
    u[: b.nb * (nc + 1)] = random_update(b, nc, seed=10).ravel()
    rng = np.random.default_rng(3)
    u[b.nb * (nc + 1):] = rng.uniform(-50, 50, (rows - b.nb) * (nc + 1))
    bhp0 = ctx.get_well_bhp()
    ctx.get_solution(u, 0.2)
    bhp1 = ctx.get_well_bhp()
    expect = bhp0.copy()
    for w in range(wells.nw):
        if wells.open[w]:
            expect[w] = bhp0[w] + u[(b.nb + wells.row_of(w)) * (nc + 1)]
    np.testing.assert_array_equal(bhp1, expect)
    ctx.close()


def test_get_solution_at_scale_matches_oracle():
    """A 60 x 220 x 10 field (132,000 cells) of the benchmark fluid."""
    from opencaeporo_amd import fluids, workloads
    gpu_util.say("[newton] GetSolution on 60x220x10")
    eos = lib.eos_from_deck(fluids.bench6_deck())
    NC = eos.nc
    nc = NC + 1
    grid = workloads.cartesian_grid(60, 220, 10, seed=33)
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=34)
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    case = dict(eos=eos, sat=fluids.make_satfunc(), rock=fluids.make_rock(), grid=grid, bulk=b)
    ctx = device_context(case)
    ctx.calrock()
    ctx.init_flash_fim(z)
    ctx.flash_fim(stats=False)
    ctx.download(b)
    u = random_update(b, nc, seed=11, scale=5.0).ravel()
    dPmax, dNmax = ctx.get_solution(u, 0.05)
    g = b.copy()
    ctx.download(g)
    from oracle.oracle_py import Oracle
    o = b.copy()
    _, (oPmax, oNmax) = Oracle().get_solution(nc, o, u, 0.05)
    for f in ("P", "Ni", "S", "xij"):
        np.testing.assert_array_equal(getattr(g, f), getattr(o, f), f)
    assert dPmax == oPmax and dNmax == oNmax
    ctx.close()
