"""HIP CalKrPc/CalRock, CalRes and AssembleMat against the oracle, through
the C-ABI: the residual and every BSR block on the same cell state, and the
whole property pass + assembly from the same primary variables."""
import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from tests import fim_case, gpu_util, parity

pytestmark = pytest.mark.gpu

DT = 5.0


def oracle_system(case, b):
    orc, grid, nc = case["orc"], case["grid"], case["NC"] + 1
    res, up, urho, uvel, norms = orc.calres(nc, grid, b, DT)
    IA, JA = orc.bsr_pattern(grid)
    val = orc.assemble(nc, grid, b, DT, up, urho, IA, JA)
    return dict(res=res, up=up, urho=urho, uvel=uvel, norms=norms, IA=IA, JA=JA, val=val)


def gpu_system(ctx, case):
    ctx.set_conn(case["grid"])
    norms = ctx.calres(DT)
    ctx.assemble(DT)
    IA, JA, val, rhs = ctx.get_bsr()
    up, urho, uvel = ctx.get_upwind()
    return dict(res=rhs, up=up, urho=urho, uvel=uvel, norms=norms, IA=IA, JA=JA, val=val)


def block_err(a, b, ncol):
    """Per block: max |a-b| / max |b| over the block."""
    A = a.reshape(-1, ncol * ncol)
    Bm = b.reshape(-1, ncol * ncol)
    s = np.abs(Bm).max(axis=1)
    s = np.where(s > 0, s, 1.0)
    return (np.abs(A - Bm).max(axis=1) / s)


@pytest.mark.parametrize("shape", [(8, 6, 5), (5, 1, 7)])
def test_calres_and_assembly_on_same_state(shape):
    gpu_util.say(f"[fim parity same state] grid {shape}")
    case = fim_case.build_case(*shape)
    b = case["bulk"]
    ref = oracle_system(case, b)
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    got = gpu_system(ctx, case)
    assert np.array_equal(got["IA"], ref["IA"]) and np.array_equal(got["JA"], ref["JA"])
    assert np.array_equal(got["up"], ref["up"])
    np.testing.assert_array_equal(got["urho"], ref["urho"])
    np.testing.assert_array_equal(got["res"], ref["res"])
    assert got["norms"].maxRelRes_V == ref["norms"].maxRelRes_V
    assert got["norms"].maxRelRes_N == ref["norms"].maxRelRes_N
    assert got["norms"].maxId_V == ref["norms"].maxId_V
    ncol = case["NC"] + 2
    err = block_err(got["val"], ref["val"], ncol)
    gpu_util.say("  max block err", err.max(), "exact blocks", np.mean(err == 0))
    assert err.max() <= 1e-13
    ctx.close()


def test_property_pass_and_assembly_from_primary_variables():
    """GPU flash + kr/Pc + rock + CalRes + AssembleMat vs the oracle's, from
    the same P, N and last-step state."""
    gpu_util.say("[fim parity pipeline]")
    case = fim_case.build_case(8, 6, 5, seed=11)
    b = case["bulk"]
    orc = case["orc"]
    o = b.copy()
    orc.flash_fim(case["eos"], o)
    orc.calkrpc(case["sat"], o)
    orc.calrock(case["rock"], o)
    ref = oracle_system(case, o)
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    got = gpu_system(ctx, case)
    g = b.copy()
    ctx.download(g)
    assert np.array_equal(g.phaseNum, o.phaseNum)
    ncol = case["NC"] + 2
    rerr = parity.scaled_err(got["res"].reshape(b.nb, ncol), ref["res"].reshape(b.nb, ncol))
    err = block_err(got["val"], ref["val"], ncol)
    gpu_util.say("  res max scaled err", rerr.max(), " block max err", err.max(),
                 " blocks within 1e-10", np.mean(err <= 1e-10))
    assert rerr.max() <= 1e-10
    assert np.mean(err <= 1e-10) >= 0.99 and err.max() <= 1e-6
    ctx.close()


def test_async_handoff_into_pinned_arrays():
    """ocpg_get_bsr_async into page-locked arrays gives what ocpg_get_bsr gives."""
    gpu_util.say("[fim] asynchronous BSR hand-off")
    case = fim_case.build_case(8, 6, 5, seed=13)
    b = case["bulk"]
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    ctx.calres(DT, norms=False)
    ctx.assemble(DT)
    IA, JA, val, rhs = ctx.get_bsr()
    arrays = ctx.pinned_bsr()
    try:
        ctx.get_bsr_async(arrays)
        ctx.synchronize()
        for ref, (_, got) in zip((IA, JA, val, rhs), arrays):
            assert np.array_equal(ref, got)
    finally:
        ctx.free_pinned(arrays)
    ctx.close()
