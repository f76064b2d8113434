"""The IsoT_FIM-shaped C++ host shim (include/ocpgpu_fim.hpp), driven from
C++ (tests/cpp/shim_driver.cpp) through one Newton iteration with wells, as
IsoT_FIM and as IsoT_FIMn:
its linear system must be the one the same C-ABI calls give, bit for bit,
and match the oracle's within the parity bound."""
import ctypes as C
import os

import numpy as np
import pytest

from opencaeporo_amd import abi, build, fluids, lib
from tests import fim_case, gpu_util

pytestmark = pytest.mark.gpu

DT = 5.0


@pytest.mark.parametrize("method", ["fim", "fimn"])
def test_shim_newton_iteration_matches_c_abi_and_oracle(method):
    gpu_util.say(f"[shim] one Newton iteration through IsoT_FIM_GPU, method {method}")
    fimn = method == "fimn"
    mcode = abi.METHOD_FIMN if fimn else abi.METHOD_FIM
    assert os.path.exists(build.SHIM_LIB), "build the library first (python -m opencaeporo_amd.build)"
    lib.load()  # libocpgpu.so first: the driver links against it
    drv = C.CDLL(build.SHIM_LIB)
    P = C.c_void_p
    drv.shim_newton_iteration.argtypes = [P, P, P, C.c_int64, P, P, P, P, P, P, C.c_double, P, P, P, P, P, C.c_int]
    case = fim_case.build_case(6, 5, 4, seed=21)
    nb, NC = case["bulk"].nb, case["NC"]
    nc, ncol = NC + 1, NC + 2
    grid = case["grid"]
    deck = fluids.spe5_deck()
    sat, rock = case["sat"], case["rock"]
    wells = fim_case.add_wells(case, seed=21)
    # inputs of the initial state and of the Newton update
    init = case["bulk"]
    P0, z, Sw = None, None, None
    from opencaeporo_amd import workloads
    P0, z, Sw = workloads.cell_inputs(nb, NC, seed=22)
    b = abi.Bulk(nb, NC)
    b.P[:] = P0
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    rng = np.random.default_rng(23)
    # the update is applied to the initial flash's moles
    ref = b.copy()
    ctx = lib.Context(case["eos"], sat, rock)
    ctx.set_method(mcode)
    ctx.set_cells(nb)
    ctx.upload(ref)
    ctx.set_conn(grid)
    ctx.calrock()
    ctx.init_flash_fim(z)
    ctx.calkrpc()
    ctx.update_last_step()
    ctx.download(ref)
    newP = ref.P + rng.uniform(-20, 20, nb)
    newNi = ref.Ni * (1 + rng.uniform(-0.01, 0.01, ref.Ni.size))
    # the C-ABI path
    up = abi.Bulk(nb, NC, fields=["P", "Ni"])
    up.P[:] = newP
    up.Ni[:] = newNi
    ctx.upload(up)
    ctx.set_wells(wells)
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    norms_c = ctx.calres(DT)
    ctx.assemble(DT)
    IA_c, JA_c, val_c, rhs_c = ctx.get_bsr()
    ctx.close()
    # the shim from C++
    nrow = nb + wells.open_count()
    IA = np.zeros(nrow + 1, dtype=np.int32)
    JA = np.zeros(len(JA_c), dtype=np.int32)
    val = np.zeros(len(val_c))
    rhs = np.zeros(len(rhs_c))
    norms = abi.ResNorm()
    conn = abi.conn_struct(grid.bId, grid.eId, grid.area)
    sb = b.copy()
    s = sb.cstruct()
    ws = wells.cstruct()
    zc = np.ascontiguousarray(z)
    rc = drv.shim_newton_iteration(C.byref(deck), C.byref(sat), C.byref(rock), nb, C.byref(conn), C.byref(s),
                                   zc.ctypes.data, C.byref(ws), newP.ctypes.data, newNi.ctypes.data, DT,
                                   IA.ctypes.data, JA.ctypes.data, val.ctypes.data, rhs.ctypes.data,
                                   C.byref(norms), mcode)
    assert rc == 0, rc
    np.testing.assert_array_equal(IA, IA_c)
    np.testing.assert_array_equal(JA, JA_c)
    np.testing.assert_array_equal(val, val_c)
    np.testing.assert_array_equal(rhs, rhs_c)
    assert norms.maxRelRes_V == norms_c.maxRelRes_V and norms.maxWellRelRes_mol == norms_c.maxWellRelRes_mol
    # and the oracle on the same primary variables
    orc = case["orc"]
    o = b.copy()
    orc.calrock(rock, o)
    (orc.init_flash_fimn if fimn else orc.init_flash_fim)(case["eos"], o, z)
    orc.calkrpc(sat, o)
    o.lNi[:] = o.Ni
    o.P[:] = newP
    o.Ni[:] = newNi
    (orc.flash_fimn if fimn else orc.flash_fim)(case["eos"], o)
    orc.calkrpc(sat, o)
    orc.calrock(rock, o)
    res_o, up_o, urho_o, _, _ = orc.calres_wells(case["eos"], wells, grid, o, DT)
    IA_o, JA_o = orc.bsr_pattern_wells(grid, wells)
    if fimn:
        val_o, rhs_o = orc.assemble_fimn(case["eos"], grid, o, DT, up_o, urho_o, IA_o, JA_o, res_o, wells=wells)
    else:
        val_o = orc.assemble_wells(case["eos"], wells, grid, o, DT, up_o, urho_o, IA_o, JA_o)
        rhs_o = res_o
    np.testing.assert_array_equal(IA, IA_o)
    np.testing.assert_array_equal(JA, JA_o)
    scale = np.abs(rhs_o).reshape(-1, ncol).max(axis=1, keepdims=True)
    err = (np.abs(rhs - rhs_o).reshape(-1, ncol) / np.where(scale > 0, scale, 1)).max()
    gpu_util.say("  rhs scaled err vs oracle", err)
    assert err < 1e-8
    from tests.test_gpu_fim import block_err
    be = block_err(val, val_o, ncol)
    gpu_util.say("  block err vs oracle max", be.max(), "median", np.median(be))
    assert np.quantile(be, 0.99) < 1e-8 and be.max() < 1e-6
