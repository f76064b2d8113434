"""ocpg_update_assemble, the overlapped Newton iteration, against the four
serial calls it stands for (ocpg_flash_fim, ocpg_calkrpc_rock, ocpg_calres,
ocpg_assemble) from the same device state: every cell field, the residual,
every BSR block, the upwinding, the norms and the flash counters must be
bit-identical.  The serial calls are held to the oracle by the other GPU
tests, so this carries their parity over to the overlapped form.

The cases make every side path run: a cold pass (every cell takes the
stability test, which runs on side stream B), warm passes (a few hundred
cells run past the bounded split slices on side stream C), and the
10-component faulted grid with wells (the perforated rows recomputed after
the join)."""
import numpy as np
import pytest

from opencaeporo_amd import abi, fluids, lib, workloads
from tests import gpu_util

pytestmark = pytest.mark.gpu

DT = 5.0


def make_field(nx, ny, nz, seed, irregular=False, method=abi.METHOD_FIM):
    deck = fluids.ten_deck() if irregular else fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    NC = eos.nc
    grid = (workloads.faulted_grid(nx, ny, nz, seed=seed, frac=0.05) if irregular
            else workloads.cartesian_grid(nx, ny, nz, seed=seed))
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    b = abi.Bulk(nb, NC, fields=["P", "T", "S", "v", "depth", "poroInit"])
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    ctx = lib.Context(eos, sat, rock)
    if method != abi.METHOD_FIM:
        ctx.set_method(method)
    ctx.set_cells(nb)
    ctx.upload(b)
    ctx.set_conn(grid)
    ctx.calrock()
    ctx.init_flash_fim(z)
    ctx.calkrpc()
    ctx.update_last_step()
    ctx.upload(abi.Bulk(nb, NC, fields=["phaseNum", "skipFlag"]))
    return ctx, grid, NC


def warm_up(ctx, grid, NC, seed):
    """The cold flash result with P + U[-5, 5] psia (a Newton iteration's input)."""
    ctx.flash_fim(stats=False)
    pb = abi.Bulk(grid.nb, NC, fields=["P"])
    ctx.download(pb)
    pb.P[:] += np.random.default_rng(seed).uniform(-5.0, 5.0, grid.nb)
    ctx.upload(pb)
    ctx.calkrpc_rock()
    return pb.P


def run_serial(ctx):
    st = ctx.flash_fim(stats=True)
    ctx.calkrpc_rock()
    nr = ctx.calres(DT, norms=True)
    ctx.assemble(DT)
    return nr, st


def run_overlapped(ctx):
    return ctx.update_assemble(DT, norms=True, stats=True)


def snapshot(ctx, grid, NC, wells):
    b = abi.Bulk(grid.nb, NC)
    ctx.download(b)
    IA, JA, val, rhs = ctx.get_bsr()
    up, urho, uvel = ctx.get_upwind()
    out = dict(bulk=b, IA=IA, JA=JA, val=val, rhs=rhs, up=up, urho=urho, uvel=uvel)
    if wells is not None:
        out["qp"], out["qw"], out["pw"] = ctx.get_well_flux()
    return out


def assert_same(a, b):
    for k in a:
        if k == "bulk":
            for f, _, _ in abi.BULK_FIELDS:
                x, y = getattr(a[k], f), getattr(b[k], f)
                if x is None:
                    continue
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"bulk field {f} differs"
        else:
            assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), \
                f"{k} differs"


def compare(ctx, grid, NC, wells=None, label=""):
    ctx.save_flash_inputs()
    ctx.restore_flash_inputs()
    n1, s1 = run_serial(ctx)
    ref = snapshot(ctx, grid, NC, wells)
    for rep in range(2):
        ctx.restore_flash_inputs()
        n2, s2 = run_overlapped(ctx)
        got = snapshot(ctx, grid, NC, wells)
        gpu_util.say(f"  {label} rep {rep}: ftype {list(s2.ftype_count)}, two-phase {s2.two_phase}, "
                     f"iters {list(s2.iters)}")
        assert_same(got, ref)
        for f in ("maxRelRes_V", "maxRelRes_N", "maxId_V", "maxId_N", "maxWellRelRes_mol"):
            assert getattr(n1, f) == getattr(n2, f), f
        assert s1.cells == s2.cells and list(s1.ftype_count) == list(s2.ftype_count)
        assert s1.two_phase == s2.two_phase and list(s1.iters) == list(s2.iters)
        assert s1.nonconverged == s2.nonconverged
    return s1


def test_overlapped_cold_pass():
    gpu_util.say("[overlap] cold pass, 60x220x4, 6-comp")
    ctx, grid, NC = make_field(60, 220, 4, seed=501)
    st = compare(ctx, grid, NC, label="cold")
    assert st.ftype_count[0] == grid.nb  # every cell took the stability test on side stream B
    ctx.close()


def test_overlapped_warm_pass():
    gpu_util.say("[overlap] warm pass, 60x220x12, 6-comp")
    ctx, grid, NC = make_field(60, 220, 12, seed=502)
    warm_up(ctx, grid, NC, seed=503)
    b = abi.Bulk(grid.nb, NC, fields=["iters"])
    compare(ctx, grid, NC, label="warm")
    ctx.download(b)
    it = b.iters.reshape(grid.nb, 5)
    # cells past the bounded slices (side stream C) and ftype-0 cells (side stream B)
    gpu_util.say(f"  cells with > 96 split iterations: {int((it[:, 2] > 96).sum())}, "
                 f"stability tests: {int((it[:, 0] > 0).sum())}")
    assert (it[:, 2] > 96).sum() > 0
    ctx.close()


def test_overlapped_with_wells_ten_components():
    gpu_util.say("[overlap] warm pass with wells, faulted 60x220x5, 10-comp")
    ctx, grid, NC = make_field(60, 220, 5, seed=504, irregular=True)
    P = warm_up(ctx, grid, NC, seed=505)
    wells = workloads.irregular_wells(grid, P, NC, seed=506)
    ctx.set_wells(wells)
    compare(ctx, grid, NC, wells=wells, label="wells")
    ctx.close()


def test_overlapped_irregular_full_size():
    """The irregular-stencil configuration at full size (1,122,000 cells,
    faulted 10-component field, 40 wells): its Newton-restart list is long
    (thousands of cells), and side stream D works it one wavefront per cell
    over the whole chip while the main stream's kernels wait for CUs."""
    gpu_util.say("[overlap] irregular stencil, faulted 60x220x85, 10-comp, 40 wells")
    ctx, grid, NC = make_field(60, 220, 85, seed=2024, irregular=True)
    P = warm_up(ctx, grid, NC, seed=2026)
    wells = workloads.irregular_wells(grid, P, NC, seed=2027)
    ctx.set_wells(wells)
    compare(ctx, grid, NC, wells=wells, label="irregular")
    ctx.close()


@pytest.mark.parametrize("irregular", [False, True])
def test_overlapped_fimn(irregular):
    """IsoT_FIMn's overlapped iteration (its finishing stage on the main
    and the side streams' cells, its assembly and right-hand side, the rows
    recomputed after the join) against its serial calls, byte for byte."""
    gpu_util.say(f"[overlap] IsoT_FIMn warm pass, irregular {irregular}")
    if irregular:
        ctx, grid, NC = make_field(60, 220, 5, seed=514, irregular=True, method=abi.METHOD_FIMN)
        P = warm_up(ctx, grid, NC, seed=515)
        wells = workloads.irregular_wells(grid, P, NC, seed=516)
        ctx.set_wells(wells)
    else:
        ctx, grid, NC = make_field(60, 220, 12, seed=512, method=abi.METHOD_FIMN)
        warm_up(ctx, grid, NC, seed=513)
        wells = None
    compare(ctx, grid, NC, wells=wells, label="fimn")
    ctx.close()
