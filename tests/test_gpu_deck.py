"""The deck path on the GPU: the reference's SPE5 decks read by
opencaeporo_amd/deck.py drive the device through the C-ABI, against the
oracle on the same inputs.

* Model.equilibrate (Bulk::InitPTSw) with the library's phase densities
  against the oracle's equilibration, and the flash from that state.
* SPE5 (147 cells, the first step's producer) and a faulted corner-point
  deck (600 cells, non-neighbour connections across its faults, an injector
  and a producer): the device's CalRock + InitFlashFIM + CalKrPc from the
  deck's feed equal the oracle's; then, from a Newton-like update of that
  state, CalRes, the assembly and the well terms equal the oracle's on the
  same cell state (the test_gpu_wells bar).
* The refined deck (70 x 70 x 30 = 147,000 cells, a producer perforated in
  ten layers): one Newton iteration through ocpg_update_assemble equals the
  serial calls bit for bit and, on the state it produced, the oracle's
  residual and Jacobian; its flash equals the oracle's phase counts."""
import os

import numpy as np
import pytest

from opencaeporo_amd import abi, deck, lib
from tests import gpu_util
from tests.test_gpu_wells import compare, gpu_system, oracle_system

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SPE5 = os.path.join(HERE, "golden", "decks", "spe5", "spe5.data")
REFINE = os.path.join(HERE, "golden", "decks", "spe5refine", "spe5-70x70x30-2y.data")
CORNER = os.path.join(HERE, "golden", "decks", "corner", "corner-faulted.data")
SPE3 = os.path.join(HERE, "golden", "decks", "spe3", "spe3.data")
DT = 5.0


def _oracle():
    from oracle.oracle_py import Oracle
    return Oracle()


def _context(m, eos):
    ctx = lib.Context(eos, m.sat, m.rock)
    ctx.set_cells(m.grid.nb)
    ctx.upload(m.bulk)
    ctx.set_conn(m.grid)
    return ctx


@pytest.mark.parametrize("path,init", [(SPE5, False), (CORNER, False), (SPE3, True)], ids=["spe5", "corner", "spe3"])
def test_deck_initial_state_and_system(path, init):
    """The SPE5 deck (7x7x3, a producer), the faulted corner-point deck
    (12x10x5, 179 non-neighbour connections, an injector and a producer) and
    the reference's SPE3 deck (9x9x4, the 9-component gas condensate with
    per-component OMEGAA / OMEGAB, a solvent injector and a gas-rate
    producer), the last one equilibrated (Model.initialize) first."""
    gpu_util.say(f"[deck] {os.path.basename(path)}: init, then CalRes + assembly with its wells")
    m = deck.load(path)
    eos = lib.eos_from_deck(m.eos_deck)
    if init:
        c0 = lib.Context(eos)
        m.initialize(c0)
        c0.close()
        for sv in m.solvent:
            w = sv["well"]
            assert m.wells.perfXi[m.wells.perfStart[w]:m.wells.perfStart[w + 1]].min() > 0
    orc = _oracle()
    o = m.bulk.copy()
    orc.calrock(m.rock, o)
    orc.init_flash_fim(eos, o, m.z)
    orc.calkrpc(m.sat, o)
    ctx = _context(m, eos)
    ctx.calrock()
    ctx.init_flash_fim(m.z)
    ctx.calkrpc()
    g = m.bulk.copy()
    ctx.download(g)
    for f in ("Ni", "S", "phaseNum", "xij", "rockVp", "kr", "Pc", "Pj", "xi", "rho", "mu"):
        np.testing.assert_allclose(getattr(g, f), getattr(o, f), rtol=1e-12, atol=0, err_msg=f)
    assert np.array_equal(g.phaseNum, o.phaseNum)
    ctx.close()
    # a Newton-like update of the deck's state, then the system on it
    b = o.copy()
    b.lNi[:] = b.Ni
    rng = np.random.default_rng(5)
    b.P[:] += rng.uniform(-20.0, 20.0, b.nb)
    b.Ni[:] *= 1 + rng.uniform(-0.01, 0.01, b.Ni.size)
    orc.flash_fim(eos, b)
    orc.calkrpc(m.sat, b)
    orc.calrock(m.rock, b)
    case = dict(eos=eos, sat=m.sat, rock=m.rock, grid=m.grid, bulk=b, NC=m.NC, orc=orc)
    ref = oracle_system(case, m.wells, b)
    ctx = lib.Context(eos, m.sat, m.rock)
    ctx.set_cells(b.nb)
    ctx.upload(b)
    compare(case, m.wells, gpu_system(ctx, case, m.wells), ref)
    ctx.close()


def test_refined_deck_newton_iteration():
    gpu_util.say("[deck] SPE5 refined 70x70x30 (147,000 cells): update_assemble vs serial vs oracle")
    m = deck.load(REFINE)
    eos = lib.eos_from_deck(m.eos_deck)
    nb, NC = m.grid.nb, m.NC

    def prepared():
        ctx = _context(m, eos)
        ctx.calrock()
        ctx.init_flash_fim(m.z)
        ctx.calkrpc()
        ctx.update_last_step()
        pb = abi.Bulk(nb, NC, fields=["P"])
        ctx.download(pb)
        pb.P[:] += np.random.default_rng(7).uniform(-30.0, 30.0, nb)
        ctx.upload(pb)
        ctx.set_wells(m.wells)
        return ctx

    ctx = prepared()
    inp = m.bulk.copy()
    ctx.download(inp)
    norms, _ = ctx.update_assemble(DT, norms=True)
    IA, JA, val, rhs = ctx.get_bsr()
    qp, qw, pw = ctx.get_well_flux()
    got = m.bulk.copy()
    ctx.download(got)
    ctx.close()
    ctx = prepared()
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    ctx.calres(DT)
    ctx.assemble(DT)
    IA2, JA2, val2, rhs2 = ctx.get_bsr()
    ctx.close()
    assert np.array_equal(IA, IA2) and np.array_equal(JA, JA2)
    assert val.tobytes() == val2.tobytes() and rhs.tobytes() == rhs2.tobytes()
    # the oracle's flash from the same input state: phase counts exactly
    orc = _oracle()
    o = inp.copy()
    orc.flash_fim(eos, o)
    assert np.array_equal(got.phaseNum, o.phaseNum)
    gpu_util.say("   two-phase cells", int(np.count_nonzero(got.phaseNum >= 3)), "of", nb)
    # the residual and Jacobian on the state the device produced
    case = dict(eos=eos, sat=m.sat, rock=m.rock, grid=m.grid, bulk=got, NC=NC, orc=orc)
    ref = oracle_system(case, m.wells, got)
    compare(case, m.wells, dict(res=rhs, norms=norms, IA=IA, JA=JA, val=val, qp=qp, qw=qw, pw=pw), ref)


@pytest.mark.parametrize("path", [SPE5, CORNER], ids=["spe5", "corner"])
def test_equilibrated_deck(path):
    """Model.equilibrate (Bulk::InitPTSw) with the library's phase
    densities equals the oracle's equilibration on its own densities; the
    device's InitFlashFIM + CalKrPc from that state equal the oracle's."""
    gpu_util.say(f"[deck] {os.path.basename(path)}: hydrostatic equilibration with the device's densities")
    m = deck.load(path)
    eos = lib.eos_from_deck(m.eos_deck)
    orc = _oracle()
    ctx = lib.Context(eos, m.sat, m.rock)
    m.equilibrate(lambda P, T, z, ph: ctx.phase_density(P, T, z, ph)[1])
    d = m.deck
    P, Sw, z = orc.init_ptsw(eos, m.sat, m.grid.depth, m.dxyz[2],
                             [d.equil[k] for k in ("Dref", "Pref", "DOWC", "PcOW", "DGOC", "PcGO")],
                             d.tables["ZMFVD"], d.rtemp)
    np.testing.assert_allclose(m.bulk.P, P, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(m.bulk.S.reshape(-1, 3)[:, 2], Sw)
    np.testing.assert_array_equal(m.z, z)
    gpu_util.say("   P range", float(m.bulk.P.min()), float(m.bulk.P.max()))
    o = m.bulk.copy()
    orc.calrock(m.rock, o)
    orc.init_flash_fim(eos, o, m.z)
    orc.calkrpc(m.sat, o)
    ctx.set_cells(m.grid.nb)
    ctx.upload(m.bulk)
    ctx.set_conn(m.grid)
    ctx.calrock()
    ctx.init_flash_fim(m.z)
    ctx.calkrpc()
    g = m.bulk.copy()
    ctx.download(g)
    # (the pressures differ in the last bits, through the two libms'
    # densities: a trace-phase kr of 1e-18 may differ by its own size)
    for f in ("Ni", "S", "xij", "rockVp", "kr", "Pc", "Pj"):
        np.testing.assert_allclose(getattr(g, f), getattr(o, f), rtol=1e-12, atol=1e-14, err_msg=f)
    assert np.array_equal(g.phaseNum, o.phaseNum)
    ctx.close()
