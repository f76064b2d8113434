"""The 147-cell, two-well, 10-step FIM case (opencaeporo_amd.workloads.
config1) set up for the oracle and for the device: the initial state from
the oracle's InitFlashFIM, the injector's surface-rate factor, the wells and
their setup."""
import numpy as np

from opencaeporo_amd import abi, fluids, lib, workloads
from oracle.oracle_py import Oracle


def build(inj_factor=None, seed=20241204, method="fim"):
    """Returns dict(eos, sat, rock, grid, bulk, wells, setup, NC, factor,
    method).  inj_factor: lbmol per Mscf of the injected gas (1000 XiPhase
    at surface conditions); the oracle's when None.  method "fimn": the
    initial state from InitFlashFIMn (IsoT_FIMn::InitReservoir)."""
    c = workloads.config1(seed=seed)
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    grid = c["grid"]
    nb = grid.nb
    NC = eos.nc
    orc = Oracle()
    b = abi.Bulk(nb, NC)
    b.P[:] = c["P"]
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = c["Sw"]
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    orc.calrock(rock, b)
    (orc.init_flash_fimn if method == "fimn" else orc.init_flash_fim)(eos, b, c["z"])
    orc.calkrpc(sat, b)
    b.lNi[:] = b.Ni
    if inj_factor is None:
        inj_factor = 1000.0 * orc.phase_density(eos, 14.7, 60.0, c["wells"][0]["injZi"], 1)[0]
    specs = [dict(w) for w in c["wells"]]
    specs[0]["ctrlRate"] = c["inj_mscf"] * inj_factor
    # Well::InitBHP: the first perforation's cell pressure
    for s in specs:
        s["bhp"] = float(b.P[s["perf"][0]])
    wells = abi.Wells(NC + 1, specs)
    setup = abi.WellSetup(*c["setup"])
    return dict(eos=eos, sat=sat, rock=rock, grid=grid, bulk=b, wells=wells, setup=setup, NC=NC,
                factor=inj_factor, method=method)


def device_context(case):
    ctx = lib.Context(case["eos"], case["sat"], case["rock"])
    if case.get("method") == "fimn":
        ctx.set_method(abi.METHOD_FIMN)
    b = case["bulk"]
    ctx.set_cells(b.nb)
    ctx.upload(b)
    ctx.set_conn(case["grid"])
    ctx.set_wells(case["wells"])
    ctx.set_well_setup(case["setup"])
    ctx.update_last_step()
    return ctx


def copy_wells(w):
    n = abi.Wells.__new__(abi.Wells)
    n.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in w.__dict__.items()})
    return n
