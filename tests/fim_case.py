"""A small FIM state on an orthogonal grid, built with the oracle: the
initial flash from (P, z, Sw), a Newton-like perturbation of P and N, a
flash, kr/Pc and rock.  Used by the residual/Jacobian parity tests."""
import numpy as np

from opencaeporo_amd import abi, fluids, workloads
from oracle.oracle_py import Oracle


def build_case(nx=8, ny=6, nz=5, seed=7, deck=None, dP=20.0, dN=0.01, grid=None, method="fim", T=None,
               bias=None):
    """grid: None (orthogonal 7-point) or "faulted" (workloads.faulted_grid,
    non-neighbour connections across thrown columns).  method "fimn": the
    IsoT_FIMn flash (InitFlashFIMn / FlashFIMn).  T (F) and bias (the feed
    the cells' compositions scatter around) default to the workloads'."""
    orc = Oracle()
    from opencaeporo_amd import lib
    deck = deck or fluids.spe5_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    NC = eos.nc
    if grid == "faulted":
        grid = workloads.faulted_grid(nx, ny, nz, seed=seed, frac=0.2)
    else:
        grid = workloads.cartesian_grid(nx, ny, nz, seed=seed)
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1, **({} if bias is None else dict(bias=bias)))
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F if T is None else T
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    orc.calrock(rock, b)
    fimn = method == "fimn"
    (orc.init_flash_fimn if fimn else orc.init_flash_fim)(eos, b, z)
    orc.calkrpc(sat, b)
    # last time step, then a Newton-like update of the primary variables
    b.lNi[:] = b.Ni
    rng = np.random.default_rng(seed + 2)
    b.P[:] = b.P + rng.uniform(-dP, dP, nb)
    b.Ni[:] = b.Ni * (1 + rng.uniform(-dN, dN, b.Ni.size))
    (orc.flash_fimn if fimn else orc.flash_fim)(eos, b)
    orc.calkrpc(sat, b)
    orc.calrock(rock, b)
    return dict(eos=eos, sat=sat, rock=rock, grid=grid, bulk=b, NC=NC, orc=orc, method=method)


def add_wells(case, spec=None, seed=11):
    """Wells on a case's grid (abi.Wells).  The default set exercises every
    branch: a rate injector, a BHP injector, a rate producer and a BHP
    producer, one cell perforated by two wells.  Injector perforation molar
    densities and the dG offsets are per-step inputs; they get plausible
    values here."""
    b, NC = case["bulk"], case["NC"]
    nc = NC + 1
    nb = b.nb
    rng = np.random.default_rng(seed)
    P = b.P
    inj_z = np.zeros(nc)
    inj_z[0] = 0.9
    inj_z[1] = 0.1
    if spec is None:
        c = [0, nb // 2, nb - 1, nb // 3, (2 * nb) // 3]
        spec = [
            dict(type=abi.WELL_INJ, mode=abi.WELL_RATE, perf=[c[0], c[1]], ctrlRate=500.0),
            dict(type=abi.WELL_PROD, mode=abi.WELL_BHP, perf=[c[2], c[1]]),
            dict(type=abi.WELL_PROD, mode=abi.WELL_RATE, perf=[c[3]], ctrlRate=20.0,
                 prodPhaseWeight=[1.0, 0.0, 0.0]),
            dict(type=abi.WELL_INJ, mode=abi.WELL_BHP, perf=[c[4]]),
        ]
    wells = []
    for s in spec:
        perf = np.asarray(s["perf"], dtype=np.int32)
        Pp = P[perf]
        inj = s["type"] == abi.WELL_INJ
        bhp = s.get("bhp", (Pp.max() + 300.0) if inj else (Pp.min() - 300.0))
        w = dict(type=s["type"], mode=s["mode"], perf=perf, bhp=bhp,
                 ctrlBHP=s.get("ctrlBHP", bhp + (25.0 if inj else -25.0)),
                 ctrlRate=s.get("ctrlRate", 100.0), open=s.get("open", 1),
                 WI=rng.uniform(5.0, 50.0, len(perf)), mult=np.ones(len(perf)),
                 dG=rng.uniform(-5.0, 5.0, len(perf)),
                 prodPhaseWeight=s.get("prodPhaseWeight", [1.0, 0.0, 0.0]))
        if inj:
            w["injZi"] = inj_z
            w["perfXi"] = rng.uniform(0.4, 0.6, len(perf))
        wells.append(w)
    return abi.Wells(nc, wells)


def deck_case_args(name):
    """build_case keywords for a named fluid: spe5, ten, and the fluids of
    the reference's spe3 (9 components, 200 F) and spe5large (7, 160 F)
    decks at their own temperatures and around their own feeds."""
    if name == "spe3":
        return dict(deck=fluids.spe3_deck(), T=fluids.SPE3["T_F"], bias=fluids.SPE3["feed"])
    if name == "spe5large":
        return dict(deck=fluids.spe5large_deck(), T=fluids.SPE5L["T_F"], bias=fluids.SPE5L["feed"])
    return dict(deck=fluids.ten_deck() if name == "ten" else fluids.spe5_deck())
