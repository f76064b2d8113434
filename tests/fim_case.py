"""A small FIM state on an orthogonal grid, built with the oracle: the
initial flash from (P, z, Sw), a Newton-like perturbation of P and N, a
flash, kr/Pc and rock.  Used by the residual/Jacobian parity tests."""
import numpy as np

from opencaeporo_amd import abi, fluids, workloads
from oracle.oracle_py import Oracle


def build_case(nx=8, ny=6, nz=5, seed=7, deck=None, dP=20.0, dN=0.01):
    orc = Oracle()
    from opencaeporo_amd import lib
    deck = deck or fluids.spe5_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    NC = eos.nc
    grid = workloads.cartesian_grid(nx, ny, nz, seed=seed)
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    orc.calrock(rock, b)
    orc.init_flash_fim(eos, b, z)
    orc.calkrpc(sat, b)
    # last time step, then a Newton-like update of the primary variables
    b.lNi[:] = b.Ni
    rng = np.random.default_rng(seed + 2)
    b.P[:] = b.P + rng.uniform(-dP, dP, nb)
    b.Ni[:] = b.Ni * (1 + rng.uniform(-dN, dN, b.Ni.size))
    orc.flash_fim(eos, b)
    orc.calkrpc(sat, b)
    orc.calrock(rock, b)
    return dict(eos=eos, sat=sat, rock=rock, grid=grid, bulk=b, NC=NC, orc=orc)
