"""Corner-point geometry (opencaeporo_amd/cornergrid.py) against the
reference's own OCP_COORD (src/CornerGrid.cpp, compiled from its source by
oracle/Makefile 'ref'; fixtures written by tests/golden/gen_corner_golden.py)
for the reference's examples/cornerpoint grid and faulted synthetic grids
that run every case of the face-overlap construction; then closed forms the
geometry must meet on its own."""
import collections

import numpy as np
import pytest

from opencaeporo_amd import cornergrid, workloads
from tests import parity

G = parity.load("corner_geometry.npz")
NAMES = [str(n) for n in G["names"]]


def grid_of(name):
    nx, ny, nz = (int(x) for x in G[f"{name}_dims"])
    return cornergrid.CornerGrid(nx, ny, nz, G[f"{name}_coord"], G[f"{name}_zcorn"])


@pytest.mark.parametrize("name", NAMES)
def test_geometry_matches_reference(name):
    cg = grid_of(name)
    rel = lambda a, b: np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    # corners, centres and face centres are the reference's arithmetic
    for k in ("depth", "dx", "dy", "dz"):
        assert rel(getattr(cg, k), G[f"{name}_{k}"]) <= 1e-15, k
    # the volume integrates the trilinear map instead of the reference's
    # closed form: equal to rounding
    assert rel(cg.v, G[f"{name}_v"]) <= 1e-12
    # the connections, in the reference's order, and their factors
    assert np.array_equal(cg.conn_begin, G[f"{name}_begin"])
    assert np.array_equal(cg.conn_end, G[f"{name}_end"])
    assert np.array_equal(cg.conn_dir, G[f"{name}_dir"])
    assert rel(cg.conn_adB, G[f"{name}_adB"]) <= 1e-14
    assert rel(cg.conn_adE, G[f"{name}_adE"]) <= 1e-14


def test_fixtures_run_every_overlap_case():
    """The fixtures pin every branch of the non-quadrilateral overlap (the
    four edge-crossing tests give twelve reachable cases)."""
    seen = collections.Counter()
    orig = cornergrid.overlap_area

    def code_of(F1, F2):
        code = 0
        edges = (((F1[0], F1[1]), (F2[0], F2[1])), ((F1[2], F1[3]), (F2[0], F2[1])),
                 ((F1[0], F1[1]), (F2[2], F2[3])), ((F1[2], F1[3]), (F2[2], F2[3])))
        for bit, (e1, e2) in enumerate(edges):
            cp = cornergrid._cross_point(e1[0], e1[1], e2[0], e2[1])
            if (cp[0] - e1[0][0]) * (cp[0] - e1[1][0]) < 0:
                code |= 1 << bit
        return code

    def counting(F1, F2):
        seen[code_of(F1, F2)] += 1
        return orig(F1, F2)

    cornergrid.overlap_area = counting
    try:
        for name in NAMES:
            grid_of(name)
    finally:
        cornergrid.overlap_area = orig
    assert set(seen) >= {1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 13, 15}, sorted(seen)


def box_arrays(nx, ny, nz, dx, dy, dz, tops):
    """COORD / ZCORN of an unfaulted box grid with vertical pillars."""
    X, Y = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dy)
    coord = np.stack([X, Y, np.full_like(X, tops - 10), X, Y, np.full_like(X, tops + nz * dz + 10)], axis=-1)
    z = np.empty((nz, 2, ny, 2, nx, 2))
    for k in range(nz):
        z[k, 0] = tops + k * dz
        z[k, 1] = tops + (k + 1) * dz
    return coord.ravel(), z.ravel()


def test_box_grid_closed_forms_and_cartesian_connections():
    """Vertical pillars, flat layers: v = dx dy dz, depth at mid-layer,
    Ad_dd = face area / half-width, and BulkConn's Akd equal to the
    orthogonal-grid path's (workloads.Grid) for the same permeabilities."""
    nx, ny, nz, dx, dy, dz, tops = 5, 4, 3, 120.0, 80.0, 15.0, 8000.0
    coord, zcorn = box_arrays(nx, ny, nz, dx, dy, dz, tops)
    cg = cornergrid.CornerGrid(nx, ny, nz, coord, zcorn)
    nb = nx * ny * nz
    np.testing.assert_allclose(cg.v, dx * dy * dz, rtol=1e-13)
    k = np.arange(nb) // (nx * ny)
    np.testing.assert_allclose(cg.depth, tops + (k + 0.5) * dz, rtol=1e-15)
    np.testing.assert_allclose(cg.dx, dx, rtol=1e-15)
    want = {1: dy * dz / (dx / 2), 2: dz * dx / (dy / 2), 3: dx * dy / (dz / 2)}
    for d, a in want.items():
        sel = cg.conn_dir == d
        np.testing.assert_allclose(cg.conn_adB[sel], a, rtol=1e-13)
        np.testing.assert_allclose(cg.conn_adE[sel], a, rtol=1e-13)
    rng = np.random.default_rng(3)
    kx, ky, kz = (np.exp(rng.normal(np.log(100), 1.0, nb)) for _ in range(3))
    ntg = rng.uniform(0.5, 1.0, nb)
    c = cornergrid.Connections(cg, kx, ky, kz, ntg=ntg)
    o = workloads.Grid(nx, ny, nz, dx, dy, dz, tops, (kx, ky, kz), ntg=ntg)
    assert np.array_equal(c.bId, o.bId) and np.array_equal(c.eId, o.eId)
    np.testing.assert_allclose(c.area, o.area, rtol=1e-13)
    np.testing.assert_allclose(c.depth, o.depth, rtol=1e-15)


def test_one_layer_throw_shifts_the_lateral_connections():
    """A vertical fault with a throw of exactly one layer: across it, cell
    (i, j, k) meets (i + 1, j, k - 1) over its whole face and nothing else
    (the layer-k cell of the other column only touches along an edge)."""
    nx, ny, nz, dx, dy, dz, tops = 2, 1, 4, 50.0, 40.0, 10.0, 5000.0
    coord, zcorn = box_arrays(nx, ny, nz, dx, dy, dz, tops)
    z = zcorn.reshape(nz, 2, ny, 2, nx, 2)
    z[:, :, :, :, 1, :] += dz  # column i = 1 drops one layer
    cg = cornergrid.CornerGrid(nx, ny, nz, coord, z.ravel())
    lat = cg.conn_dir == 1
    pairs = set(zip(cg.conn_begin[lat].tolist(), cg.conn_end[lat].tolist()))
    cell = lambda i, k: k * nx + i
    assert pairs == {tuple(sorted((cell(0, k), cell(1, k - 1)))) for k in range(1, nz)}
    np.testing.assert_allclose(cg.conn_adB[lat], dy * dz / (dx / 2), rtol=1e-13)
