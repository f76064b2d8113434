"""The deck reader (opencaeporo_amd/deck.py) on the reference's SPE5 decks,
pinned by what the decks and the mathematics fix: the fluid equals the
SPE5 constants the flash fixtures were made with (tests/golden, from the
reference's own EoS reader), cell geometry and transmissibilities equal
their closed forms on these uniform blocks, Peaceman's well index its
closed form for an isotropic horizontal permeability, and the table
interpolation the two-point line through the rows around the depth."""
import math
import os

import numpy as np
import pytest

from opencaeporo_amd import abi, deck, fluids

HERE = os.path.dirname(os.path.abspath(__file__))
SPE5 = os.path.join(HERE, "golden", "decks", "spe5", "spe5.data")
REFINE = os.path.join(HERE, "golden", "decks", "spe5refine", "spe5-70x70x30-2y.data")
CORNER = os.path.join(HERE, "golden", "decks", "corner", "corner-faulted.data")
SPE3 = os.path.join(HERE, "golden", "decks", "spe3", "spe3.data")


def test_tokens_and_defaults():
    assert deck._tokens("  'PROD1'   2*   3   3   1*   0.5   3*   / trailing text") == \
        ["PROD1", "2*", "3", "3", "1*", "0.5", "3*", "/"]
    assert deck._tokens("-- a comment") == [] and deck._tokens("# another") == []
    assert deck._tokens("70 70 30 / -- origin") == ["70", "70", "30", "/"]
    assert deck._defaults(["A", "2*", "3*7", "x"]) == ["A", None, None, "7", "7", "7", "x"]


def test_spe5_fluid_and_tables_are_the_pinned_ones():
    m = deck.load(SPE5)
    assert m.NC == 6 and m.deck.names == fluids.SPE5["names"]
    for k in ("Tc", "Pc", "Zc", "MW", "Acf"):
        assert m.comp[k] == fluids.SPE5[k], k
    assert m.bic == fluids.SPE5["BIC_lower"]
    assert m.controls == fluids.EOS_CONTROLS
    assert m.pvtw == fluids.PVTW_SPE5
    assert m.deck.tables["SWOF"] == fluids.SWOF_SPE5
    assert m.deck.tables["SGOF"] == fluids.SGOF_SPE5
    assert m.deck.rock == fluids.ROCK_SPE5
    # the EosDeck the library reads is the one the SPE5 fixtures use
    ref = fluids.spe5_deck()
    for f, _ in abi.EosDeck._fields_:
        a, b = getattr(m.eos_deck, f), getattr(ref, f)
        if hasattr(a, "__len__"):
            assert list(np.ravel(a)) == list(np.ravel(b)), f
        else:
            assert a == b, f


def test_spe5_grid_closed_forms():
    m = deck.load(SPE5)
    g = m.grid
    nx, ny, nz = 7, 7, 3
    assert (g.nx, g.ny, g.nz) == (nx, ny, nz)
    dz = np.repeat([20.0, 30.0, 50.0], nx * ny)
    np.testing.assert_array_equal(g.v, 500.0 * 500.0 * dz)
    np.testing.assert_array_equal(g.depth, np.repeat([8335.0, 8360.0, 8400.0], nx * ny))
    # 6x7x3 + 7x6x3 lateral and 7x7x2 vertical connections
    assert g.nconn == 2 * 6 * 7 * 3 + 7 * 7 * 2
    kx = np.repeat([500.0, 50.0, 200.0], nx * ny)
    kz = np.repeat([50.0, 50.0, 25.0], nx * ny)
    for b_, e_, a in zip(g.bId, g.eId, g.area):
        if e_ - b_ in (1, nx):             # lateral: T = k * dz (both halves the same)
            T = kx[b_] * 2 * 500.0 * dz[b_] / 500.0
            assert a == pytest.approx(T / 2, rel=1e-15)
        else:                               # vertical: series of two half-cells
            T1 = kz[b_] * 2 * 500.0 * 500.0 / dz[b_]
            T2 = kz[e_] * 2 * 500.0 * 500.0 / dz[e_]
            assert a == pytest.approx(T1 * T2 / (T1 + T2), rel=1e-14)
    np.testing.assert_array_equal(m.bulk.poroInit, 0.3)
    np.testing.assert_array_equal(m.bulk.P, 4000.0)
    np.testing.assert_array_equal(m.bulk.T, 160.0)
    np.testing.assert_array_equal(m.bulk.S.reshape(-1, 3)[:, 2], 0.2)
    # ZMFVD is flat in depth here
    np.testing.assert_array_equal(m.z, np.tile([0.5, 0.03, 0.07, 0.2, 0.15, 0.05], (g.nb, 1)))


def test_spe5_well_from_the_first_step():
    m = deck.load(SPE5)
    w = m.wells
    assert w.nw == 1 and list(w.perfCell) == [m.cell(7, 7, 3)] == [146]
    assert w.type[0] == abi.WELL_PROD and w.mode[0] == abi.WELL_RATE
    assert w.ctrlRate[0] == 12000.0 and w.ctrlBHP[0] == 1000.0 and w.bhp[0] == 4000.0
    np.testing.assert_array_equal(w.prodPhaseWeight, [1.0, 0.0, 0.0])
    # Peaceman, vertical well, kx = ky = k: r_o = 0.14 sqrt(dx^2 + dy^2), kh = k dz
    ro = 0.14 * math.hypot(500.0, 500.0)
    wi = 1.12712e-3 * 2 * 3.141592653 * (200.0 * 50.0) / math.log(ro / 0.25)
    assert w.perfWI[0] == pytest.approx(wi, rel=1e-14)


def test_refined_deck():
    m = deck.load(REFINE)
    nx, ny, nz = 70, 70, 30
    g = m.grid
    assert g.nb == 147000 and g.nconn == 2 * 69 * 70 * 30 + 70 * 70 * 29
    layer = np.arange(g.nb) // (nx * ny)
    np.testing.assert_array_equal(g.v, 50.0 * 50.0 * np.where(layer < 10, 2.0, np.where(layer < 20, 3.0, 5.0)))
    assert g.depth[0] == 8326.0 and g.depth[-1] == pytest.approx(8325 + 20 + 30 + 50 - 2.5, rel=1e-15)
    w = m.wells
    assert list(w.perfCell) == [m.cell(70, 70, k) for k in range(21, 31)]
    ro = 0.14 * math.hypot(50.0, 50.0)
    wi = 1.12712e-3 * 2 * 3.141592653 * (200.0 * 5.0) / math.log(ro / 0.25)
    np.testing.assert_allclose(w.perfWI, wi, rtol=1e-14)
    np.testing.assert_array_equal(m.z[0], [0.5, 0.03, 0.07, 0.2, 0.15, 0.05])


def test_table_evaluation():
    t = [[1000.0, 0.1, 5.0], [2000.0, 0.3, 4.0], [4000.0, 0.7, 0.0]]
    np.testing.assert_array_equal(deck._eval0(t, 500.0), [0.1, 5.0])
    np.testing.assert_array_equal(deck._eval0(t, 4000.0), [0.7, 0.0])
    np.testing.assert_array_equal(deck._eval0(t, 9000.0), [0.7, 0.0])
    for x in (1000.0, 1500.0, 2500.0, 3999.0):
        np.testing.assert_allclose(deck._eval0(t, x),
                                   [np.interp(x, [r[0] for r in t], [r[c] for r in t]) for c in (1, 2)],
                                   rtol=1e-15, atol=1e-15)


def _write(tmp_path, text, eos=None):
    """A deck in tmp_path beside a copy of the SPE5 EGOIL.in (INCLUDE takes
    a relative name: '/' ends a record)."""
    inc = os.path.join(os.path.dirname(SPE5), "EGOIL.in")
    (tmp_path / "EGOIL.in").write_text(eos if eos is not None else open(inc).read())
    p = tmp_path / "d.data"
    p.write_text(text)
    return str(p)


def test_unsupported_decks_are_rejected(tmp_path):
    base = open(SPE5, encoding="latin-1").read()
    with pytest.raises(deck.DeckError, match="corner-point"):
        deck.load(_write(tmp_path, base.replace("GRID ", "GRID\nCOORD\n1 2 3 /\n/\n", 1)))
    with pytest.raises(deck.DeckError, match="ACTNUM"):
        deck.load(_write(tmp_path, base.replace("GRID ", "GRID\nACTNUM\n147*1 /\n", 1)))
    with pytest.raises(deck.DeckError, match="inactive"):
        deck.load(_write(tmp_path, base.replace("'PORO'   0.3   4*  1  3", "'PORO'   0.0   4*  1  3")))
    eos = open(os.path.join(os.path.dirname(SPE5), "EGOIL.in")).read()
    with pytest.raises(deck.DeckError, match="NRSP"):
        deck.load(_write(tmp_path, base, eos=eos[:eos.index("NRSP")]))
    # the unchanged deck from the same place reads
    assert deck.load(_write(tmp_path, base)).grid.nb == 147


def test_corner_point_deck():
    """COORD / ZCORN decks: the cells and connections are the reference's
    corner-point geometry (the corner_deck case of
    tests/golden/corner_geometry.npz, written by OCP_COORD itself) with
    BulkConn::CalAkd's harmonic transmissibilities, in BulkConn order;
    Peaceman's well index takes the corner-point cell sizes."""
    from opencaeporo_amd import cornergrid
    from tests import parity
    m = deck.load(CORNER)
    g = m.grid
    G = parity.load("corner_geometry.npz")
    np.testing.assert_allclose(g.v, G["corner_deck_v"], rtol=1e-12)
    np.testing.assert_allclose(g.depth, G["corner_deck_depth"], rtol=1e-15)
    np.testing.assert_array_equal(m.bulk.depth, g.depth)
    b, e, d = G["corner_deck_begin"], G["corner_deck_end"], G["corner_deck_dir"]
    aB, aE = G["corner_deck_adB"], G["corner_deck_adE"]
    kx, ky, kz = m.perm
    k = lambda n: np.choose(d - 1, (kx[n], ky[n], kz[n]))
    area = 1 / (1 / (k(b) * aB) + 1 / (k(e) * aE))  # NTG is 1 in this deck
    order = np.lexsort((e, b))
    assert np.array_equal(g.bId, b[order]) and np.array_equal(g.eId, e[order])
    np.testing.assert_allclose(g.area, area[order], rtol=1e-14)
    # faults: connections between cells that are not grid neighbours
    nx, ny, _ = m.deck.dims
    step = np.abs(g.eId - g.bId)
    assert np.count_nonzero((step != 1) & (step != nx) & (step != nx * ny)) > 100
    # the producer's first perforation (I = 11, J = 9, K = 3), vertical:
    # Peaceman with the cell's corner-point sizes
    w = m.deck.wells[0]
    n = m.cell(11, 9, 3)
    dx, dy, dz = (a[n] for a in m.dxyz)
    cg = cornergrid.CornerGrid(*m.deck.dims, m.deck.coord, m.deck.zcorn)
    assert (dx, dy, dz) == (cg.dx[n], cg.dy[n], cg.dz[n])
    kk = kx[n] * ky[n]
    ro = 0.28 * math.sqrt(dx * dx * math.sqrt(ky[n] / kx[n]) + dy * dy * math.sqrt(kx[n] / ky[n]))
    ro /= (kx[n] / ky[n]) ** 0.25 + (ky[n] / kx[n]) ** 0.25
    wi = deck.CONV2 * 2 * deck.PI * dz * math.sqrt(kk) / math.log(ro / (0.5 / 2))
    assert w["name"] == "PROD1"
    assert m.wells.perfWI[0] == pytest.approx(wi, rel=1e-12)


def test_spe3_deck_fluid_and_solvent_injector():
    """The reference's SPE3 deck: its 9-component fluid equals the SPE3
    constants the 9-component flash fixtures were made with (per-component
    OMEGAA / OMEGAB included), and its solvent injector takes the WELLSTRE
    composition, with the surface-rate factor 1000 xi(14.7, 60 F) and the
    perforations' density from prepare_injectors."""
    m = deck.load(SPE3)
    ref = fluids.spe3_deck()
    e = m.eos_deck
    assert e.nc == 9 and e.has_OmegaA and e.has_OmegaB
    for f in ("Tc", "Pc", "MW", "Acf", "Zc", "OmegaA", "OmegaB"):
        assert list(getattr(e, f)[:9]) == list(getattr(ref, f)[:9]), f
    assert list(e.BIC_lower[:36]) == list(ref.BIC_lower[:36])
    assert m.deck.rtemp == fluids.SPE3["T_F"]
    w = m.wells
    assert w.nw == 2 and w.type[0] == abi.WELL_INJ and w.type[1] == abi.WELL_PROD
    assert list(w.prodPhaseWeight[3:6]) == [0.0, 1.0, 0.0]  # GRAT
    np.testing.assert_allclose(w.injZi[:10], m.deck.streams["Solvent"][:10])
    seen = []

    def xi(P, T, z, phase):
        seen.append((P, T, phase))
        return 0.25 if P == 14.7 else 0.5

    m.prepare_injectors(xi)
    assert w.ctrlRate[0] == 4700.0 * 1000.0 * 0.25
    assert np.all(w.perfXi[:w.perfStart[1]] == 0.5)
    assert seen[0] == (14.7, 60.0, 1) and seen[1] == (float(w.bhp[0]), 200.0, 1)
