"""Input decks for the compositional FIM path: ECLIPSE-style keyword files
read into the library's inputs (the EoS, the saturation tables, the rock,
the grid and its connections, the wells and their first-step controls, and
the initial cell state).

The keyword semantics follow the reference's readers:

* tokens and defaults: ReadLine / DealDefault (src/UtilInput.cpp:16-72):
  ``--`` and ``#`` start comments, quotes are dropped, ``n*`` is n defaults
  and ``n*v`` n copies of v;
* grid: DIMENS, EQUALS / COPY / MULTIPLY with 1-based boxes, and the array
  keywords DX DY DZ TOPS PORO NTG PERMX PERMY PERMZ (src/ParamReservoir.cpp);
  geometry and connections as Grid::SetupOrthogonalGrid and
  BulkConn::Setup compute them (workloads.Grid); or corner-point geometry,
  COORD and ZCORN, with the non-neighbour connections across faults, as
  Grid::SetupCornerGrid computes it (opencaeporo_amd/cornergrid.py);
* fluid: COMPS, CNAMES, TCRIT, PCRIT, ZCRIT, VCRIT, MW, ACF, BIC (strict
  lower triangle), LBCCOEF, RR, SSMSTA, NRSTA, SSMSP, NRSP, PVTW, and INCLUDE
  (src/ParamReservoir.cpp:1280-1360, the EoS solver keywords are required as
  in src/MixtureComp.cpp:140-159);
* SWOF, SGOF, ROCK (LINEAR01: Pref, cp1; src/ParamReservoir.cpp:520-548),
  RTEMP, ZI, ZMFVD, EQUIL;
* wells: WELSPECS (name, group, I, J, depth), COMPDAT (name, I, J, K1, K2,
  WI, diameter, Kh, skin, direction; src/ParamWell.cpp:80-143), WCONPROD /
  WCONINJE (src/ParamWell.cpp:14-47), WELLSTRE; the controls are those in
  force for the first TSTEP.  The well index is Peaceman's
  (Well::CalWI_Peaceman, src/Well.cpp:93-161) and the initial BHP the
  pressure of the first perforated cell (include/Well.hpp:94).

Not read: inactive cells (ACTNUM, or PORO x NTG below the reference's
activity threshold): such decks raise DeckError.  The loaded state starts
every cell at EQUIL's datum pressure, with the composition of ZI or of
ZMFVD at its depth (OCPTable::Eval_All0: linear, flat outside the table),
the temperature of RTEMP and the water saturation of SWOF's first row;
Model.equilibrate replaces it with the reference's hydrostatic
equilibration (Bulk::InitPTSw, src/Bulk.cpp:443-1064; opencaeporo_amd/
equil.py), with the phase densities from the library.  Injection of a
solvent stream (WELLSTRE) in the first step raises DeckError.
"""
import math
import os

import numpy as np

from . import abi, cornergrid, equil, fluids, workloads

CONV2 = 1.12712E-3          # include/OCPConst.hpp:60
CONV1 = 5.61458             # include/OCPConst.hpp:59
PI = 3.141592653            # include/OCPConst.hpp:38 (the reference's truncated value)


class DeckError(ValueError):
    pass


def _tokens(line):
    """ReadLine's split of one line (src/UtilInput.cpp:14-48): lines opening
    with '#' or '--' are comments, the text after the first '/' is dropped
    and the '/' kept as a token, quotes and commas separate tokens.  A '--'
    comment after data is dropped too."""
    s = line.lstrip(" \t\r")
    if not s or s.startswith("#") or s.startswith("--"):
        return []
    s = s.split("--", 1)[0]
    cut = s.find("/")
    if cut >= 0:
        s = s[:cut] + " /"
    return s.replace("'", " ").replace(",", " ").replace('"', " ").split()


def _defaults(tokens):
    """DealDefault: 'n*' -> n x None, 'n*v' -> n x v (src/UtilInput.cpp:50-72)."""
    out = []
    for t in tokens:
        if "*" in t:
            head, _, tail = t.partition("*")
            n = int(head) if head.isdigit() else 0
            if n == 0:
                out.append(t)
            else:
                out += [tail if tail else None] * n
        else:
            out.append(t)
    return out


def _text(path):
    """A deck's lines; bytes outside ASCII (comments in other encodings) are
    kept as latin-1 characters, which no keyword or value uses."""
    with open(path, encoding="latin-1") as f:
        return f.read().splitlines()


class _Lines:
    def __init__(self, path):
        self.stack = [(path, _text(path), 0)]

    def include(self, name):
        base = os.path.dirname(self.stack[-1][0])
        p = os.path.join(base, name)
        self.stack.append((p, _text(p), 0))

    def next(self):
        """Next line's tokens (skipping empty and comment lines), or None."""
        while self.stack:
            path, lines, i = self.stack[-1]
            if i >= len(lines):
                self.stack.pop()
                continue
            self.stack[-1] = (path, lines, i + 1)
            toks = _tokens(lines[i])
            if toks:
                return toks
        return None

    def push_back(self, toks):
        path, lines, i = self.stack[-1]
        self.stack[-1] = (path, lines[:i] + [" ".join(toks)] + lines[i:], i)

    def record(self):
        """Tokens up to and including a '/', over as many lines as needed
        (None at the end of the file)."""
        out = []
        while True:
            t = self.next()
            if t is None:
                return out or None
            out += t
            if "/" in t:
                return out[: out.index("/") + 1]

    def records(self):
        """Records until a lone '/'."""
        recs = []
        while True:
            r = self.record()
            if r is None or r == ["/"]:
                return recs
            recs.append(_defaults(r[:-1]))

    def rows(self):
        """Table rows, one a line, until a '/' (which may end the last row)."""
        rows = []
        while True:
            t = self.next()
            if t is None:
                return rows
            if t[0] == "/":
                return rows
            end = "/" in t
            vals = [float(x) for x in _defaults(t[: t.index("/")] if end else t)]
            if vals:
                rows.append(vals)
            if end:
                return rows


_ARRAYS = ("DX", "DY", "DZ", "TOPS", "PORO", "NTG", "PERMX", "PERMY", "PERMZ")
_VECTORS = ("TCRIT", "PCRIT", "ZCRIT", "VCRIT", "MW", "ACF", "BIC", "LBCCOEF", "ZI", "SSHIFT", "OMEGAA", "OMEGAB")


class Deck:
    """The keywords of one deck that the compositional FIM path reads."""

    def __init__(self, path):
        self.path = path
        self.dims = None
        self.nc = None
        self.arrays = {}
        self.vec = {}
        self.names = []
        self.controls = {}
        self.tables = {}
        self.rock = None
        self.rtemp = None
        self.equil = None
        self.wells = []          # WELSPECS order
        self.streams = {}
        self.corner = False
        self.coord = None
        self.zcorn = None
        self.actnum = False
        self._read(path)

    # ---- grid arrays ----
    def _array(self, name):
        nb = self.nb
        if name not in self.arrays:
            self.arrays[name] = np.full(nb, np.nan)
        return self.arrays[name]

    @property
    def nb(self):
        nx, ny, nz = self.dims
        return nx * ny * nz

    def _box(self, vals):
        nx, ny, nz = self.dims
        full = (1, nx, 1, ny, 1, nz)
        b = [int(v) if v is not None else full[i] for i, v in enumerate((vals + [None] * 6)[:6])]
        i1, i2, j1, j2, k1, k2 = b
        if not (1 <= i1 <= i2 <= nx and 1 <= j1 <= j2 <= ny and 1 <= k1 <= k2 <= nz):
            raise DeckError(f"box {b} outside the {nx}x{ny}x{nz} grid")
        kk, jj, ii = np.meshgrid(np.arange(k1 - 1, k2), np.arange(j1 - 1, j2), np.arange(i1 - 1, i2), indexing="ij")
        return (kk * ny * nx + jj * nx + ii).ravel()

    def _read(self, path):
        L = _Lines(path)
        while True:
            t = L.next()
            if t is None:
                break
            kw = t[0].upper()
            if kw == "DIMENS":
                r = L.record() if len(t) == 1 else t[1:]
                self.dims = tuple(int(x) for x in r[:3])
            elif kw == "COMPS":
                self.nc = int((L.record() if len(t) == 1 else t[1:])[0])
            elif kw == "INCLUDE":
                r = L.record() if len(t) == 1 else t[1:]
                L.include(r[0])
            elif kw in ("COORD", "ZCORN"):
                self.corner = True
                vals = _defaults(L.record()[:-1])
                if any(v is None for v in vals):
                    raise DeckError(f"{kw} has defaulted values")
                setattr(self, kw.lower(), np.array([float(v) for v in vals]))
            elif kw == "ACTNUM":
                self.actnum = True
                L.record()
            elif kw == "EQUALS":
                for r in L.records():
                    a = self._array(r[0].upper())
                    a[self._box(r[2:])] = float(r[1])
            elif kw == "COPY":
                for r in L.records():
                    src = self._array(r[0].upper())
                    dst = self._array(r[1].upper())
                    idx = self._box(r[2:])
                    dst[idx] = src[idx]
            elif kw == "MULTIPLY":
                for r in L.records():
                    a = self._array(r[0].upper())
                    a[self._box(r[2:])] *= float(r[1])
            elif kw in _ARRAYS:
                vals = _defaults(L.record()[:-1])
                a = self._array(kw)
                n = min(len(vals), self.nb)
                a[:n] = [float(v) for v in vals[:n]]
            elif kw == "CNAMES":
                self.names = [x for x in L.record()[:-1]]
            elif kw in _VECTORS:
                self.vec[kw] = [float(v) for v in _defaults(L.record()[:-1])]
            elif kw in ("RR", "SSMSTA", "NRSTA", "SSMSP", "NRSP"):
                self.controls[kw] = [float(v) for v in L.record()[:-1]]
            elif kw in ("SWOF", "SGOF", "PVTW", "ZMFVD"):
                self.tables[kw] = L.rows()
            elif kw == "ROCK":
                r = L.record()
                if r[0].upper() not in ("LINEAR01",):
                    raise DeckError(f"ROCK model {r[0]} is not the linear one")
                self.rock = dict(Pref=float(r[1]), cp1=float(r[2]), cp2=0.0)
                nxt = L.next()          # the closing lone '/' of the ROCK table
                if nxt is not None and nxt[0] != "/":
                    L.push_back(nxt)
            elif kw == "RTEMP":
                self.rtemp = float(L.record()[0])
            elif kw == "EQUIL":
                r = [float(v) if v is not None else 0.0 for v in _defaults(L.record()[:-1])]
                self.equil = dict(Dref=r[0], Pref=r[1], DOWC=r[2], PcOW=r[3], DGOC=r[4], PcGO=r[5])
            elif kw == "WELSPECS":
                for r in L.records():
                    self.wells.append(dict(name=r[0], I=int(r[2]), J=int(r[3]),
                                           depth=float(r[4]) if r[4] is not None else None,
                                           perf=[], opt=None))
            elif kw == "COMPDAT":
                for r in L.records():
                    r = (r + [None] * 10)[:10]
                    for w in self._match(r[0]):
                        for k in range(int(r[3]), int(r[4]) + 1):
                            I = w["I"] if r[1] is None or r[2] is None else int(r[1])
                            J = w["J"] if r[1] is None or r[2] is None else int(r[2])
                            w["perf"].append(dict(
                                I=I, J=J, K=k,
                                WI=float(r[5]) if r[5] is not None else -1.0,
                                diameter=float(r[6]) if r[6] is not None else 1.0,
                                kh=float(r[7]) if r[7] is not None else -1.0,
                                skin=float(r[8]) if r[8] is not None else 0.0,
                                direction=(r[9] or "z").lower()))
            elif kw == "WCONPROD":
                for r in L.records():
                    for w in self._match(r[0]):
                        if w["opt"] is None:
                            w["opt"] = dict(type="PROD", state=r[1], mode=r[2],
                                            rate=float(r[3]) if r[3] is not None else 1e10,
                                            bhp=float(r[4]))
            elif kw == "WCONINJE":
                for r in L.records():
                    for w in self._match(r[0]):
                        if w["opt"] is None:
                            w["opt"] = dict(type="INJ", fluid=r[1], state=r[2], mode=r[3],
                                            rate=float(r[4]) if r[4] is not None else 1e10,
                                            bhp=float(r[5]))
            elif kw == "WELLSTRE":
                for r in L.records():
                    self.streams[r[0]] = [float(v) for v in r[1:]]
            elif kw == "TSTEP":
                # the controls of the first step are set; later steps are not read
                break

    def _match(self, pattern):
        """Wells named by a WCONPROD / COMPDAT pattern ('PROD*' matches a prefix)."""
        if pattern.endswith("*"):
            return [w for w in self.wells if w["name"].startswith(pattern[:-1])]
        return [w for w in self.wells if w["name"] == pattern]


def _eval0(table, x):
    """OCPTable::Eval_All0 (src/OCPTable.cpp:88-124): columns 1.. at x,
    linear on the row interval [x_r, x_r+1) holding x (slope first, then
    y_r + slope * (x - x_r)), the first row below the table, the last row
    at or above its end."""
    t = np.asarray(table, dtype=np.float64)
    if x < t[0, 0]:
        return t[0, 1:].copy()
    for r in range(1, len(t)):
        if x < t[r, 0]:
            r0 = r - 1
            slope = (t[r0 + 1, 1:] - t[r0, 1:]) / (t[r0 + 1, 0] - t[r0, 0])
            return t[r0, 1:] + slope * (x - t[r0, 0])
    return t[-1, 1:].copy()


class Model:
    """A deck turned into the library's inputs: ``eos_deck`` (abi.EosDeck),
    ``sat`` (abi.SatFunc), ``rock`` (abi.Rock), ``grid`` (workloads.Grid),
    ``bulk`` (abi.Bulk with P, T, S, v, depth, poroInit), ``z`` (nb x NC feed
    for ocpg_init_flash_fim) and ``wells`` (abi.Wells, or None)."""

    def __init__(self, deck):
        d = deck
        if d.dims is None or d.nc is None:
            raise DeckError("DIMENS and COMPS are required")
        if d.actnum:
            raise DeckError("ACTNUM (inactive cells) is not read")
        for kw in ("RR", "SSMSTA", "NRSTA", "SSMSP", "NRSP"):
            if kw not in d.controls:
                raise DeckError(f"the EoS solver keyword {kw} is required")
        NC = d.nc
        nx, ny, nz = d.dims
        nb = d.nb
        self.deck = d
        self.NC = NC
        # ---- fluid ----
        need = ("TCRIT", "PCRIT", "MW", "ACF")
        for kw in need:
            if len(d.vec.get(kw, [])) != NC:
                raise DeckError(f"{kw} needs {NC} values")
        comp = dict(Tc=d.vec["TCRIT"], Pc=d.vec["PCRIT"], MW=d.vec["MW"], Acf=d.vec["ACF"],
                    Zc=d.vec.get("ZCRIT", [0.0] * NC))
        # per-component EoS constants (MixtureComp's OmegaA / OmegaB)
        for kw, key in (("OMEGAA", "OmegaA"), ("OMEGAB", "OmegaB")):
            if kw in d.vec:
                if len(d.vec[kw]) != NC:
                    raise DeckError(f"{kw} needs {NC} values")
                comp[key] = d.vec[kw]
        if "ZCRIT" not in d.vec:
            raise DeckError("ZCRIT is required (VCRIT-only fluids are not read)")
        bic = d.vec.get("BIC", [0.0] * (NC * (NC - 1) // 2))
        if len(bic) != NC * (NC - 1) // 2:
            raise DeckError("BIC needs the strict lower triangle")
        c = d.controls
        controls = dict(rr_maxit=int(c["RR"][0]), rr_tol=c["RR"][1],
                        ssmsta_maxit=int(c["SSMSTA"][0]), ssmsta_tol=c["SSMSTA"][1],
                        ssmsta_eyt=c["SSMSTA"][2],
                        nrsta_maxit=int(c["NRSTA"][0]), nrsta_tol=c["NRSTA"][1],
                        ssmsp_maxit=int(c["SSMSP"][0]), ssmsp_tol=c["SSMSP"][1],
                        nrsp_maxit=int(c["NRSP"][0]), nrsp_tol=c["NRSP"][1])
        pvtw = d.tables.get("PVTW")
        if not pvtw:
            raise DeckError("PVTW is required")
        self.pvtw = pvtw
        self.eos_deck = fluids.make_deck(comp, bic, controls=controls, pvtw=pvtw,
                                         lbc=d.vec.get("LBCCOEF"))
        self.comp = comp
        self.bic = bic
        self.controls = controls
        swof, sgof = d.tables.get("SWOF"), d.tables.get("SGOF")
        if not swof or not sgof:
            raise DeckError("SWOF and SGOF are required")
        self.sat = fluids.make_satfunc(swof, sgof)
        if d.rock is None:
            raise DeckError("ROCK is required")
        self.rock = fluids.make_rock(**d.rock)
        # ---- grid ----
        arr = {}
        need = ("PORO", "PERMX", "PERMY", "PERMZ") if d.corner else ("DX", "DY", "DZ", "PORO", "PERMX", "PERMY",
                                                                       "PERMZ")
        for kw in need:
            a = d.arrays.get(kw)
            if a is None or np.isnan(a).any():
                raise DeckError(f"{kw} is not set on every cell")
            arr[kw] = a
        ntg = d.arrays.get("NTG")
        ntg = np.ones(nb) if ntg is None else np.where(np.isnan(ntg), 1.0, ntg)
        if (arr["PORO"] * ntg < 1e-8).any():
            raise DeckError("the deck has inactive cells (PORO x NTG ~ 0)")
        perm = (arr["PERMX"], arr["PERMY"], arr["PERMZ"])
        if d.corner:
            # Grid::SetupCornerGrid (src/Grid.cpp:200-245) and BulkConn::CalAkd
            if d.coord is None or d.zcorn is None:
                raise DeckError("corner-point geometry needs both COORD and ZCORN")
            try:
                cg = cornergrid.CornerGrid(nx, ny, nz, d.coord, d.zcorn)
                self.grid = cornergrid.Connections(cg, *perm, ntg=ntg)
            except cornergrid.CornerError as e:
                raise DeckError(str(e)) from None
            self.dxyz = (cg.dx, cg.dy, cg.dz)
        else:
            tops = d.arrays.get("TOPS")
            if tops is None or np.isnan(tops[: nx * ny]).any():
                raise DeckError("TOPS is not set on the top layer")
            self.grid = workloads.Grid(nx, ny, nz, arr["DX"], arr["DY"], arr["DZ"], tops[: nx * ny], perm, ntg=ntg)
            self.dxyz = (arr["DX"], arr["DY"], arr["DZ"])
        self.ntg = ntg
        self.perm = perm
        # ---- initial state ----
        if d.rtemp is None or d.equil is None:
            raise DeckError("RTEMP and EQUIL are required")
        depth = self.grid.depth
        if "ZI" in d.vec:
            z = np.tile(np.asarray(d.vec["ZI"][:NC], dtype=np.float64), (nb, 1))
        elif "ZMFVD" in d.tables:
            z = np.array([_eval0(d.tables["ZMFVD"], x) for x in depth])[:, :NC]
        else:
            raise DeckError("ZI or ZMFVD is required")
        self.z = np.ascontiguousarray(z)
        b = abi.Bulk(nb, NC)
        b.P[:] = d.equil["Pref"]
        b.T[:] = d.rtemp
        b.S.reshape(nb, 3)[:, 2] = swof[0][0]
        b.v[:] = self.grid.v
        b.depth[:] = depth
        b.poroInit[:] = arr["PORO"]
        self.bulk = b
        self.wells = self._wells(d, b)

    def equilibrate(self, density):
        """Bulk::InitPTSw (opencaeporo_amd/equil.py): replaces the datum-
        pressure start with the hydrostatic equilibrium of EQUIL: P, the
        water saturation and the ZMFVD feed of every cell.  density(P, T,
        z, phase) -> lb/ft3, e.g. from a library context:
        ``lambda P, T, z, ph: ctx.phase_density(P, T, z, ph)[1]``.  The
        wells' initial BHP follows their first perforation's new pressure."""
        d = self.deck
        if "ZMFVD" not in d.tables:
            raise DeckError("the equilibration reads the composition from ZMFVD")
        P, Sw, z = equil.init_ptsw(self.grid.depth, self.dxyz[2], d.equil, d.tables["ZMFVD"], d.rtemp,
                                   d.tables["SWOF"], d.tables["SGOF"], density)
        b = self.bulk
        b.P[:] = P
        S = b.S.reshape(b.nb, 3)
        S[:, 0] = 0.0
        S[:, 1] = 0.0
        S[:, 2] = Sw
        self.z = np.ascontiguousarray(z[:, :self.NC])
        if self.wells is not None:
            cells = self.wells.perfCell[self.wells.perfStart[:-1]]
            self.wells.bhp[:] = P[cells]
        return self

    def prepare_injectors(self, xi):
        """Molar densities of the solvent injectors' stream (gas phase):
        the surface-rate factor 1000 xi(14.7 psia, 60 F) that turns the
        control rate (Mscf/day) into lbmol/day (MixtureComp::SetupWellOpt)
        and the perforations' density at the well's BHP and the reservoir
        temperature (Well::CalFlux's injector xi; dG is zero at the first
        step).  xi(P, T, z, phase) -> lbmol/ft3."""
        for s in self.solvent:
            w = s["well"]
            factor = 1000.0 * xi(14.7, 60.0, s["zi"], 1)
            self.wells.ctrlRate[w] = s["rate"] * factor
            p0, p1 = self.wells.perfStart[w], self.wells.perfStart[w + 1]
            self.wells.perfXi[p0:p1] = xi(float(self.wells.bhp[w]), self.deck.rtemp, s["zi"], 1)
        self.solvent_prepared = True
        return self

    def initialize(self, ctx):
        """The deck's initial state with the library's densities (context
        ctx): the hydrostatic equilibration when the deck has ZMFVD, then
        the solvent injectors' densities at the new BHPs."""
        if "ZMFVD" in self.deck.tables:
            self.equilibrate(lambda P, T, z, ph: ctx.phase_density(P, T, z, ph)[1])
        self.prepare_injectors(lambda P, T, z, ph: ctx.phase_density(P, T, z, ph)[0])
        return self

    def cell(self, I, J, K):
        nx, ny, _ = self.deck.dims
        return (K - 1) * nx * ny + (J - 1) * nx + (I - 1)

    def peaceman_wi(self, p, n):
        """Well::CalWI_Peaceman for perforation p in cell n."""
        dx, dy, dz = (a[n] for a in self.dxyz)
        dz = dz * self.ntg[n]
        kx, ky, kz = (a[n] for a in self.perm)
        if p["direction"] == "x":
            kk, kr, a, b = ky * kz, ky / kz, dy, dz
            h = dx
        elif p["direction"] == "y":
            kk, kr, a, b = kz * kx, kz / kx, dz, dx
            h = dy
        else:
            kk, kr, a, b = kx * ky, kx / ky, dx, dy
            h = dz
        if kk <= 0:
            raise DeckError("Peaceman's well index needs positive permeabilities")
        ro = 0.28 * math.pow(a * a * math.pow(1 / kr, 0.5) + b * b * math.pow(kr, 0.5), 0.5)
        ro /= math.pow(kr, 0.25) + math.pow(1 / kr, 0.25)
        kh = p["kh"] if p["kh"] >= 0 else h * math.pow(kk, 0.5)
        return CONV2 * (2 * PI) * kh / (math.log(ro / (p["diameter"] / 2.0)) + p["skin"])

    def _wells(self, d, b):
        self.solvent = []
        if not d.wells:
            return None
        nc = self.NC + 1
        specs = []
        for w in d.wells:
            opt = w["opt"]
            if opt is None or not w["perf"]:
                raise DeckError(f"well {w['name']} has no control or no perforation")
            cells = [self.cell(p["I"], p["J"], p["K"]) for p in w["perf"]]
            # Well::CalWI_Peaceman stops at the first perforation given a WI
            wi, given = [], False
            for p, n in zip(w["perf"], cells):
                given = given or p["WI"] > 0
                wi.append(p["WI"] if given else self.peaceman_wi(p, n))
            inj = opt["type"] == "INJ"
            mode = abi.WELL_BHP if opt["mode"].upper() == "BHP" else abi.WELL_RATE
            bhp = float(b.P[cells[0]])
            spec = dict(type=abi.WELL_INJ if inj else abi.WELL_PROD, mode=mode, perf=cells, bhp=bhp,
                        ctrlBHP=opt["bhp"], ctrlRate=opt["rate"], open=int(opt["state"].upper() == "OPEN"),
                        WI=np.asarray(wi), mult=np.ones(len(cells)), dG=np.zeros(len(cells)))
            if inj:
                fluid = opt["fluid"].upper()
                if fluid in ("WAT", "WATER"):
                    zi = np.zeros(nc)
                    zi[-1] = 1.0
                    spec["injZi"] = zi
                    # MixtureComp::XiPhase for water at the perforation pressure
                    # (the well's BHP: dG is zero at the first step)
                    Pw0, bw0, cbw = self.pvtw[0][0], self.pvtw[0][1], self.pvtw[0][2]
                    bw = bw0 * (1 - cbw * (bhp - Pw0))
                    spec["perfXi"] = np.full(len(cells), 1 / (CONV1 * bw))
                else:
                    # a solvent stream (WELLSTRE) injected as gas
                    # (MixtureComp::SetupWellOpt, src/MixtureComp.cpp:770-800):
                    # its molar densities come from prepare_injectors
                    stream = next((v for k, v in d.streams.items() if k.upper() == fluid), None)
                    if stream is None:
                        raise DeckError(f"well {w['name']}: injected fluid {opt['fluid']} is not a WELLSTRE stream")
                    zi = np.zeros(nc)
                    k = min(len(stream), nc)
                    zi[:k] = stream[:k]
                    spec["injZi"] = zi
                    spec["perfXi"] = np.zeros(len(cells))
                    self.solvent.append(dict(well=len(specs), zi=zi, rate=opt["rate"]))
            else:
                m = opt["mode"].upper()
                spec["prodPhaseWeight"] = {"ORAT": [1.0, 0.0, 0.0], "GRAT": [0.0, 1.0, 0.0],
                                           "WRAT": [0.0, 0.0, 1.0], "LRAT": [1.0, 0.0, 1.0],
                                           "BHP": [1.0, 0.0, 0.0]}.get(m)
                if spec["prodPhaseWeight"] is None:
                    raise DeckError(f"well {w['name']}: production mode {m} is not read")
            specs.append(spec)
        return abi.Wells(nc, specs)


def load(path):
    """Read a deck and build its Model."""
    return Model(Deck(path))
