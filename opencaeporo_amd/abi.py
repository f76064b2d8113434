"""ctypes mirror of include/ocpgpu.h.

The structures here must match the C header field for field; the library
exports ocpg_sizeof() so a test can check the layout on both sides.
"""
import ctypes as C

import numpy as np

MAXNC = 16
NP = 3
MAXTAB = 64

OK = 0
EINVAL = -1
EUNSUP = -2
EHIP = -3
ESTATE = -4
ENODEV = -5
ETOOBIG = -6

_dblN = C.c_double * MAXNC


class Eos(C.Structure):
    _fields_ = [
        ("nc", C.c_int32),
        ("pvtw_rows", C.c_int32),
        ("Tc", _dblN),
        ("Pc", _dblN),
        ("Vc", _dblN),
        ("MW", _dblN),
        ("Acf", _dblN),
        ("OmegaA", _dblN),
        ("OmegaB", _dblN),
        ("Vshift", _dblN),
        ("Vcvis", _dblN),
        ("BIC", C.c_double * (MAXNC * MAXNC)),
        ("LBCcoef", C.c_double * 5),
        ("ssmsta_maxit", C.c_int32), ("ssmsta_tol", C.c_double), ("ssmsta_eyt", C.c_double),
        ("nrsta_maxit", C.c_int32), ("nrsta_tol", C.c_double),
        ("ssmsp_maxit", C.c_int32), ("ssmsp_tol", C.c_double),
        ("nrsp_maxit", C.c_int32), ("nrsp_tol", C.c_double),
        ("rr_maxit", C.c_int32), ("rr_tol", C.c_double),
        ("pvtw", (C.c_double * 5) * MAXTAB),
        ("std_rhow", C.c_double),
    ]


class EosDeck(C.Structure):
    _fields_ = [
        ("nc", C.c_int32),
        ("pvtw_rows", C.c_int32),
        ("Tc", _dblN),
        ("Pc", _dblN),
        ("MW", _dblN),
        ("Acf", _dblN),
        ("has_Vc", C.c_int32), ("Vc", _dblN),
        ("has_Zc", C.c_int32), ("Zc", _dblN),
        ("has_OmegaA", C.c_int32), ("OmegaA", _dblN),
        ("has_OmegaB", C.c_int32), ("OmegaB", _dblN),
        ("has_Vshift", C.c_int32), ("Vshift", _dblN),
        ("has_Vcvis", C.c_int32), ("Vcvis", _dblN),
        ("has_Zcvis", C.c_int32), ("Zcvis", _dblN),
        ("BIC_lower", C.c_double * (MAXNC * (MAXNC - 1) // 2)),
        ("LBCcoef", C.c_double * 5),
        ("ssmsta_maxit", C.c_int32), ("ssmsta_tol", C.c_double), ("ssmsta_eyt", C.c_double),
        ("nrsta_maxit", C.c_int32), ("nrsta_tol", C.c_double),
        ("ssmsp_maxit", C.c_int32), ("ssmsp_tol", C.c_double),
        ("nrsp_maxit", C.c_int32), ("nrsp_tol", C.c_double),
        ("rr_maxit", C.c_int32), ("rr_tol", C.c_double),
        ("pvtw", (C.c_double * 5) * MAXTAB),
        ("std_rhow", C.c_double),
    ]


class SatFunc(C.Structure):
    _fields_ = [
        ("swof_rows", C.c_int32),
        ("sgof_rows", C.c_int32),
        ("swof", (C.c_double * 4) * MAXTAB),
        ("sgof", (C.c_double * 4) * MAXTAB),
    ]


class Rock(C.Structure):
    _fields_ = [("Pref", C.c_double), ("cp1", C.c_double), ("cp2", C.c_double)]


_P = C.c_void_p
BULK_FIELDS = [
    ("P", "f8", "nb"), ("T", "f8", "nb"), ("Ni", "f8", "nb*nc"), ("lNi", "f8", "nb*nc"),
    ("S", "f8", "nb*np"), ("phaseNum", "i4", "nb"), ("xij", "f8", "nb*np*nc"),
    ("phaseExist", "u1", "nb*np"), ("Nt", "f8", "nb"), ("vf", "f8", "nb"), ("vfP", "f8", "nb"),
    ("vfi", "f8", "nb*nc"), ("rho", "f8", "nb*np"), ("xi", "f8", "nb*np"), ("mu", "f8", "nb*np"),
    ("rhoP", "f8", "nb*np"), ("xiP", "f8", "nb*np"), ("muP", "f8", "nb*np"),
    ("rhox", "f8", "nb*np*nc"), ("xix", "f8", "nb*np*nc"), ("mux", "f8", "nb*np*nc"),
    ("pSderExist", "u1", "nb*np"), ("pVnumCom", "i4", "nb*np"), ("bRowSizedSdP", "i4", "nb"),
    ("dSec_dPri", "f8", "nb*lend"),
    ("skipFlag", "u1", "nb"), ("skipMinEig", "f8", "nb"), ("skipP", "f8", "nb"),
    ("skipT", "f8", "nb"), ("skipZi", "f8", "nb*NC"),
    ("kr", "f8", "nb*np"), ("Pc", "f8", "nb*np"), ("Pj", "f8", "nb*np"),
    ("dKr_dS", "f8", "nb*np*np"), ("dPcj_dS", "f8", "nb*np*np"),
    ("v", "f8", "nb"), ("depth", "f8", "nb"), ("poroInit", "f8", "nb"), ("poro", "f8", "nb"),
    ("poroP", "f8", "nb"), ("rockVp", "f8", "nb"),
    ("ftype", "i4", "nb"), ("iters", "i4", "nb*5"), ("nonconv", "i4", "nb"),
    ("nj", "f8", "nb*np"), ("res_n", "f8", "nb*np*ncol"), ("resPc", "f8", "nb"),
]


class BulkC(C.Structure):
    _fields_ = [("nb", C.c_int64)] + [(f, _P) for f, _, _ in BULK_FIELDS]


class Conn(C.Structure):
    _fields_ = [("nconn", C.c_int64), ("bId", _P), ("eId", _P), ("area", _P)]


class ResNorm(C.Structure):
    _fields_ = [("maxRelRes_V", C.c_double), ("maxRelRes_N", C.c_double),
                ("maxId_V", C.c_int64), ("maxId_N", C.c_int64), ("maxWellRelRes_mol", C.c_double)]


class NRUpdate(C.Structure):
    _fields_ = [("NRdPmax", C.c_double), ("NRdNmax", C.c_double)]


class NRChange(C.Structure):
    _fields_ = [("NRdSmax", C.c_double), ("NRdSmax_cell", C.c_int64), ("maxNRdSSP", C.c_double),
                ("maxNRdSSP_cell", C.c_int64)]


WELL_INJ, WELL_PROD = 0, 1
# ocpg_set_method: IsoT_FIM / IsoT_FIMn
METHOD_FIM, METHOD_FIMN = 0, 1
WELL_BHP, WELL_RATE = 0, 1


class WellsC(C.Structure):
    _fields_ = [("nw", C.c_int32), ("type", _P), ("open", _P), ("mode", _P), ("unweight", _P),
                ("bhp", _P), ("ctrlBHP", _P), ("ctrlRate", _P), ("injZi", _P), ("prodPhaseWeight", _P),
                ("perfStart", _P), ("perfCell", _P), ("perfWI", _P), ("perfMult", _P), ("perfDG", _P),
                ("perfXi", _P), ("Psurf", C.c_double), ("Tsurf", C.c_double)]


class Wells:
    """Numpy-backed ocpg_wells (include/ocpgpu.h): per-well arrays of length
    nw (injZi nw*nc, prodPhaseWeight nw*3) and per-perforation arrays in
    well-major order."""

    INT = ("type", "open", "mode", "unweight", "perfStart", "perfCell")
    DBL = ("bhp", "ctrlBHP", "ctrlRate", "injZi", "prodPhaseWeight", "perfWI", "perfMult", "perfDG",
           "perfXi")

    def __init__(self, nc, wells, Psurf=14.7, Tsurf=60.0):
        """wells: list of dicts with keys type, mode, bhp, ctrlBHP, ctrlRate,
        perf (cells), WI, dG, and optionally open, unweight, mult, injZi,
        prodPhaseWeight, perfXi."""
        self.nc = nc
        self.nw = len(wells)
        self.Psurf, self.Tsurf = Psurf, Tsurf
        i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
        f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)
        self.type = i32([w["type"] for w in wells])
        self.open = i32([w.get("open", 1) for w in wells])
        self.mode = i32([w["mode"] for w in wells])
        self.unweight = i32([w.get("unweight", 0) for w in wells])
        self.bhp = f64([w["bhp"] for w in wells])
        self.ctrlBHP = f64([w["ctrlBHP"] for w in wells])
        self.ctrlRate = f64([w["ctrlRate"] for w in wells])
        self.injZi = f64(np.concatenate([np.asarray(w.get("injZi", np.zeros(nc)), float) for w in wells])
                         if wells else np.zeros(0))
        self.prodPhaseWeight = f64(np.concatenate([np.asarray(w.get("prodPhaseWeight", [1, 0, 0]), float)
                                                   for w in wells]) if wells else np.zeros(0))
        counts = [len(w["perf"]) for w in wells]
        self.perfStart = i32(np.concatenate([[0], np.cumsum(counts)]))
        self.nperf = int(self.perfStart[-1])
        cat = lambda key, default: f64(np.concatenate(
            [np.asarray(w.get(key, np.full(len(w["perf"]), default)), float) for w in wells])
            if wells else np.zeros(0))
        self.perfCell = i32(np.concatenate([w["perf"] for w in wells]) if wells else np.zeros(0))
        self.perfWI = cat("WI", 1.0)
        self.perfMult = cat("mult", 1.0)
        self.perfDG = cat("dG", 0.0)
        self.perfXi = cat("perfXi", 0.0)

    def cstruct(self):
        s = WellsC()
        s.nw = self.nw
        for f in self.INT + self.DBL:
            a = getattr(self, f)
            setattr(s, f, a.ctypes.data if a.size else None)
        s.Psurf, s.Tsurf = self.Psurf, self.Tsurf
        return s

    def open_count(self):
        return int(np.count_nonzero(self.open))

    def row_of(self, w):
        """Index of open well w among the open wells (its row after the bulk rows)."""
        return int(np.count_nonzero(self.open[:w]))


class WellSetupC(C.Structure):
    _fields_ = [("depth", _P), ("injTemp", _P), ("injPhase", _P), ("initMode", _P)]


class WellSetup:
    """Numpy-backed ocpg_well_setup: per well its BHP reference depth, the
    injected fluid's temperature (F) and phase (1 gas, 2 water) and the
    scheduled control mode."""

    def __init__(self, depth, injTemp, injPhase, initMode):
        self.depth = np.ascontiguousarray(depth, dtype=np.float64)
        self.injTemp = np.ascontiguousarray(injTemp, dtype=np.float64)
        self.injPhase = np.ascontiguousarray(injPhase, dtype=np.int32)
        self.initMode = np.ascontiguousarray(initMode, dtype=np.int32)

    def cstruct(self):
        s = WellSetupC()
        for f in ("depth", "injTemp", "injPhase", "initMode"):
            setattr(s, f, getattr(self, f).ctypes.data)
        return s


class MaxChange(C.Structure):
    _fields_ = [("dPmax", C.c_double), ("dNmax", C.c_double), ("dSmax", C.c_double), ("eVmax", C.c_double),
                ("dBHPmax", C.c_double)]


WELL_SUCCESS, WELL_NEGATIVE_PRESSURE, WELL_SWITCH = 0, 1, 2
BULK_NEGATIVE_PRESSURE, BULK_NEGATIVE_COMPONENTS_MOLES = -1, -3


class FlashStats(C.Structure):
    _fields_ = [("cells", C.c_int64), ("ftype_count", C.c_int64 * 3), ("two_phase", C.c_int64),
                ("iters", C.c_int64 * 5), ("nonconverged", C.c_int64)]


def max_len_dsdp(nc):
    """maxLendSdP of src/OCPFluidMethod.cpp:1056: (nc+1)^2 * np."""
    return (nc + 1) * (nc + 1) * NP


class Bulk:
    """Host arrays in the reference's Bulk layout (cell-major), numpy-owned.

    ``Bulk.cstruct()`` returns the ocpg_bulk view handed to C.  Fields absent
    from ``fields`` stay NULL.
    """

    def __init__(self, nb, NC, fields=None):
        self.nb = int(nb)
        self.NC = int(NC)
        self.nc = self.NC + 1
        dims = {"nb": self.nb, "nc": self.nc, "np": NP, "NC": self.NC, "ncol": self.nc + 1,
                "lend": max_len_dsdp(self.nc)}
        self.arrays = {}
        for name, dt, shape in BULK_FIELDS:
            if fields is not None and name not in fields:
                continue
            n = 1
            for tok in shape.split("*"):
                n *= dims[tok] if tok in dims else int(tok)
            self.arrays[name] = np.zeros(n, dtype=dt)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def copy(self):
        b = Bulk.__new__(Bulk)
        b.nb, b.NC, b.nc = self.nb, self.NC, self.nc
        b.arrays = {k: v.copy() for k, v in self.arrays.items()}
        return b

    def cstruct(self):
        s = BulkC()
        s.nb = self.nb
        for name, _, _ in BULK_FIELDS:
            a = self.arrays.get(name)
            setattr(s, name, a.ctypes.data if a is not None else None)
        return s


def conn_struct(bId, eId, area):
    s = Conn()
    s.nconn = len(bId)
    s.bId = bId.ctypes.data
    s.eId = eId.ctypes.data
    s.area = area.ctypes.data
    return s
