"""Shared helpers of the parity tests: fixture loading and the comparison
rules (integer fields exact; floating fields by a per-cell scaled error)."""
import os

import numpy as np

from opencaeporo_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

INT_FIELDS = ["phaseNum", "phaseExist", "pSderExist", "pVnumCom", "bRowSizedSdP", "ftype",
              "skipFlag"]
FLOAT_FIELDS = ["S", "xij", "rho", "xi", "mu", "rhoP", "xiP", "muP", "rhox", "xix", "mux", "Nt",
                "vf", "vfP", "vfi", "dSec_dPri", "skipP", "skipT", "skipZi"]
STATE_FIELDS = INT_FIELDS + FLOAT_FIELDS + ["skipMinEig", "iters"]
FIMN_FIELDS = ["nj", "res_n", "resPc"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def eos_from_fixture(d):
    """ocpg_eos with the values the reference constructor derived."""
    e = abi.Eos()
    NC = int(d["deck_nc"])
    e.nc = NC
    for i in range(NC):
        e.Tc[i] = d["deck_Tc"][i]
        e.Pc[i] = d["deck_Pc"][i]
        e.MW[i] = d["deck_MW"][i]
        e.Acf[i] = d["deck_Acf"][i]
        e.Vc[i] = d["eos_Vc"][i]
        e.OmegaA[i] = d["eos_OmegaA"][i]
        e.OmegaB[i] = d["eos_OmegaB"][i]
        e.Vshift[i] = d["eos_Vshift"][i]
        e.Vcvis[i] = d["eos_Vcvis"][i]
        for k in range(NC):
            e.BIC[i * abi.MAXNC + k] = d["eos_BIC"][i, k]
    for q in range(5):
        e.LBCcoef[q] = d["eos_LBCcoef"][q]
    c = d["deck_controls"]
    (e.ssmsta_maxit, e.ssmsta_tol, e.ssmsta_eyt, e.nrsta_maxit, e.nrsta_tol, e.ssmsp_maxit,
     e.ssmsp_tol, e.nrsp_maxit, e.nrsp_tol, e.rr_maxit, e.rr_tol) = (
        int(c[0]), c[1], c[2], int(c[3]), c[4], int(c[5]), c[6], int(c[7]), c[8], int(c[9]), c[10])
    pv = d["deck_pvtw"]
    e.pvtw_rows = len(pv)
    for r in range(len(pv)):
        for q in range(5):
            e.pvtw[r][q] = pv[r][q]
    e.std_rhow = float(d["deck_std_rhow"])
    return e


def set_field(b, f, a):
    dst = getattr(b, f)
    if f == "dSec_dPri":
        L = abi.max_len_dsdp(b.nc)
        dst.reshape(b.nb, L)[:, :a.shape[1]] = a
    else:
        dst[:] = a


def bulk_from_fixture(d, phase):
    """Bulk holding a pass's inputs; for 'warm', the cold outputs first."""
    NC = int(d["deck_nc"])
    nb = len(d[f"{phase}_in_P"])
    b = abi.Bulk(nb, NC)
    if phase == "warm":
        for f in STATE_FIELDS:
            set_field(b, f, d["cold_out_" + f])
    for key in d:
        if key.startswith(phase + "_in_"):
            set_field(b, key[len(phase) + 4:], d[key])
    return b


def field(b, f, used=None):
    a = getattr(b, f)
    if f == "dSec_dPri":
        L = abi.max_len_dsdp(b.nc)
        a = a.reshape(b.nb, L)[:, :used]
    return a.reshape(b.nb, -1)


def scaled_err(a, ref):
    """Per cell: max |a-ref| / max |ref| over the cell's entries of the field
    (a field that is zero in a cell is compared absolutely)."""
    a = a.reshape(ref.shape[0], -1)
    ref = ref.reshape(ref.shape[0], -1)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale = np.where(scale > 0, scale, 1.0)
    return (np.abs(a - ref) / scale).max(axis=1)


def worst_scaled_err(b, d, prefix):
    """Per cell: the largest scaled error over the floating state fields
    (with IsoT_FIMn's when d holds them)."""
    worst = np.zeros(b.nb)
    for f in FLOAT_FIELDS:
        ref = d[prefix + f]
        got = field(b, f, used=ref.shape[1] if f == "dSec_dPri" else None)
        worst = np.maximum(worst, scaled_err(got, ref.reshape(b.nb, -1)))
    if prefix + "res_n" in d:
        for e in fimn_errs(b, d, prefix).values():
            worst = np.maximum(worst, e)
    return worst


# (input, relative perturbation) pairs of rounding_sensitivity: a few units in
# the last place of P, T, the first component's moles and all moles.
SENS_PERTURBATIONS = [("P", 2e-16), ("P", -2e-16), ("P", 1e-15), ("P", -1e-15), ("T", 1e-15), ("T", -1e-15),
                      ("N0", 2e-16), ("N0", -2e-16), ("N", 2e-16), ("N", -2e-16)]


def rounding_sensitivity(flash, bulk, perturbations=SENS_PERTURBATIONS):
    """Per cell: how far the FP64 flash's outputs move when its inputs (P, T,
    moles) move by a few units in the last place.  ``flash(b)`` runs the
    flash in place on a copy of ``bulk``.  A cell whose phase count changes
    under a perturbation gets +inf.  This is the cell's conditioning: two
    correct FP64 implementations that round differently (other libm, other
    LAPACK) differ by about this much, whatever their quality."""
    from tests import gpu_util
    base = bulk.copy()
    flash(base)
    ref = gpu_util.snapshot_dict(base, "s_", fimn=base.resPc is not None and bool(base.res_n.any()))
    sens = np.zeros(bulk.nb)
    for what, e in perturbations:
        p = bulk.copy()
        if what == "P":
            p.P[:] = bulk.P * (1 + e)
        elif what == "T":
            p.T[:] = bulk.T * (1 + e)
        else:
            Ni = p.Ni.reshape(bulk.nb, bulk.nc)
            if what == "N0":
                Ni[:, 0] *= 1 + e
            else:
                Ni[:, :] *= 1 + e
        flash(p)
        w = worst_scaled_err(p, ref, "s_")
        w[(field(p, "phaseNum") != field(base, "phaseNum")).any(axis=1)] = np.inf
        sens = np.maximum(sens, w)
    return sens


def fimn_errs(b, d, prefix):
    """Per-cell errors of the IsoT_FIMn fields: nj over the existing phases
    relative to Nt; res_n relative to the larger of its own magnitude and
    the variable's scale (one for the saturation rows, the first phaseNum;
    Nt for the n_ij rows); resPc relative to the larger of |resPc| and vf.
    At a converged split res_n is at rounding level and is compared on the
    variables' scale; where the split is poorly converged it is large and
    compared relatively."""
    nb = b.nb
    Nt = np.abs(d[prefix + "Nt"]).reshape(nb, 1)
    vf = np.abs(d[prefix + "vf"])
    ex = d[prefix + "phaseExist"].reshape(nb, -1) != 0
    nj = np.where(ex, np.abs(b.nj.reshape(nb, -1) - d[prefix + "nj"].reshape(nb, -1)), 0.0) / Nt
    rows = d[prefix + "bRowSizedSdP"].reshape(nb)
    nS = d[prefix + "phaseNum"].reshape(nb)
    rr = np.arange(b.res_n.size // nb)[None, :]
    ref = d[prefix + "res_n"].reshape(nb, -1)
    dr = np.abs(b.res_n.reshape(nb, -1) - ref)
    dr = np.where(rr < rows[:, None], dr, 0.0)
    dr = dr / np.maximum(np.abs(ref), np.where(rr < nS[:, None], 1.0, Nt))
    rpc = np.abs(b.resPc - d[prefix + "resPc"]) / np.maximum(np.abs(d[prefix + "resPc"]), vf)
    return {"nj": nj.max(axis=1), "res_n": dr.max(axis=1), "resPc": rpc}


def cell_field_errs(b, d, prefix, c):
    """Scaled error of each floating field in cell c (diagnostics)."""
    out = {}
    for f in FLOAT_FIELDS:
        ref = d[prefix + f]
        got = field(b, f, used=ref.shape[1] if f == "dSec_dPri" else None)
        out[f] = scaled_err(got, ref.reshape(b.nb, -1))[c]
    if prefix + "res_n" in d:
        for f, e in fimn_errs(b, d, prefix).items():
            out[f] = e[c]
    return out


def compare_state(b, d, prefix, tol, int_exact=True, skip_tol=1e-4, frac=1.0, loose=None, cell_tol=None,
                  max_over=0):
    """Integer fields must match in every cell.  A floating field must be
    within ``tol`` (per-cell scaled error) in at least ``frac`` of the cells
    and within ``loose`` everywhere (``loose`` defaults to ``tol``).  With
    ``cell_tol`` (one bound per cell), every cell but at most ``max_over``
    must also be within its own bound.  Returns a report dict; raises
    AssertionError on a breach."""
    loose = tol if loose is None else loose
    report = {}
    worst = np.zeros(b.nb)
    for f in INT_FIELDS:
        ref = d[prefix + f].reshape(b.nb, -1)
        got = field(b, f).astype(ref.dtype)
        bad = np.any(got != ref, axis=1)
        report[f] = int(bad.sum())
        if int_exact:
            assert not bad.any(), f"{f}: {bad.sum()} cells differ, first {np.flatnonzero(bad)[:5]}"
    if prefix + "nonconv" in d:
        # the per-cell non-convergence flag (ePEC = 1e8) must match exactly
        bad = b.nonconv != d[prefix + "nonconv"]
        report["nonconv"] = int(bad.sum())
        report["nonconv_cells"] = int(b.nonconv.sum())
        assert not bad.any(), f"nonconv: {bad.sum()} cells differ, first {np.flatnonzero(bad)[:5]}"
    for f in FLOAT_FIELDS:
        ref = d[prefix + f]
        got = field(b, f, used=ref.shape[1] if f == "dSec_dPri" else None)
        err = scaled_err(got, ref.reshape(b.nb, -1))
        worst = np.maximum(worst, err)
        report[f] = float(err.max())
        assert err.max() <= loose, f"{f}: max scaled error {err.max():.3e} > {loose:.1e} at cell {err.argmax()}"
    if prefix + "res_n" in d:
        # IsoT_FIMn: nj of the existing phases and the flash's correction,
        # scaled by the cell's moles (n_ij rows) or by one (saturation rows)
        err = fimn_errs(b, d, prefix)
        for f, e in err.items():
            worst = np.maximum(worst, e)
            report[f] = float(e.max())
            assert e.max() <= loose, f"{f}: max scaled error {e.max():.3e} > {loose:.1e} at cell {e.argmax()}"
    within = float(np.mean(worst <= tol))
    report["frac_within_tol"] = within
    report["cells_beyond_tol"] = [int(i) for i in np.flatnonzero(worst > tol)]
    assert within >= frac, f"only {within:.4f} of cells within {tol:.0e} (need {frac})"
    if cell_tol is not None:
        over = np.flatnonzero(worst > cell_tol)
        report["cells_over_own_bound"] = [int(i) for i in over]
        if len(over):
            it_ref = d[prefix + "iters"].reshape(b.nb, 5)
            detail = {int(c): dict(iters=field(b, "iters")[c].tolist(), ref_iters=it_ref[c].tolist(),
                                   fields={f: float(v) for f, v in cell_field_errs(b, d, prefix, c).items()
                                           if v > cell_tol[c]})
                      for c in over[:5]}
        assert len(over) <= max_over, (f"{len(over)} cells beyond their conditioning bound, first {over[:5]}: "
                                f"err {worst[over[:5]]} bound {cell_tol[over[:5]]}; {detail}")
    ref = d[prefix + "skipMinEig"]
    got = b.skipMinEig
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    err[(ref == 0) & (got == 0)] = 0
    report["skipMinEig"] = float(err.max())
    assert err.max() <= skip_tol, f"skipMinEig: {err.max():.3e}"
    it_ref = d[prefix + "iters"].reshape(b.nb, 5)
    report["iters_mismatch_cells"] = int(np.any(field(b, "iters") != it_ref, axis=1).sum())
    return report
