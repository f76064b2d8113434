"""The oracle's IsoT_FIMn flash (FlashFIMn / InitFlashFIMn) against the
reference's own, and checks that pin it independently of the reference.

Golden vectors: tests/golden/gen_flash_golden.py --fimn (the reference's
MixtureComp compiled from its sources).  Integer fields exact; floating
fields as in tests/test_oracle_golden.py, plus nj, the secondary-variable
correction res_n and resPc.

Independent pins (what the mathematics fixes):
* FlashFIMn and FlashFIM run the same phase equilibrium, so phase counts,
  labels, S, x_ij, xi, rho, mu and the skip state are identical, bit for bit.
* CalVfiVfp_full03 and CalVfiVfp_full02 are two derivations of the same
  total derivatives dvf/dN, dvf/dP (through dn_ij/dN from the 15 x 15 system
  of CaldXsdXp03 versus the NC x NC system of CalVfiVfp_full02), and so are
  the saturation rows of CaldXsdXpAPI03 and CaldXsdXpAPI02p (dS/dP, dS/dN):
  they agree to rounding.
* The n_ij rows are dn_ij/d(P, N): with x_ij = n_ij / nj they reproduce
  FIM's dx_ij rows by the quotient rule, and each component's rows over the
  phases sum to dN_i/dN_k = delta_ik (mass balance).
* d xi / d n_ij, d rho / d n_ij, d mu / d n_ij (with the exact LBC
  derivative variant) equal central finite differences of xi, rho, mu over
  the phase moles n_ij of the EoS at fixed P, T.
* res_n is the Newton correction of the flash equations: at a converged
  split it is at the level of the split's tolerance (median 1e-10 of the
  cell's moles, nowhere above 1e-6), not O(1).
"""
import numpy as np
import pytest

from opencaeporo_amd import abi
from oracle.oracle_py import Oracle
from tests import parity

TOL = 1e-8
FRAC = 0.97
LOOSE = 1e-4
CASES = ["flashn_spe5.npz", "flashn_ten.npz", "flashn_spe3.npz", "flashn_spe5large.npz",
         "flashn_limits_bench6.npz"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_fimn_flash_matches_reference(orc, name, phase):
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, phase)
    orc.flash_fimn(eos, b)
    rep = parity.compare_state(b, d, f"{phase}_out_", TOL, frac=FRAC, loose=LOOSE)
    assert rep["iters_mismatch_cells"] <= 0.05 * b.nb, rep


def test_fimn_initflash_matches_reference(orc):
    d = parity.load("initflashn_spe5.npz")
    eos = parity.eos_from_fixture(parity.load("flash_spe5.npz"))
    nb = len(d["init_in_P"])
    b = abi.Bulk(nb, int(d["deck_nc"]))
    b.P[:] = d["init_in_P"]
    b.T[:] = d["init_in_T"]
    b.S[:] = d["init_in_S"]
    b.rockVp[:] = d["init_rockVp"]
    orc.init_flash_fimn(eos, b, d["init_z"])
    parity.compare_state(b, d, "init_out_", TOL)
    err = parity.scaled_err(b.Ni.reshape(nb, -1), d["init_out_Ni"].reshape(nb, -1))
    assert err.max() <= TOL


def _both(orc, name, phase):
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, phase)
    f, n = b.copy(), b.copy()
    orc.flash_fim(eos, f)
    orc.flash_fimn(eos, n)
    return eos, f, n


@pytest.mark.parametrize("name", CASES)
def test_fimn_equilibrium_equals_fim(orc, name):
    _, f, n = _both(orc, name, "warm")
    for fld in ["phaseNum", "phaseExist", "S", "xij", "xi", "rho", "mu", "skipFlag", "skipMinEig",
                "bRowSizedSdP", "pVnumCom", "Nt", "vf"]:
        assert np.array_equal(getattr(f, fld), getattr(n, fld)), fld


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_fimn_total_derivatives_equal_fim(orc, name, phase):
    """vfi, vfP and the saturation rows of dSec_dPri: two derivations."""
    _, f, n = _both(orc, name, phase)
    nb, nc = f.nb, f.nc
    ncol = nc + 1
    e = parity.scaled_err(n.vfi.reshape(nb, -1), f.vfi.reshape(nb, -1))
    e = np.maximum(e, np.abs(n.vfP - f.vfP) / np.abs(f.vfP))
    L = abi.max_len_dsdp(nc)
    DF = f.dSec_dPri.reshape(nb, L)
    DN = n.dSec_dPri.reshape(nb, L)
    nS = f.phaseNum  # saturation rows
    for c in range(nb):
        k = nS[c] * ncol
        a, r = DN[c, :k], DF[c, :k]
        e[c] = max(e[c], np.abs(a - r).max() / max(np.abs(r).max(), 1e-300))
    # rounding level, amplified by the two derivative systems' conditioning
    # in the near-critical cells (7% of the CO2 fluid's warm pass, up to
    # 9.3e-9 there)
    assert np.mean(e <= 1e-10) >= 0.9 and e.max() <= 1e-6, (np.mean(e <= 1e-10), e.max())


@pytest.mark.parametrize("name", CASES)
def test_fimn_nij_rows_quotient_rule_and_balance(orc, name):
    """x = n / nj: dx_ij = (dn_ij - x_ij sum_k dn_kj) / nj, FIM's x rows; and
    sum over phases of dn_ij / dN_k = delta_ik, dn_ij / dP summed = 0."""
    _, f, n = _both(orc, name, "warm")
    nb, nc = f.nb, f.nc
    NC, ncol = nc - 1, nc + 1
    L = abi.max_len_dsdp(nc)
    DF = f.dSec_dPri.reshape(nb, L)
    DN = n.dSec_dPri.reshape(nb, L)
    ex = n.phaseExist.reshape(nb, 3)
    checked = 0
    for c in range(nb):
        nS = int(n.phaseNum[c])
        phases = [j for j in range(2) if ex[c, j]]
        sum_rows = np.zeros((NC, ncol))
        for q, j in enumerate(phases):
            base = (nS + q * NC) * ncol
            dn = DN[c, base:base + NC * ncol].reshape(NC, ncol)
            dx = DF[c, base:base + NC * ncol].reshape(NC, ncol)
            nj = n.nj.reshape(nb, 3)[c, j]
            x = n.xij.reshape(nb, 3, nc)[c, j, :NC]
            quot = (dn - x[:, None] * dn.sum(axis=0)[None, :]) / nj
            scale = max(np.abs(dx).max(), 1e-300)
            assert np.abs(quot - dx).max() <= 1e-8 * scale, (c, j, np.abs(quot - dx).max() / scale)
            sum_rows += dn
        if len(phases) == 2:
            want = np.zeros((NC, ncol))
            want[:, 1:1 + NC] = np.eye(NC)
            assert np.abs(sum_rows - want).max() <= 1e-9, (c, np.abs(sum_rows - want).max())
            checked += 1
    assert checked > 0.3 * nb


def test_fimn_single_phase_rows_are_identity(orc):
    """One hydrocarbon phase: n_i = N_i, so dn_i / dN_k = delta_ik, dn/dP = 0
    (CaldXsdXpAPI03's NP = 1 branch) and res_n, resPc are zero."""
    _, f, n = _both(orc, "flashn_spe5.npz", "warm")
    nb, nc = n.nb, n.nc
    NC, ncol = nc - 1, nc + 1
    L = abi.max_len_dsdp(nc)
    DN = n.dSec_dPri.reshape(nb, L)
    one = np.flatnonzero(n.phaseNum == 2)
    assert len(one) > 10
    want = np.zeros((NC, ncol))
    want[:, 1:1 + NC] = np.eye(NC)
    for c in one:
        assert np.array_equal(DN[c, 2 * ncol:(2 + NC) * ncol].reshape(NC, ncol), want)
        assert not n.res_n.reshape(nb, -1)[c].any() and n.resPc[c] == 0


@pytest.mark.parametrize("name", CASES)
def test_fimn_correction_is_rounding_level(orc, name):
    """res_n solves -dF/ds res = F at the converged split: rounding level."""
    _, _, n = _both(orc, name, "warm")
    nb = n.nb
    two = n.phaseNum == 3
    conv = two & (n.nonconv == 0) if n.nonconv is not None else two
    r = np.abs(n.res_n.reshape(nb, -1)).max(axis=1) / n.Nt
    assert np.median(r[conv]) <= 1e-10 and np.all(r[conv] <= 1e-6), (np.median(r[conv]), r[conv].max())
    assert np.all(np.abs(n.resPc[conv]) <= 1e-6 * n.vf[conv])


@pytest.mark.parametrize("name", CASES)
def test_fimn_phase_derivatives_match_finite_differences(name):
    """xix, rhox, mux (by n_ij at fixed P, T) and xiP, rhoP, muP (by P at
    fixed n_ij) of every existing phase against central differences of the
    phase's xi, rho and LBC mu (the exact-LBC variant, whose viscosity
    derivative has no integer-zero term)."""
    from oracle.oracle_py import EXACT_SO
    ex_orc = Oracle(EXACT_SO)
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    b = parity.bulk_from_fixture(d, "warm")
    ex_orc.flash_fimn(eos, b)
    nb, nc = b.nb, b.nc
    NC = nc - 1
    ex = b.phaseExist.reshape(nb, 3)
    xij = b.xij.reshape(nb, 3, nc)
    nj = b.nj.reshape(nb, 3)
    D = {f: getattr(b, f).reshape(nb, 3, nc) for f in ("xix", "rhox", "mux")}
    DP = {f: getattr(b, f).reshape(nb, 3) for f in ("xiP", "rhoP", "muP")}
    worst = 0.0
    for c in range(0, nb, 4):
        P, T = b.P[c], b.T[c]
        for j in range(2):
            if not ex[c, j]:
                continue
            n = xij[c, j, :NC] * nj[c, j]
            props = lambda nn, PP=P: ex_orc.phase_props(eos, PP, T, nn / nn.sum())
            base = props(n)
            for k in range(NC):
                h = 1e-6 * nj[c, j]
                up, dn = n.copy(), n.copy()
                up[k] += h
                dn[k] -= h
                fd = (props(up) - props(dn)) / (2 * h)
                for q, f in enumerate(("xix", "rhox", "mux")):
                    scale = max(np.abs(D[f][c, j, :NC]).max(), 1e-300)
                    err = abs(fd[q] - D[f][c, j, k]) / scale
                    worst = max(worst, err)
                    assert err < 1e-5, (c, j, k, f, fd[q], D[f][c, j, k])
            hP = 1e-6 * P
            fdP = (props(n, P + hP) - props(n, P - hP)) / (2 * hP)
            for q, f in enumerate(("xiP", "rhoP", "muP")):
                err = abs(fdP[q] - DP[f][c, j]) / max(abs(DP[f][c, j]), 1e-300)
                assert err < 1e-5, (c, j, f, fdP[q], DP[f][c, j])
            assert np.allclose(base, [b.xi.reshape(nb, 3)[c, j], b.rho.reshape(nb, 3)[c, j],
                                      b.mu.reshape(nb, 3)[c, j]], rtol=1e-12)
    print(f"worst relative FD disagreement {worst:.1e}")
