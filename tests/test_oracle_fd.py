"""Pins the oracle's residual and Jacobian (restating IsoT_FIM::CalRes and
AssembleMatBulksNew, which the reference cannot build here without FASP)
against mathematics: each Jacobian column equals the central finite
difference of the residual when the primary variable (P or N_i of one cell)
is perturbed and the flash, kr/Pc and rock are re-evaluated.

The reference's LBC viscosity derivative drops a term (an integer -(1 / 2)),
so this check uses the oracle variant built with that term restored
(liboracle_exact.so); the default oracle differs from it only there, which
the second test confirms.  It also checks the invariant of the assembly:
the connection part of block (e, b) is the negative of b's diagonal share.
"""
import numpy as np
import pytest

from oracle.oracle_py import EXACT_SO, Oracle
from tests import fim_case

DT = 5.0


def system(orc, case, b):
    nc = case["NC"] + 1
    ncol = nc + 1
    grid = case["grid"]
    res, up, urho, _, _ = orc.calres(nc, grid, b, DT)
    IA, JA = orc.bsr_pattern(grid)
    val = orc.assemble(nc, grid, b, DT, up, urho, IA, JA).reshape(-1, ncol, ncol)
    return -res, IA, JA, val


def props(orc, case, b):
    w = b.copy()
    orc.flash_fim(case["eos"], w)
    orc.calkrpc(case["sat"], w)
    orc.calrock(case["rock"], w)
    return w


@pytest.mark.parametrize("shape,seed", [((3, 3, 3), 5), ((4, 2, 3), 9)])
def test_jacobian_matches_finite_differences(shape, seed):
    case = fim_case.build_case(*shape, seed=seed, dP=2.0, dN=0.001)
    orc = Oracle(EXACT_SO)
    base = props(orc, case, case["bulk"])
    R, IA, JA, val = system(orc, case, base)
    nb = base.nb
    nc = case["NC"] + 1
    ncol = nc + 1
    worst = 0.0
    for cell in range(nb):
        for col in range(ncol):
            xp, xm = base.copy(), base.copy()
            if col == 0:
                h = 1e-4 * base.P[cell]
                xp.P[cell] += h
                xm.P[cell] -= h
            else:
                k = cell * nc + col - 1
                h = 1e-4 * max(base.Ni[k], 1e-3)
                xp.Ni[k] += h
                xm.Ni[k] -= h
            fd = (system(orc, case, props(orc, case, xp))[0] -
                  system(orc, case, props(orc, case, xm))[0]) / (2 * h)
            fd = fd.reshape(nb, ncol)
            # finite-difference noise level: rounding of the residual over 2h
            floor = max(1e-6 * np.abs(fd).max(), 1e-12 * np.abs(R).max() / h)
            for r in range(nb):
                ks = [k for k in range(IA[r], IA[r + 1]) if JA[k] == cell]
                if not ks:
                    assert np.abs(fd[r]).max() <= 1e-8 * np.abs(fd).max()
                    continue
                j = val[ks[0]][:, col]
                scale = max(np.abs(j).max(), np.abs(fd[r]).max(), floor, 1e-300)
                worst = max(worst, np.abs(j - fd[r]).max() / scale)
    assert worst < 1e-4, worst


def test_default_oracle_differs_only_in_viscosity_terms():
    case = fim_case.build_case(3, 3, 3, seed=5)
    a = props(Oracle(), case, case["bulk"])
    e = props(Oracle(EXACT_SO), case, case["bulk"])
    for f in ("S", "xij", "rho", "xi", "mu", "rhoP", "xiP", "muP", "rhox", "xix", "vfi", "vfP",
              "dSec_dPri"):
        assert np.array_equal(getattr(a, f), getattr(e, f)), f
    assert not np.array_equal(a.mux, e.mux)


def test_connection_blocks_are_antisymmetric():
    """Block (e, b) = -(b's diagonal share of the connection), and the
    diagonal is the accumulation block plus the signed shares
    (src/OCPFluidMethod.cpp:1806-1834)."""
    case = fim_case.build_case(4, 3, 2, seed=3)
    orc = Oracle()
    b = case["bulk"]
    R, IA, JA, val = system(orc, case, b)
    nb, ncol = b.nb, case["NC"] + 2
    for n in range(nb):
        diag = val[IA[n]].copy()
        acc = np.zeros((ncol, ncol))
        acc[0, 0] = b.v[n] * b.poroP[n] - b.vfP[n]
        acc[0, 1:] = -b.vfi.reshape(nb, -1)[n]
        for i in range(1, ncol):
            acc[i, i] = 1
        share = np.zeros((ncol, ncol))
        for k in range(IA[n] + 1, IA[n + 1]):
            m = JA[k]
            km = [q for q in range(IA[m], IA[m + 1]) if JA[q] == n][0]
            # the block the neighbour row holds for column n is -(n's share)
            share += -val[km]
        np.testing.assert_allclose(diag, acc + share, rtol=1e-12, atol=1e-12 * np.abs(diag).max())
