"""Failure paths of the flash on the oracle (CPU): a NaN in a cell's input
ends the pass normally, the cell is flagged (nonconv = 2: results not
finite) and every other cell's result is bit for bit the clean pass's; the
iteration caps' non-convergence flags equal the reference's own (ePEC =
1e8, from the goldens made with caps in tests/golden/gen_flash_golden.py)."""
import numpy as np
import pytest

from oracle.oracle_py import Oracle
from tests import parity

FIELDS = ("phaseNum", "phaseExist", "S", "xij", "rho", "xi", "mu", "dSec_dPri", "skipFlag", "skipMinEig",
          "nonconv", "iters")


@pytest.mark.parametrize("phase", ["cold", "warm"])
def test_nan_cells_are_flagged_and_contained(phase):
    d = parity.load("flash_bench6.npz")
    eos = parity.eos_from_fixture(d)
    orc = Oracle()
    b = parity.bulk_from_fixture(d, phase)
    clean = b.copy()
    orc.flash_fim(eos, clean)
    bad = b.copy()
    nc = bad.nc
    bad.Ni.reshape(-1, nc)[5, 2] = np.nan   # a moles entry
    bad.P[17] = np.nan                      # a pressure
    orc.flash_fim(eos, bad)
    assert list(bad.nonconv[[5, 17]]) == [2, 2]
    assert np.all(clean.nonconv != 2)
    others = np.ones(b.nb, bool)
    others[[5, 17]] = False
    for f in FIELDS:
        a = getattr(bad, f).reshape(b.nb, -1)[others]
        c = getattr(clean, f).reshape(b.nb, -1)[others]
        assert np.array_equal(a, c), f


@pytest.mark.parametrize("name", ["flash_limits_bench6.npz", "flash_maxit0_bench6.npz"])
def test_capped_iterations_flag_what_the_reference_flags(name):
    d = parity.load(name)
    eos = parity.eos_from_fixture(d)
    orc = Oracle()
    for phase in ("cold", "warm"):
        b = parity.bulk_from_fixture(d, phase)
        orc.flash_fim(eos, b)
        want = d[f"{phase}_out_nonconv"]
        assert np.array_equal(b.nonconv, want)
        # iteration paths agree in the large majority of cells (the LAPACK
        # layer decides a few convergence tests differently, as without caps)
        diff = np.any(b.iters.reshape(-1, 5) != d[f"{phase}_out_iters"].reshape(-1, 5), axis=1)
        assert diff.sum() <= 0.05 * b.nb
    if "maxit0" in name:
        # every cap at 0: no cell whose flash runs converges (the skipped
        # cells of the warm pass have nothing to converge)
        cold = parity.bulk_from_fixture(d, "cold")
        orc.flash_fim(eos, cold)
        assert np.all(cold.nonconv == 1)
