"""The oracle's residual and assembly computed by rows on several OpenMP
threads (the CPU baseline's parallel form) equal the one-thread result
byte for byte: every row sums its connections' terms in the reference's
connection order whatever the thread count."""
import numpy as np
import pytest

from tests import fim_case


@pytest.mark.parametrize("grid", [None, "faulted"])
def test_threads_do_not_change_the_system(grid):
    case = fim_case.build_case(12, 10, 6, seed=31, grid=grid)
    b, orc, g = case["bulk"], case["orc"], case["grid"]
    nc = case["NC"] + 1
    IA, JA = orc.bsr_pattern(g)
    out = []
    for nt in (1, 3, 8):
        res, up, urho, uvel, norms = orc.calres(nc, g, b, 5.0, nthreads=nt)
        val = orc.assemble(nc, g, b, 5.0, up, urho, IA, JA, nthreads=nt)
        out.append((res, up, urho, uvel, val, (norms.maxRelRes_V, norms.maxRelRes_N, norms.maxId_V, norms.maxId_N)))
    for o in out[1:]:
        for a, r in zip(o[:5], out[0][:5]):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(r).view(np.uint8))
        assert o[5] == out[0][5]
