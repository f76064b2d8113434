"""ocpg_mlp_gemm_noalloc (the split-K product with its partial tiles in the
library's static per-device buffer) against ocpg_mlp_gemm with a caller
workspace: bit-identical results on two calls in a row, the second smaller
than the first, and the size limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from opencaeporo_amd import abi, lib
from tests import gpu_util

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 1 << 22   # OCPG_MLP_SCRATCH_FLOATS


def _gemm_pair(t, L, stream, M, N, K, splitK, rng):
    A = t.tensor(rng.normal(size=(M, K)).astype(np.float32), device="cuda:0")
    B = t.tensor(rng.normal(size=(K, N)).astype(np.float32), device="cuda:0")
    C0 = t.tensor(rng.normal(size=(M, N)).astype(np.float32), device="cuda:0")
    want = C0.clone()
    got = C0.clone()
    part = t.zeros(splitK * M * N, dtype=t.float32, device="cuda:0")
    p = lambda x: C.c_void_p(x.data_ptr())
    rc = L.ocpg_mlp_gemm(stream, M, N, K, p(A), K, 1, p(B), N, 1, p(want), N, 1, 1, splitK, p(part))
    assert rc == abi.OK
    rc = L.ocpg_mlp_gemm_noalloc(stream, M, N, K, p(A), K, 1, p(B), N, 1, p(got), N, 1, 1, splitK)
    assert rc == abi.OK
    t.cuda.synchronize()
    return want.cpu().numpy(), got.cpu().numpy()


def test_noalloc_gemm_matches_workspace_gemm():
    import torch as t
    L = lib.load()
    stream = C.c_void_p(t.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(11)
    # the second call is smaller: stale partial tiles of the first must not leak in
    for (M, N, K, splitK) in [(300, 96, 4096, 8), (37, 20, 700, 3), (64, 64, 64, 1)]:
        gpu_util.say(f"[no-alloc] M={M} N={N} K={K} splitK={splitK}")
        want, got = _gemm_pair(t, L, stream, M, N, K, splitK, rng)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32))
    # Z * M * N above the buffer: refused on the host, nothing launched
    M = N = 1024
    A = t.zeros((M, 256), dtype=t.float32, device="cuda:0")
    Bm = t.zeros((256, N), dtype=t.float32, device="cuda:0")
    Cm = t.zeros((M, N), dtype=t.float32, device="cuda:0")
    p = lambda x: C.c_void_p(x.data_ptr())
    rc = L.ocpg_mlp_gemm_noalloc(stream, M, N, 256, p(A), 256, 1, p(Bm), N, 1, p(Cm), N, 1, 0, 8)
    assert 8 * M * N > SCRATCH_FLOATS and rc == abi.ETOOBIG
    # splitK = 1 needs no buffer whatever the size
    rc = L.ocpg_mlp_gemm_noalloc(stream, M, N, 256, p(A), 256, 1, p(Bm), N, 1, p(Cm), N, 1, 0, 1)
    assert rc == abi.OK
    t.cuda.synchronize()
