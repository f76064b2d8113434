"""Helpers of the GPU tests: run the HIP path and the oracle on the same
inputs."""
import sys

import numpy as np

from opencaeporo_amd import abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# Floating-point parity of the HIP path against the FP64 oracle: per-cell
# scaled error 1e-10 in at least 99% of cells; the remainder are near-critical
# cells whose split Newton system is ill-conditioned (the same cells where the
# oracle and the reference's MKL-backed flash differ), bounded loosely.
GPU_TOL = 1e-10
GPU_FRAC = 0.99
GPU_LOOSE = 1e-4


def gpu_flash(ctx, bulk, init_z=None):
    """Upload ``bulk``, run the flash (or InitFlash with ``init_z``), download
    into a copy.  Returns (bulk_out, stats)."""
    out = bulk.copy()
    ctx.set_cells(bulk.nb)
    ctx.upload(bulk)
    stats = None
    if init_z is None:
        stats = ctx.flash_fim(stats=True)
    else:
        ctx.init_flash_fim(init_z)
    ctx.download(out)
    return out, stats


def snapshot_dict(b, prefix):
    """Fixture-like dict of a bulk's state fields (for parity.compare_state)."""
    from tests import parity
    d = {}
    for f in parity.STATE_FIELDS:
        a = getattr(b, f)
        if f == "dSec_dPri":
            L = abi.max_len_dsdp(b.nc)
            used = int(b.bRowSizedSdP.max()) * (b.nc + 1)
            a = a.reshape(b.nb, L)[:, :used]
        d[prefix + f] = a.copy()
    return d
