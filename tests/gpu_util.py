"""Helpers of the GPU tests: run the HIP path and the oracle on the same
inputs."""
import sys

import numpy as np

from opencaeporo_amd import abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# Floating-point parity of the HIP path against the FP64 oracle: per-cell
# scaled error 1e-10, and in every cell at most max(1e-10, SENS_FACTOR x the
# cell's rounding sensitivity) (parity.rounding_sensitivity: the oracle's own
# response to a 1-ulp change of P).  Near-critical cells amplify last-bit
# differences (GPU libm vs glibc) far beyond 1e-10; the oracle and the
# reference's MKL-backed flash differ there by the same amount.  At least
# GPU_FRAC of the cells must meet the flat 1e-10, and all of them GPU_LOOSE.
GPU_TOL = 1e-10
GPU_FRAC = 0.9
GPU_LOOSE = 1e-4
SENS_FACTOR = 10.0


def cell_bounds(sens):
    return np.maximum(GPU_TOL, SENS_FACTOR * sens)


def gpu_flash(ctx, bulk, init_z=None, method=None):
    """Upload ``bulk``, run the flash (or InitFlash with ``init_z``), download
    into a copy.  Returns (bulk_out, stats).  method: abi.METHOD_FIMN for
    the IsoT_FIMn flash (set before the cells)."""
    out = bulk.copy()
    if method is not None:
        ctx.set_method(method)
    ctx.set_cells(bulk.nb)
    ctx.upload(bulk)
    stats = None
    if init_z is None:
        stats = ctx.flash_fim(stats=True)
    else:
        ctx.init_flash_fim(init_z)
    ctx.download(out)
    return out, stats


def snapshot_dict(b, prefix, fimn=False):
    """Fixture-like dict of a bulk's state fields (for parity.compare_state);
    fimn: also IsoT_FIMn's nj, res_n and resPc."""
    from tests import parity
    d = {}
    for f in parity.STATE_FIELDS + ["nonconv"] + (parity.FIMN_FIELDS if fimn else []):
        a = getattr(b, f)
        if f == "dSec_dPri":
            L = abi.max_len_dsdp(b.nc)
            used = int(b.bRowSizedSdP.max()) * (b.nc + 1)
            a = a.reshape(b.nb, L)[:, :used]
        d[prefix + f] = a.copy()
    return d
