"""Benchmark: cell flash + Jacobian updates per second on one MI355X.

Workload (default): the 1,122,000-cell standalone-kernel configuration — a
60 x 220 x 85 orthogonal grid with 7-point TPFA connections (3,329,000),
the 6-component PR fluid (C1, C3, C6, C10, C15, C20; zero BIPs except C1 with
C15/C20 at 0.05), T = 160 F, P ~ U[1000, 5000] psia, z_i ~ Exp(1) * bias
(about half the cells two-phase), Sw ~ U[0.1, 0.4], ln k ~ N(ln 100, 1.5^2),
FP64 throughout, Jacobian in the reference's 8 x 8 BSR blocks (6 hydrocarbon
components + water + pressure).  Synthetic, seeded inputs.

One step = one Newton iteration of the reference's FIM loop on every cell:
restore the iteration's input state (P, N, S, phase count, x_ij, skip memory;
device-to-device), FlashFIM + PassFlashValue, CalKrPc + CalRock, CalRes,
AssembleMat -- by default through ocpg_update_assemble, which overlaps the
flash's slowest cells with the rest of the iteration (--serial: the four
calls one after the other; identical results).  The input state is the "warm" one: the cold flash result with
P perturbed by U[-5, 5] psia.  Inputs are resident in HBM before timing.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from opencaeporo_amd import abi, fluids, lib, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
DT = 5.0
# Algorithmic FP64 work of the warm flash pass per cell (add, sub, mul, div,
# sqrt; transcendental calls counted apart), measured with the oracle's
# counting build (oracle/liboracle_count.so, FlopCounter.count_flash_stages)
# on this exact workload, all 1,122,000 cells: 21,584,690,202 flops and
# 542,748,484 transcendentals in total, of which PhaseSplit 13,922,832,587
# flops and 292,374,158 transcendentals.
WARM_FLOPS_PER_CELL = 21584690202 / 1122000
WARM_SPLIT_FLOPS_PER_CELL = 13922832587 / 1122000
WARM_SPLIT_TRANS_PER_CELL = 292374158 / 1122000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(nx, ny, nz, seed=2024, irregular=False):
    t0 = time.time()
    deck = fluids.ten_deck() if irregular else fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    NC = eos.nc
    grid = (workloads.faulted_grid(nx, ny, nz, seed=seed, frac=0.05) if irregular
            else workloads.cartesian_grid(nx, ny, nz, seed=seed))
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    fields = ["P", "T", "S", "v", "depth", "poroInit"]
    b = abi.Bulk(nb, NC, fields=fields)
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    ctx = lib.Context(eos, sat, rock)
    ctx.set_cells(nb)
    ctx.upload(b)
    ctx.set_conn(grid)
    log(f"[bench] host setup {time.time() - t0:.1f}s, device bytes {ctx.device_bytes() / 1e9:.2f} GB")
    # initial state: rock, InitFlashFIM (Ni from z, Sw, pore volume), kr/Pc
    ctx.calrock()
    ctx.init_flash_fim(z)
    ctx.calkrpc()
    ctx.update_last_step()
    # cold pass: no previous phase information, empty skip memory
    cold = abi.Bulk(nb, NC, fields=["phaseNum", "skipFlag"])
    ctx.upload(cold)
    st = ctx.flash_fim(stats=True)
    log(f"[bench] cold flash: two-phase {st.two_phase / nb:.3f}, ftype {list(st.ftype_count)}, "
        f"iters {list(st.iters)}")
    # warm input: P + U[-5, 5]
    rng = np.random.default_rng(seed + 2)
    pb = abi.Bulk(nb, NC, fields=["P"])
    ctx.download(pb)
    pb.P[:] += rng.uniform(-5.0, 5.0, nb)
    ctx.upload(pb)
    wells = None
    if irregular:
        wells = workloads.irregular_wells(grid, pb.P, NC, seed + 3)
        ctx.set_wells(wells)
    ctx.calkrpc_rock()
    ctx.calres(DT, norms=False)
    ctx.save_flash_inputs()
    ctx.synchronize()
    return ctx, grid, nb, NC, wells


def step(ctx, serial=False):
    ctx.restore_flash_inputs()
    if serial:
        ctx.flash_fim(stats=False)
        ctx.calkrpc_rock()
        ctx.calres(DT, norms=False)
        ctx.assemble(DT)
    else:
        ctx.update_assemble(DT)


def cpu_baseline(nthreads=None, sample_layers=6, seed=2024):
    """The oracle (CPU restatement) on the host: the same warm pass on a
    60 x 220 x sample_layers slice; flash with OpenMP over cells, the rest
    single-threaded as the restatement runs it."""
    from oracle.oracle_py import Oracle
    orc = Oracle()
    nthreads = nthreads or min(os.cpu_count() or 1, 16)
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    NC = eos.nc
    grid = workloads.cartesian_grid(60, 220, sample_layers, seed=seed)
    nb = grid.nb
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = grid.v
    b.depth[:] = grid.depth
    b.poroInit[:] = 0.3
    orc.calrock(rock, b, nthreads)
    orc.init_flash_fim(eos, b, z, nthreads)
    orc.calkrpc(sat, b, nthreads)
    b.lNi[:] = b.Ni
    b.phaseNum[:] = 0
    b.skipFlag[:] = 0
    orc.flash_fim(eos, b, nthreads)
    rng = np.random.default_rng(seed + 2)
    b.P[:] += rng.uniform(-5.0, 5.0, nb)
    orc.calkrpc(sat, b, nthreads)
    orc.calrock(rock, b, nthreads)
    warm = b.copy()
    times = []
    for _ in range(3):
        w = warm.copy()
        t0 = time.perf_counter()
        orc.flash_fim(eos, w, nthreads)
        orc.calkrpc(sat, w, nthreads)
        orc.calrock(rock, w, nthreads)
        nc = NC + 1
        res, up, urho, _, _ = orc.calres(nc, grid, w, DT)
        IA, JA = orc.bsr_pattern(grid)
        orc.assemble(nc, grid, w, DT, up, urho, IA, JA)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return dict(value=nb / t, unit="cell-updates/s", cores=nthreads, kind="port",
                sample=f"oracle warm pass on 60x220x{sample_layers} = {nb} cells "
                       f"(flash + kr/Pc + rock with {nthreads} OpenMP threads, CalRes + "
                       f"assembly 1 thread), median of {len(times)}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nz", type=int, default=85, help="layers of the 60 x 220 x nz grid")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serial", action="store_true",
                    help="one Newton iteration as four serial calls (flash, kr/Pc + rock, CalRes, assembly) "
                         "instead of ocpg_update_assemble")
    ap.add_argument("--irregular", action="store_true",
                    help="the irregular-stencil configuration: 10-component fluid, faulted grid with "
                         "non-neighbour connections, 40 wells (a parity/measurement case, not the headline)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    nx, ny, nz = 60, 220, a.nz
    ctx, grid, nb, NC, wells = setup(nx, ny, nz, irregular=a.irregular)
    for _ in range(a.warmup):
        step(ctx, a.serial)
    ctx.synchronize()
    ctx.timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(ctx, a.serial)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    kern = {k: ctx.timer(k) for k in ("restore", "flash", "flash_stability", "flash_split", "flash_finish",
                                      "flash_tail", "krpc_rock", "calres", "assemble", "fixup",
                                      "update_assemble")}
    ctx.timing(False)
    per = {k: (v[0] / max(v[1], 1)) for k, v in kern.items()}
    log("[bench] per-launch ms: " + ", ".join(f"{k} {v:.3f}" for k, v in per.items()))
    value = nb / (ms_step / 1e3)
    nnzb = ctx.nnzb()
    ncol = NC + 2
    # roofline of the assembly kernel (HBM-bound): algorithmic bytes per launch
    c_bar = grid.nconn / nb
    asm_bytes = nnzb * ncol * ncol * 8 + nnzb * 4 * 3 + (nb + 1) * 4 + nb * 8 * (
        1 + 3 * 3 + 2 * 3 * (NC + 1) + 2 * 9 + 2 + 3 + (3 + 2 * NC) * ncol) * 1.0 + grid.nconn * 3 * 12
    flash_ms = per["flash"]
    launches = {k: v for k, v in per.items() if k not in ("flash", "update_assemble")}
    dominant = max(launches, key=launches.get)
    # roofline of the dominant kernel: the split (FP64-bound; vector and
    # matrix FP64 peaks coincide on MI355X) or the assembly (HBM-bound)
    split_flops = WARM_SPLIT_FLOPS_PER_CELL * nb
    roof_split = dict(kernel="flashSplitKernel", bound="mfma", note="FP64 VALU (no MFMA): vector and matrix "
                      "FP64 peaks coincide on MI355X; flops exclude the transcendental calls",
                      achieved=split_flops / (per["flash_split"] / 1e3) / 1e12, peak=FP64_PEAK_TFLOPS,
                      unit="TFLOP/s", traffic=None, flops_per_launch=split_flops,
                      transcendentals_per_launch=WARM_SPLIT_TRANS_PER_CELL * nb)
    roof_asm = dict(kernel="assembleKernel", bound="hbm", achieved=asm_bytes / (per["assemble"] / 1e3) / 1e9,
                    peak=HBM_PEAK_GBS, unit="GB/s", traffic=None, bytes_per_launch=asm_bytes)
    for r in (roof_split, roof_asm):
        r["frac"] = r["achieved"] / r["peak"]
    roof = roof_split if per["flash_split"] >= per["assemble"] else roof_asm
    other = roof_asm if roof is roof_split else roof_split
    if a.irregular:
        # the flop tally was measured on the 6-component workload only
        roof, other = roof_asm, None
    cpu = None
    if rank == 0 and not a.no_cpu and not a.irregular:
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline is informative only
            log(f"[bench] cpu baseline failed: {e}")
    if a.irregular:
        workload = (f"irregular stencil: {nx}x{ny}x{nz} faulted grid ({grid.nconn} connections incl. "
                    f"non-neighbour), 10-comp PR, {wells.nw} wells / {wells.nperf} perforations, warm "
                    f"Newton pass, {ncol}x{ncol} BSR")
    else:
        workload = (f"standalone flash+assembly, {nx}x{ny}x{nz} grid, {nb} cells, "
                    f"{grid.nconn} connections, 6-comp PR, warm Newton pass, 8x8 BSR")
    out = {
        "metric": ("cell flash+Jacobian updates/s (1.1M cells, 10-comp, irregular, wells)" if a.irregular
                   else "cell flash+Jacobian updates/s (1.1M cells, 6-comp)"),
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "cells": nb, "connections": grid.nconn, "block": ncol,
                   "kernel_ms": per, "dominant_kernel": dominant, "flash_ms": flash_ms,
                   "flash_flops_per_cell": WARM_FLOPS_PER_CELL, "asm_bytes_per_launch": asm_bytes,
                   "c_bar": c_bar, "roofline_other": other},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
