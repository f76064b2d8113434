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
FP64_PEAK_TFS = 78.6      # MI355X FP64 vector
DT = 5.0
# Algorithmic FP64 work of the warm flash pass per cell (add, sub, mul, div,
# sqrt; transcendental calls counted apart), measured with the oracle's
# counting build (oracle/liboracle_count.so, FlopCounter.count_flash_stages)
# on this exact workload, all 1,122,000 cells: 21,584,690,202 flops and
# 542,748,484 transcendentals in total.
WARM_FLOPS_PER_CELL = 21584690202 / 1122000
# HBM bytes per launch of the assembly and of the CalRes + record-packing
# launch on the default workload: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE,
# scaled by what the counters read for a device-to-device copy of known size
# (profiles/assembly_calres_1p1M_6comp_traffic.md)
ASM_TRAFFIC_BYTES = 3.840e9 + 3.983e9
CALRES_PACK_TRAFFIC_BYTES = 6.149e9 + 3.259e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(nx, ny, nz, seed=2024, irregular=False, model=None, method=abi.METHOD_FIM, equilibrate=True):
    """model: a deck.Model (the workload read from an input deck) instead of
    the synthetic generators, started from its hydrostatic equilibration
    (Bulk::InitPTSw, with the library's phase densities) when it has ZMFVD
    and equilibrate is set.  method: abi.METHOD_FIMN for IsoT_FIMn's flash
    and assembly (InitFlashFIMn for the initial state)."""
    t0 = time.time()
    if model is not None:
        eos = lib.eos_from_deck(model.eos_deck)
        sat, rock, grid = model.sat, model.rock, model.grid
        if equilibrate and "ZMFVD" in model.deck.tables:
            c0 = lib.Context(eos)
            model.equilibrate(lambda P, T, z, ph: c0.phase_density(P, T, z, ph)[1])
            c0.close()
            log(f"[bench] equilibrated: P in [{model.bulk.P.min():.1f}, {model.bulk.P.max():.1f}] psia")
    else:
        deck = fluids.ten_deck() if irregular else fluids.bench6_deck()
        eos = lib.eos_from_deck(deck)
        sat = fluids.make_satfunc()
        rock = fluids.make_rock()
        grid = (workloads.faulted_grid(nx, ny, nz, seed=seed, frac=0.05) if irregular
                else workloads.cartesian_grid(nx, ny, nz, seed=seed))
    NC = eos.nc
    nb = grid.nb
    fields = ["P", "T", "S", "v", "depth", "poroInit"]
    b = abi.Bulk(nb, NC, fields=fields)
    if model is not None:
        z = model.z
        for f in fields:
            getattr(b, f)[:] = getattr(model.bulk, f)
    else:
        P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
        b.P[:] = P
        b.T[:] = workloads.T_F
        b.S.reshape(nb, 3)[:, 2] = Sw
        b.v[:] = grid.v
        b.depth[:] = grid.depth
        b.poroInit[:] = 0.3
    ctx = lib.Context(eos, sat, rock)
    if method != abi.METHOD_FIM:
        ctx.set_method(method)
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
    if model is not None:
        wells = model.wells
    elif irregular:
        wells = workloads.irregular_wells(grid, pb.P, NC, seed + 3)
    if wells is not None:
        ctx.set_wells(wells)
    ctx.calkrpc_rock()
    ctx.calres(DT, norms=False)
    ctx.save_flash_inputs()
    ctx.synchronize()
    return ctx, grid, nb, NC, wells


def fimn_leg(nx, ny, nz, steps, warmup):
    """IsoT_FIMn on the same field: the initial state from InitFlashFIMn,
    then warm Newton iterations (FlashFIMn + PassFlashValue, kr/Pc + rock,
    CalRes, IsoT_FIMn::AssembleMat with its right-hand side), timed like the
    headline: the overlapped iteration (ocpg_update_assemble), then the
    serial calls."""
    ctx, grid, nb, NC, _ = setup(nx, ny, nz, method=abi.METHOD_FIMN)
    out = {"method": "IsoT_FIMn", "steps": steps}
    for serial in (False, True):
        for _ in range(warmup):
            step(ctx, serial)
        ctx.synchronize()
        ctx.timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(ctx, serial)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        kern = {k: ctx.timer(k) for k in ("restore", "flash", "flash_stability", "flash_split", "flash_finish",
                                          "krpc_rock", "calres", "pack", "assemble", "fixup")}
        ctx.timing(False)
        per = {k: (v[0] / max(v[1], 1)) for k, v in kern.items()}
        form = "serial" if serial else "overlapped"
        log(f"[bench] IsoT_FIMn {form}: {ms:.2f} ms per step; per-launch ms: " +
            ", ".join(f"{k} {v:.3f}" for k, v in per.items()))
        if serial:
            out["serial_ms_per_step"] = ms
            out["serial_kernel_ms"] = per
        else:
            out["ms_per_step"] = ms
            out["cell_updates_per_s"] = nb / (ms / 1e3)
            out["kernel_ms"] = per
    ctx.close()
    return out


def learned_leg(ctx, nb, NC, reps=3):
    """The learned form on the benchmark's own field: the flash alone
    (ocpg_flash_fim) against ocpg_flash_fim_learned (the network's fused
    inference from the cells' P, T and moles, then the guided flash), cold
    (no previous phase information) and warm (the timed Newton pass's
    input), with the network trained on the GPU on 10M flash results
    (python -m opencaeporo_amd.learned --out ...).  Medians of reps timed
    passes; the agreement of the converged results."""
    from opencaeporo_amd import learned
    path = learned.weights_path(NC)
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    ctx.set_learned([d[f"W{l}"] for l in range(4)], [d[f"b{l}"] for l in range(4)])
    cold = abi.Bulk(nb, NC, fields=["phaseNum", "skipFlag"])
    out = {"weights": os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))}
    for kind in ("cold", "warm"):
        res, got = {}, {}
        for mode in ("plain", "learned"):
            ts = []
            for rep in range(reps + 1):
                ctx.restore_flash_inputs()
                if kind == "cold":
                    ctx.upload(cold)
                ctx.synchronize()
                ctx.timing(rep == reps)
                t0 = time.perf_counter()
                st = ctx.flash_fim(stats=True) if mode == "plain" else ctx.flash_fim_learned(stats=True)
                ctx.synchronize()
                if rep > 0:
                    ts.append((time.perf_counter() - t0) * 1e3)
            stages = {k: ctx.timer(k)[0] for k in ("flash_stability", "flash_split", "flash_finish", "mlp")}
            ctx.timing(False)
            o = abi.Bulk(nb, NC, fields=["phaseNum", "S", "xij"])
            ctx.download(o)
            got[mode] = o
            res[mode] = {"flash_ms": float(np.median(ts)), "stage_ms": stages,
                         "iters": dict(zip(["ssm_sta", "nr_sta", "ssm_sp", "nr_sp", "rr"], [int(v) for v in st.iters])),
                         "two_phase": int(st.two_phase), "nonconverged": int(st.nonconverged)}
            log(f"[bench] learned leg, {kind} {mode}: {res[mode]['flash_ms']:.2f} ms, iterations {res[mode]['iters']}")
        same = got["plain"].phaseNum == got["learned"].phaseNum
        dS = np.abs(got["plain"].S.reshape(nb, 3)[same] - got["learned"].S.reshape(nb, 3)[same])
        res["agreement"] = {"phase_count_differs": int((~same).sum()), "max_abs_S_diff": float(dS.max()),
                            "p99_abs_S_diff": float(np.quantile(dS.max(axis=1), 0.99))}
        res["gain"] = res["plain"]["flash_ms"] / res["learned"]["flash_ms"]
        res["inference_ms"] = res["learned"]["stage_ms"]["mlp"]
        out[kind] = res
    ctx.set_learned(None)
    return out


def step(ctx, serial=False):
    ctx.restore_flash_inputs()
    step_body(ctx, serial)


def step_body(ctx, serial=False):
    if serial:
        ctx.flash_fim(stats=False)
        ctx.calkrpc_rock()
        ctx.calres(DT, norms=False)
        ctx.assemble(DT)
    else:
        ctx.update_assemble(DT)


# Fields a warm Newton pass starts from (what ocpg_save_flash_inputs keeps,
# plus the static cell data and the last step's moles CalRes reads).
PASS_INPUTS = ["P", "T", "Ni", "lNi", "S", "phaseNum", "xij", "skipFlag", "skipMinEig", "skipP", "skipT", "skipZi",
               "v", "depth", "poroInit"]


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))
    return model, os.cpu_count() or 1, max(threads, 1)


def cpu_baseline(ctx, grid, NC, reps=5, slice_layers=2, seed=2024):
    """The oracle (CPU restatement, OpenMP over cells for the flash and over
    rows for CalRes and the assembly) timed on the host on the GPU's own warm
    input state, downloaded from the device: every host thread the process
    may use (the box grants OMP_NUM_THREADS) on all cells, and one thread on
    the first slice_layers layers (the same cell distribution; the all-thread
    rate on that slice is reported beside it).  A cold pass is the same input
    with no previous phase information and empty skip memory."""
    from oracle.oracle_py import Oracle
    orc = Oracle()
    eos = ctx.eos
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    model, nproc, threads = host_cpu()
    nb = grid.nb
    nc = NC + 1
    ctx.restore_flash_inputs()
    warm = abi.Bulk(nb, NC)
    inp = abi.Bulk(nb, NC, fields=PASS_INPUTS)
    ctx.download(inp)
    for f in PASS_INPUTS:
        getattr(warm, f)[:] = getattr(inp, f)
    del inp
    IA, JA = orc.bsr_pattern(grid)
    val = np.zeros(int(IA[-1]) * (nc + 1) ** 2)

    def sub(b, n1):
        """The first n1 cells (whole layers) with their own connections."""
        s = abi.Bulk(n1, NC)
        for k, a in b.arrays.items():
            per = a.size // nb
            s.arrays[k][:] = a[:n1 * per]
        m = grid.eId < n1
        g = type("G", (), {})()
        g.nb, g.bId, g.eId, g.area = n1, grid.bId[m], grid.eId[m], grid.area[m]
        g.nconn = int(m.sum())
        return s, g

    def timed(b0, g, nt, cold, n_rep, vbuf):
        IA_, JA_ = (IA, JA) if g is grid else orc.bsr_pattern(g)
        if vbuf is None or vbuf.size != int(IA_[-1]) * (nc + 1) ** 2:
            vbuf = np.zeros(int(IA_[-1]) * (nc + 1) ** 2)
        times = []
        for _ in range(n_rep):
            w = b0.copy()
            if cold:
                w.phaseNum[:] = 0
                w.skipFlag[:] = 0
            t0 = time.perf_counter()
            orc.flash_fim(eos, w, nt)
            orc.calkrpc(sat, w, nt)
            orc.calrock(rock, w, nt)
            res, up, urho, _, _ = orc.calres(nc, g, w, DT, nthreads=nt)
            orc.assemble(nc, g, w, DT, up, urho, IA_, JA_, nthreads=nt, out=vbuf)
            times.append(time.perf_counter() - t0)
        return g.nb / float(np.median(times)), times

    warm_all, t_wa = timed(warm, grid, threads, False, reps, val)
    n1 = 60 * 220 * slice_layers
    sb, sg = sub(warm, n1)
    del val
    cold_all, t_ca = timed(sb, sg, threads, True, reps, None)
    warm_slice_all, _ = timed(sb, sg, threads, False, reps, None)
    warm_one, t_w1 = timed(sb, sg, 1, False, reps, None)
    cold_one, t_c1 = timed(sb, sg, 1, True, reps, None)
    return dict(value=warm_all, unit="cell-updates/s", cores=threads, kind="port",
                sample=(f"oracle warm Newton pass (FlashFIM + kr/Pc + rock + CalRes + assembly) on the GPU's own "
                        f"warm input, all {nb} cells with {threads} OpenMP threads, median of {reps}; the cold "
                        f"pass and the one-thread rates on the first {slice_layers} layers ({n1} cells)"),
                cpu_model=model, nproc=nproc, threads=threads,
                warm_all_threads=warm_all, cold_all_threads_slice=cold_all,
                warm_all_threads_slice=warm_slice_all, warm_one_thread_slice=warm_one,
                cold_one_thread_slice=cold_one, reps=reps,
                seconds=dict(warm_all=t_wa, cold_all_slice=t_ca, warm_one_slice=t_w1, cold_one_slice=t_c1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nz", type=int, default=85, help="layers of the 60 x 220 x nz grid")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--handoff-steps", type=int, default=3,
                    help="steps timed again with the BSR copied to page-locked host arrays (PCIe-inclusive rate)")
    ap.add_argument("--cold-steps", type=int, default=3,
                    help="cold passes (no previous phase information) timed after the warm steps")
    ap.add_argument("--serial", action="store_true",
                    help="one Newton iteration as four serial calls (flash, kr/Pc + rock, CalRes, assembly) "
                         "instead of ocpg_update_assemble")
    ap.add_argument("--irregular", action="store_true",
                    help="the irregular-stencil configuration: 10-component fluid, faulted grid with "
                         "non-neighbour connections, 40 wells (a parity/measurement case, not the headline)")
    ap.add_argument("--serving-requests", type=int, default=384,
                    help="requests of the serving-form leg (p50/p99 latency and capacity, continuous batching "
                         "vs one at a time, 256 realisations of 1e3..1e6 cells); 0 skips it")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-form leg")
    ap.add_argument("--no-fimn", action="store_true", help="skip the IsoT_FIMn leg")
    ap.add_argument("--no-equil", action="store_true",
                    help="with --deck: start every cell at the datum pressure instead of the hydrostatic "
                         "equilibration")
    ap.add_argument("--fimn-only", action="store_true",
                    help="time only the IsoT_FIMn leg (--steps, --warmup) and print its record")
    ap.add_argument("--deck", default=None,
                    help="read the workload from an input deck (opencaeporo_amd/deck.py), e.g. "
                         "tests/golden/decks/spe5refine/spe5-70x70x30-2y.data (a measurement case, not the headline)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (across GPUs each would "
              "take a slab of cells and exchange the records of its boundary layer once per Newton "
              "iteration)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    nx, ny, nz = 60, 220, a.nz
    if a.fimn_only:
        print(json.dumps(fimn_leg(nx, ny, nz, a.steps, a.warmup)), flush=True)
        return
    model = None
    if a.deck:
        from opencaeporo_amd import deck
        model = deck.load(a.deck)
        nx, ny, nz = model.deck.dims
    side_case = a.irregular or model is not None
    ctx, grid, nb, NC, wells = setup(nx, ny, nz, irregular=a.irregular, model=model, equilibrate=not a.no_equil)
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
                                      "flash_tail", "flash_restarts", "krpc_rock", "calres", "pack", "assemble",
                                      "fixup", "update_assemble")}
    ctx.timing(False)
    per = {k: (v[0] / max(v[1], 1)) for k, v in kern.items()}
    # the same iteration with the system handed to the host (FASP's
    # dBSRmat arrays, page-locked) each step: the PCIe-inclusive rate
    handoff_ms = None
    if not a.serial and a.handoff_steps > 0:
        arrays = ctx.pinned_bsr()
        ctx.synchronize()
        t0h = time.perf_counter()
        for _ in range(a.handoff_steps):
            step(ctx, False)
            ctx.get_bsr_async(arrays)
        ctx.synchronize()
        handoff_ms = (time.perf_counter() - t0h) * 1e3 / a.handoff_steps
        ctx.free_pinned(arrays)
        log(f"[bench] with the BSR handed to pinned host arrays: {handoff_ms:.2f} ms per step")
    log("[bench] per-launch ms: " + ", ".join(f"{k} {v:.3f}" for k, v in per.items()))
    # the cold pass (no previous phase information, empty skip memory: every
    # cell takes the full stability test) on the same input, timed per step
    cold_ms = None
    if not side_case and a.cold_steps > 0:
        flags = abi.Bulk(nb, NC, fields=["phaseNum", "skipFlag"])
        ts = []
        for _ in range(a.cold_steps):
            ctx.restore_flash_inputs()
            ctx.upload(flags)
            ctx.synchronize()
            t0c = time.perf_counter()
            step_body(ctx, a.serial)
            ctx.synchronize()
            ts.append((time.perf_counter() - t0c) * 1e3)
        cold_ms = float(np.median(ts))
        log(f"[bench] cold pass: {cold_ms:.2f} ms per step (median of {len(ts)})")
    learned_res = None
    if not side_case and not a.no_learned:
        try:
            learned_res = learned_leg(ctx, nb, NC)
        except Exception as e:  # informative only
            log(f"[bench] learned leg failed: {e}")
    value = nb / (ms_step / 1e3)
    nnzb = ctx.nnzb()
    nc, ncol = NC + 1, NC + 2
    # compact dSec_dPri rows per cell (3 + 2 NC with two hydrocarbon phases,
    # 2 + NC with one) from the last iteration's state
    rb = abi.Bulk(nb, NC, fields=["bRowSizedSdP"])
    ctx.download(rb)
    rows_mean = float(rb.bRowSizedSdP.mean())
    # Algorithmic bytes of one launch of the assembly kernel (every value it
    # needs read once, every block written once): the BSR blocks, the
    # pattern (JA, connection and mirror slot per block; IA, bulkEnd per
    # row), each cell's neighbour record (10 + 3 nc doubles: Pj, xi, kr/mu
    # per phase, xij, depth; three phase flags), its own-row record (dSec_dPri
    # rows x ncol, dPcj_dS, vfi, vfP, v, poroP, its derivatives xiP, rhoP,
    # muP/mu, xi/mu, dKr_dS, rhox, xix, mux/mu; 7 counts) and per connection
    # Akd and its upwinding (3 cells, 3 densities).
    rec_bytes = (10 + 3 * nc) * 8 + 3
    own_bytes = (rows_mean * ncol + 9 + nc + 3 + 21 + 9 * nc) * 8 + 7 * 4
    asm_bytes = (nnzb * ncol * ncol * 8 + nnzb * 12 + nb * 8 + nb * (rec_bytes + own_bytes)
                 + grid.nconn * (8 + 3 * 4 + 3 * 8))
    # The record packing reads the same cell values from the
    # structure-of-arrays state (mu and xi once) and writes the records
    # (doubles); CalRes reads each cell's pore and fluid volume, moles, last
    # moles, densities, pressures, depth, flags and upwind properties once,
    # writes the residual and per connection the upwinding (3 cells, 3
    # densities, 3 velocities) from the area.
    pack_bytes = (nb * (rec_bytes + own_bytes - 6 * 8)
                  + nb * ((10 + 3 * nc + 3) + rows_mean * ncol + 9 + nc + 10 + 21 + 9 * nc) * 8)
    calres_bytes = (nb * ((2 + 2 * nc + 3 + 3 + 1 + 9 + 3 * nc) * 8 + 3 + (nc + 1) * 8)
                    + grid.nconn * (8 + 3 * 4 + 3 * 8 + 3 * 8))
    flash_ms = per["flash"]
    launches = {k: v for k, v in per.items() if k not in ("flash", "update_assemble", "flash_stability",
                                                          "flash_split", "flash_finish", "flash_tail",
                                                          "flash_restarts")}
    dominant = max(launches, key=launches.get)
    # The dominant kernel: the assembly, the longest launch on the main
    # stream (the flash stages are several launches each, and the side
    # streams' kernels run beside the main stream).  Its bound by algorithmic
    # work is HBM; it reaches a fraction of it because each row's wavefront
    # works through its connections one after the other (~1,190 VALU and
    # ~990 SALU instructions per row, rocprofv3 SQ_INSTS_VALU / SQ_INSTS_SALU)
    # at five wavefronts per SIMD.
    roof = dict(kernel="assembleKernel", bound="hbm", achieved=asm_bytes / (per["assemble"] / 1e3) / 1e9,
                peak=HBM_PEAK_GBS, unit="GB/s", traffic=None, bytes_per_launch=asm_bytes)
    if a.serial:
        # separate launches: CalRes (ocpg_calres), the packing (ocpg_assemble)
        roof_other = [dict(kernel="calResKernel", bound="hbm", bytes_per_launch=calres_bytes, ms=per["calres"]),
                      dict(kernel="packRecordsKernel", bound="hbm", bytes_per_launch=pack_bytes, ms=per["pack"])]
    else:
        roof_other = [dict(kernel="calResPackKernel", bound="hbm", bytes_per_launch=calres_bytes + pack_bytes,
                           ms=per["calres"])]
    other = []
    for r in roof_other:
        ms = r.pop("ms")
        if ms > 0:
            r.update(achieved=r["bytes_per_launch"] / (ms / 1e3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                     traffic=None)
            r["frac"] = r["achieved"] / r["peak"]
            other.append(r)
    roof["frac"] = roof["achieved"] / roof["peak"]
    if not side_case and nz == 85 and not a.serial:
        # HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
        # on this workload, scaled by the copy-kernel calibration
        # (profiles/assembly_calres_1p1M_6comp_traffic.md)
        roof["traffic"] = ASM_TRAFFIC_BYTES
        for r in other:
            if r["kernel"] == "calResPackKernel":
                r["traffic"] = CALRES_PACK_TRAFFIC_BYTES
    # the flash against FP64 vector peak (its algorithmic flops per warm cell,
    # over the main stream's flash time), and the whole step against the
    # step's roofline time: flash flops at FP64 peak plus the assembly's
    # algorithmic bytes at HBM peak
    flash_roof = step_roof = None
    if not side_case and nz == 85:
        flops = WARM_FLOPS_PER_CELL * nb
        flash_roof = dict(bound="fp64", flops_per_launch=flops, achieved=flops / (per["flash"] / 1e3) / 1e12,
                          peak=FP64_PEAK_TFS, unit="TFLOP/s")
        flash_roof["frac"] = flash_roof["achieved"] / flash_roof["peak"]
        t_roof = flops / (FP64_PEAK_TFS * 1e12) + asm_bytes / (HBM_PEAK_GBS * 1e9)
        step_roof = dict(t_roof_ms=t_roof * 1e3, t_measured_ms=ms_step, frac=t_roof * 1e3 / ms_step)
    cpu = None
    if rank == 0 and not a.no_cpu and not side_case:
        try:
            t0c = time.time()
            cpu = cpu_baseline(ctx, grid, NC)
            log(f"[bench] cpu baseline ({cpu['threads']} threads, {cpu['cpu_model']}): warm "
                f"{cpu['warm_all_threads']:.4g} cell-updates/s all threads, one thread "
                f"{cpu['warm_one_thread_slice']:.4g} ({time.time() - t0c:.1f}s)")
        except Exception as e:  # the baseline is informative only
            log(f"[bench] cpu baseline failed: {e}")
    if model is not None:
        workload = (f"input deck {os.path.basename(a.deck)}: {nx}x{ny}x{nz} grid, {nb} cells, "
                    f"{grid.nconn} connections, {NC}-comp PR, {wells.nw if wells else 0} wells / "
                    f"{wells.nperf if wells else 0} perforations, warm Newton pass, {ncol}x{ncol} BSR")
    elif a.irregular:
        workload = (f"irregular stencil: {nx}x{ny}x{nz} faulted grid ({grid.nconn} connections incl. "
                    f"non-neighbour), 10-comp PR, {wells.nw} wells / {wells.nperf} perforations, warm "
                    f"Newton pass, {ncol}x{ncol} BSR")
    else:
        workload = (f"standalone flash+assembly, {nx}x{ny}x{nz} grid, {nb} cells, "
                    f"{grid.nconn} connections, 6-comp PR, warm Newton pass, 8x8 BSR")
    out = {
        "metric": ("cell flash+Jacobian updates/s (input deck)" if model is not None
                   else "cell flash+Jacobian updates/s (1.1M cells, 10-comp, irregular, wells)" if a.irregular
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
        "data": "input deck" if model is not None else "synthetic",
        "config": {"workload": workload,
                   "cells": nb, "connections": grid.nconn, "block": ncol,
                   "kernel_ms": per, "dominant_kernel": dominant, "flash_ms": flash_ms,
                   "flash_flops_per_cell": WARM_FLOPS_PER_CELL, "asm_bytes_per_launch": asm_bytes,
                   "c_bar": grid.nconn / nb, "dsec_rows_mean": rows_mean, "roofline_other": other,
                   "cold_ms_per_step": cold_ms,
                   "learned": learned_res,
                   "cold_cell_updates_per_s": (nb / (cold_ms / 1e3)) if cold_ms else None,
                   "roofline_flash": flash_roof,
                   "roofline_step": step_roof,
                   "pcie_inclusive_ms_per_step": handoff_ms,
                   "pcie_inclusive_cell_updates_per_s": (nb / (handoff_ms / 1e3)) if handoff_ms else None},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    ctx.close()
    if not side_case and not a.no_fimn:
        # IsoT_FIMn (secondary variables S and n_ij) on the same field
        try:
            out["config"]["fimn"] = fimn_leg(nx, ny, nz, max(3, a.steps // 2), 1)
        except Exception as e:  # informative only
            log(f"[bench] IsoT_FIMn leg failed: {e}")
    if rank == 0 and a.serving_requests > 0 and not side_case and not a.serial:
        # the serving form beside the headline: 256 ensemble realisations of
        # 1e3..1e6 cells, one Newton iteration per request, admission by a
        # 200 GB device-memory budget (opencaeporo_amd/serving.py)
        try:
            from opencaeporo_amd import serving
            sv = serving.benchmark(n_real=256, cells_lo=1e3, cells_hi=1e6, n_req=a.serving_requests,
                                   budget_gb=200.0, log=log, load=0.8)
            keys = ("p50_ms", "p99_ms", "mean_ms", "throughput_req_s", "cell_updates_s", "admissions", "evictions")
            out["config"]["serving"] = {
                "realisations": 256, "cells": "log-uniform in [1e3, 1e6]",
                "total_cells": sv["config"]["total_cells"], "requests": a.serving_requests,
                "budget_gb": 200.0, "pool_cells": sv["config"]["pool_cells"],
                "offered_rate_per_s": sv["config"]["rate_per_s"], "offered_load_vs_one_at_a_time": 0.8,
                "capacity": sv["capacity"],
                "one_at_a_time": {k: sv["one_at_a_time"][k] for k in keys},
                "continuous": {k: sv["continuous"][k] for k in keys + ("mean_batch", "max_batch")},
                "pool_one_at_a_time": {k: sv["pool_one_at_a_time"][k] for k in keys},
                "gain": sv["gain"]}
        except Exception as e:  # informative only
            log(f"[bench] serving leg failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
