"""Serving form: continuous batching of Newton-iteration requests from many
ensemble realisations on one GPU.

Each realisation is a reservoir model of its own size (its own grid and
fluid state, no wells) driven through the reference's FIM Newton loop; a
request asks for one Newton iteration of one realisation (flash + kr/Pc +
rock + CalRes + AssembleMat on its resident state, from its saved iteration
inputs).

Continuous batching (`Pool`, `BatchServer`):

* one C-ABI context is a pool of cells (`ocpg_pool_init`); a realisation is
  placed at a free range of it (`ocpg_pool_place`) and stays resident
  between its iterations;
* admission is by the pool's capacity, i.e. the device-memory budget: when
  no free range fits, the least recently used idle realisation is evicted
  (its state goes back to host memory) and readmitted on its next request;
* a batch is every request that has arrived when the GPU finishes the
  previous one; its realisations' masks are switched on
  (`ocpg_pool_activate`) and the batch runs as ONE replay of an iteration
  graph captured once for the whole pool, so a batch of many small
  realisations costs one set of launches;
* while a batch runs, the realisations of waiting requests are placed into
  free ranges, overlapping their uploads with the GPU's work.

One at a time (`Server` with max_inflight=1): the plain GPU path, each
realisation on a context of its own with its own captured iteration,
admitted and evicted by the same budget, one request per launch.

`benchmark()` measures both modes' capacity (all requests queued at once)
and their p50/p99 latency (arrival to completion) for the same Poisson
arrivals.

    python -m opencaeporo_amd.serving --realisations 256 --cells-hi 1e6
"""
import argparse
import collections
import ctypes as C
import json
import sys
import time

import numpy as np

from . import abi, fluids, lib, workloads

DT = 5.0
# Fields that carry a realisation between admissions: the Newton iteration's
# inputs, the last time step's moles and the static geometry.
STATE_FIELDS = ["P", "T", "Ni", "lNi", "S", "phaseNum", "xij", "skipFlag", "skipMinEig", "skipP", "skipT",
                "skipZi", "v", "depth", "poroInit"]


class Realisation:
    """One ensemble member: an nx x ny x nz grid of about `cells` cells with
    its own seeded permeability, pressure and composition fields."""

    def __init__(self, rid, cells, seed, eos, sat, rock, nz=10):
        self.rid = rid
        self.seed = seed
        self.eos, self.sat, self.rock = eos, sat, rock
        nxy = max(1, int(round(cells / nz)))
        nx = max(1, int(round(np.sqrt(nxy))))
        ny = max(1, int(round(nxy / nx)))
        self.grid = workloads.cartesian_grid(nx, ny, nz, seed=seed)
        self.nb = self.grid.nb
        self.NC = eos.nc
        self.ctx = None
        self.host = None      # state while not resident
        self.image = None     # page-locked image for a pool (ocpg_pool_export)
        self.image_bytes = 0
        self.bytes = 0
        self.busy = False
        self.last_used = 0.0

    # ---- life cycle ----
    def initialize(self):
        """The realisation's starting state: InitFlashFIM of its feed, a cold
        flash, and a Newton-like pressure change; kept on the host."""
        nb, NC = self.nb, self.NC
        P, z, Sw = workloads.cell_inputs(nb, NC, seed=self.seed + 1)
        b = abi.Bulk(nb, NC, fields=["P", "T", "S", "v", "depth", "poroInit"])
        b.P[:] = P
        b.T[:] = workloads.T_F
        b.S.reshape(nb, 3)[:, 2] = Sw
        b.v[:] = self.grid.v
        b.depth[:] = self.grid.depth
        b.poroInit[:] = 0.3
        ctx = lib.Context(self.eos, self.sat, self.rock)
        ctx.set_cells(nb)
        ctx.upload(b)
        ctx.set_conn(self.grid)
        ctx.calrock()
        ctx.init_flash_fim(z)
        ctx.calkrpc()
        ctx.update_last_step()
        ctx.upload(abi.Bulk(nb, NC, fields=["phaseNum", "skipFlag"]))
        ctx.flash_fim(stats=False)
        pb = abi.Bulk(nb, NC, fields=["P"])
        ctx.download(pb)
        pb.P[:] += np.random.default_rng(self.seed + 2).uniform(-5.0, 5.0, nb)
        ctx.upload(pb)
        self.host = abi.Bulk(nb, NC, fields=STATE_FIELDS)
        ctx.download(self.host)
        ctx.close()

    def admit(self):
        """Make the realisation resident: upload its state, capture its
        Newton iteration."""
        ctx = lib.Context(self.eos, self.sat, self.rock)
        ctx.set_cells(self.nb)
        ctx.upload(self.host)
        ctx.set_conn(self.grid)
        ctx.save_flash_inputs()
        ctx.capture_iteration(DT)
        self.ctx = ctx
        self.bytes = ctx.device_bytes()

    def evict(self):
        """Back to host memory: the state the next iteration starts from (the
        saved iteration inputs: primary variables, phase state, skip memory)
        and the last time step, so readmission changes nothing."""
        self.ctx.restore_flash_inputs()
        self.ctx.download(self.host)
        self.ctx.close()
        self.ctx = None
        self.bytes = 0

    @property
    def resident(self):
        return self.ctx is not None

    def bytes_needed(self):
        # measured on admission; before that, the library's per-cell footprint
        return self.bytes or int(self.nb * 10_500)


class Request:
    __slots__ = ("rid", "arrival", "start", "end")

    def __init__(self, rid, arrival):
        self.rid, self.arrival, self.start, self.end = rid, arrival, None, None


class Server:
    """Continuous batching over resident realisations within a device-memory
    budget.  max_inflight = 1 gives one-at-a-time service."""

    def __init__(self, realisations, budget_bytes, max_inflight=None):
        self.real = {r.rid: r for r in realisations}
        self.budget = budget_bytes
        self.max_inflight = max_inflight or len(realisations)
        self.used = sum(r.bytes for r in realisations if r.resident)
        self.admissions = 0
        self.evictions = 0
        self.max_concurrent = 0

    def _make_room(self, need):
        idle = sorted((r for r in self.real.values() if r.resident and not r.busy), key=lambda r: r.last_used)
        for r in idle:
            if self.used + need <= self.budget:
                break
            self.used -= r.bytes
            r.evict()
            self.evictions += 1
        return self.used + need <= self.budget

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (sorted by arrival, seconds from the start) in real
        time.  Returns the requests with start/end times filled in."""
        pending = collections.deque()
        inflight = []
        t0 = clock()
        i = 0
        n = len(requests)
        done = 0
        while done < n:
            now = clock() - t0
            while i < n and requests[i].arrival <= now:
                pending.append(requests[i])
                i += 1
            # admit and launch in arrival order; a busy realisation's requests wait
            started = []
            for q in pending:
                if len(inflight) >= self.max_inflight:
                    break
                r = self.real[q.rid]
                if r.busy:
                    continue
                if not r.resident:
                    if not self._make_room(r.bytes_needed()):
                        break
                    r.admit()
                    self.used += r.bytes
                    self.admissions += 1
                r.ctx.launch_iteration()
                r.busy = True
                q.start = clock() - t0
                inflight.append(q)
                started.append(q)
            for q in started:
                pending.remove(q)
            self.max_concurrent = max(self.max_concurrent, len(inflight))
            # completions
            still = []
            for q in inflight:
                r = self.real[q.rid]
                if r.ctx.iteration_done():
                    q.end = clock() - t0
                    r.busy = False
                    r.last_used = q.end
                    done += 1
                else:
                    still.append(q)
            inflight = still
            if not started and len(inflight) == len(still) and (i >= n or requests[i].arrival > clock() - t0):
                time.sleep(0.0002)
        return requests


class Pool:
    """One device context holding many realisations at reserved cell ranges
    (ocpg_pool_*), with one captured Newton iteration over every active
    realisation.  Ranges are handed out first fit."""

    def __init__(self, eos, sat, rock, capacity, max_neighbours=6):
        self.ctx = lib.Context(eos, sat, rock)
        self.ctx.pool_init(capacity, max_neighbours)
        self.ctx.capture_iteration(DT)
        self.capacity = int(capacity)
        self.free = [[0, self.capacity]]      # sorted [base, length]
        self.bytes = self.ctx.device_bytes()

    def alloc(self, nb):
        for k, (b, n) in enumerate(self.free):
            if n >= nb:
                if n == nb:
                    self.free.pop(k)
                else:
                    self.free[k] = [b + nb, n - nb]
                return b
        return None

    def prime(self, realisations, log=lambda *a: None):
        """Builds each realisation's image (placed once through the cell
        layout, then exported), so that admissions are image imports; the
        pool is left empty."""
        ctx = self.ctx
        for k, r in enumerate(realisations):
            if r.image is not None:
                continue
            if r.nb > self.capacity:
                raise ValueError(f"realisation {r.rid} ({r.nb} cells) exceeds the pool")
            ctx.pool_place(0, r.host, r.grid)
            r.image_bytes = ctx.pool_image_bytes(0)
            p = ctx.lib.ocpg_host_alloc(r.image_bytes)
            if not p:
                raise lib.OcpgError(f"ocpg_host_alloc({r.image_bytes}) failed")
            r.image = p
            ctx.pool_export(0, p, all=True)
            ctx.pool_sync_copies()
            ctx.pool_remove(0)
        ctx.synchronize()

    def free_images(self, realisations):
        for r in realisations:
            if r.image is not None:
                self.ctx.lib.ocpg_host_free(C.c_void_p(r.image))
                r.image = None

    def release(self, base, nb):
        self.free.append([base, nb])
        self.free.sort()
        merged = []
        for b, n in self.free:
            if merged and merged[-1][0] + merged[-1][1] == b:
                merged[-1][1] += n
            else:
                merged.append([b, n])
        self.free = merged

    def close(self):
        self.ctx.close()


class BatchServer:
    """Continuous batching on a Pool: every request that has arrived when
    the GPU finishes a batch joins the next one (one request per realisation
    per batch; a second request for the same realisation waits for the
    next), and the whole batch is one replay of the captured iteration over
    the active realisations.  While a batch runs, the realisations of the
    waiting requests are admitted (placed into free ranges, evicting the
    least recently used idle ones when the pool is full), so their uploads
    overlap the GPU's work.  max_batch = 1 serves one request per launch on
    the same pool."""

    def __init__(self, realisations, pool, max_batch=None):
        """pool: an empty Pool on which every realisation has an image
        (Pool.prime)."""
        self.real = {r.rid: r for r in realisations}
        self.pool = pool
        self.max_batch = max_batch or len(realisations)
        self.base = {}               # rid -> cell base while resident
        self.active = set()          # rids whose mask is on
        self.inflight = set()        # rids in the running batch
        self.last_used = {}
        self.admissions = 0
        self.evictions = 0
        self.batches = 0
        self.max_batch_seen = 0

    def _evict(self, rid):
        """Back to host memory: the state part of its image is refreshed (on
        the pool's copy stream) and its range freed."""
        r = self.real[rid]
        ctx = self.pool.ctx
        base = self.base.pop(rid)
        if rid in self.active:
            ctx.pool_activate(base, False)
            self.active.discard(rid)
        ctx.pool_export(base, r.image, all=False)
        ctx.pool_remove(base)
        self.pool.release(base, r.nb)
        self.evictions += 1

    def _admit(self, rid, keep, evict=True):
        """Place realisation rid; evicts idle residents outside `keep` (least
        recently used first) until a range fits.  False if none does.  An
        eviction waits for the batch in flight, so the overlapped admissions
        take free ranges only (evict=False)."""
        r = self.real[rid]
        if r.nb > self.pool.capacity:
            raise ValueError(f"realisation {rid} ({r.nb} cells) exceeds the pool")
        base = self.pool.alloc(r.nb)
        while base is None:
            if not evict:
                return False
            idle = [q for q in self.base if q not in keep and q not in self.inflight]
            if not idle:
                return False
            self._evict(min(idle, key=lambda q: self.last_used.get(q, 0.0)))
            base = self.pool.alloc(r.nb)
        self.pool.ctx.pool_import(base, r.image)
        self.base[rid] = base
        self.admissions += 1
        return True

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (sorted by arrival, seconds from the start) in real
        time; fills in their start/end times."""
        ctx = self.pool.ctx
        pending = collections.deque()
        batch = []
        t0 = clock()
        i, n, done = 0, len(requests), 0
        while done < n:
            now = clock() - t0
            while i < n and requests[i].arrival <= now:
                pending.append(requests[i])
                i += 1
            if batch and ctx.iteration_done():
                end = clock() - t0
                for q in batch:
                    q.end = end
                    self.last_used[q.rid] = end
                done += len(batch)
                batch = []
                self.inflight = set()
            if not batch and pending:
                chosen, seen = [], set()
                for q in pending:
                    if len(chosen) >= self.max_batch:
                        break
                    if q.rid in seen:
                        continue
                    if q.rid not in self.base and not self._admit(q.rid, keep=seen):
                        continue
                    chosen.append(q)
                    seen.add(q.rid)
                if chosen:
                    for q in chosen:
                        pending.remove(q)
                    for rid in self.active - seen:
                        ctx.pool_activate(self.base[rid], False)
                    for rid in seen - self.active:
                        ctx.pool_activate(self.base[rid], True)
                    self.active = set(seen)
                    ctx.launch_iteration()
                    start = clock() - t0
                    for q in chosen:
                        q.start = start
                    batch = chosen
                    self.inflight = set(seen)
                    self.batches += 1
                    self.max_batch_seen = max(self.max_batch_seen, len(chosen))
                    continue
            if batch:
                # overlap: admit one waiting realisation while the GPU works
                for q in pending:
                    if q.rid not in self.base:
                        if not self._admit(q.rid, keep=self.inflight, evict=False):
                            time.sleep(0.0001)
                        break
                else:
                    time.sleep(0.0001)
            elif i < n:
                time.sleep(min(0.0002, max(0.0, requests[i].arrival - (clock() - t0))))
        return requests

    def system(self, rid):
        """The realisation's system from its last iteration (own indexing)."""
        return self.pool.ctx.pool_system(self.base[rid], self.real[rid].nb)

    def clear(self):
        """Evicts every resident realisation (the pool is empty again)."""
        for rid in list(self.base):
            self._evict(rid)
        self.pool.ctx.pool_sync_copies()


def make_ensemble(n, cells_lo, cells_hi, seed, eos, sat, rock):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(cells_lo), np.log(cells_hi), n)).astype(int)
    return [Realisation(k, int(sizes[k]), seed + 1000 * (k + 1), eos, sat, rock) for k in range(n)]


def poisson_requests(n_req, rate, n_real, seed):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    rids = rng.integers(0, n_real, n_req)
    return [Request(int(r), float(a)) for r, a in zip(rids, t)]


def summarize(reqs, ens):
    lat = np.array([(q.end - q.arrival) * 1e3 for q in reqs])
    span = max(q.end for q in reqs) - min(q.arrival for q in reqs)
    cells = sum(ens[q.rid].nb for q in reqs)
    return dict(requests=len(reqs), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), throughput_req_s=len(reqs) / span, cell_updates_s=cells / span,
                makespan_s=span)


def _reset(ens):
    for r in ens:
        if r.resident:
            r.evict()
        r.busy = False


def pool_bytes_per_cell(eos, sat, rock, probe=1 << 16):
    """Device bytes per cell of a pool (measured on a small one)."""
    p = Pool(eos, sat, rock, probe)
    b = p.bytes / probe
    p.close()
    return b


def benchmark(n_real=256, cells_lo=1e3, cells_hi=1e6, n_req=1024, rate=None, budget_gb=96.0, seed=7,
              log=lambda *a: None, load=0.8):
    """The serving form on an ensemble of n_real realisations with cell counts
    log-uniform in [cells_lo, cells_hi], within a device-memory budget:

    * one_at_a_time: the plain GPU path, one request at a time (each
      realisation on a context of its own with its captured iteration,
      admitted and evicted by the budget);
    * continuous: continuous batching on one pool of the same budget
      (BatchServer), admissions by image import;
    * pool_one_at_a_time: the same pool and admissions, one request per
      launch (what batching alone gains).

    Each mode's capacity (every request queued at once) is measured, and
    each serves the same Poisson arrivals at `load` x the one-at-a-time
    capacity (or at `rate` requests/s)."""
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    ens = make_ensemble(n_real, cells_lo, cells_hi, seed, eos, sat, rock)
    t0 = time.time()
    for k, r in enumerate(ens):
        r.initialize()
        if (k + 1) % 32 == 0:
            log(f"[serving] initialised {k + 1}/{n_real} realisations ({time.time() - t0:.1f}s)")
    total = sum(r.nb for r in ens)
    by_id = {r.rid: r for r in ens}
    budget = budget_gb * 1e9
    per_cell = pool_bytes_per_cell(eos, sat, rock)
    capacity = int(min(budget / per_cell, (1 << 31) - 1))
    capacity = max(capacity, max(r.nb for r in ens))
    log(f"[serving] {n_real} realisations, {total} cells (largest {max(r.nb for r in ens)}), initial states in "
        f"{time.time() - t0:.1f}s; budget {budget_gb:g} GB = pool of {capacity} cells ({per_cell:.0f} B/cell)")
    out = {}

    def one_at_a_time(reqs):
        _reset(ens)
        srv = Server(ens, budget, max_inflight=1)
        srv.run(reqs)
        s = summarize(reqs, by_id)
        s.update(admissions=srv.admissions, evictions=srv.evictions)
        _reset(ens)
        return s

    pool = None

    def batched(reqs, max_batch=None):
        srv = BatchServer(ens, pool, max_batch=max_batch)
        srv.run(reqs)
        s = summarize(reqs, by_id)
        s.update(admissions=srv.admissions, evictions=srv.evictions, batches=srv.batches,
                 mean_batch=len(reqs) / max(srv.batches, 1), max_batch=srv.max_batch_seen)
        srv.clear()
        return s

    burst = lambda: poisson_requests(n_req, 1e9, n_real, seed + 5)
    # the per-realisation contexts first: they and the pool never share the GPU
    cap1 = one_at_a_time(burst())
    if rate is None:
        rate = load * cap1["throughput_req_s"]
    arrivals = lambda: poisson_requests(n_req, rate, n_real, seed + 5)
    out["one_at_a_time"] = one_at_a_time(arrivals())
    t1 = time.time()
    pool = Pool(eos, sat, rock, capacity)
    pool.prime(ens)
    log(f"[serving] pool of {capacity} cells ({pool.bytes / 1e9:.1f} GB) and {n_real} images "
        f"({sum(r.image_bytes for r in ens) / 1e9:.1f} GB page-locked) in {time.time() - t1:.1f}s")
    capB = batched(burst())
    capP = batched(burst(), max_batch=1)
    log(f"[serving] capacity (all {n_req} requests queued at once): one at a time {cap1['throughput_req_s']:.1f} "
        f"req/s, pool one at a time {capP['throughput_req_s']:.1f} req/s, continuous batching "
        f"{capB['throughput_req_s']:.1f} req/s")
    out["capacity"] = {"one_at_a_time_req_s": cap1["throughput_req_s"],
                       "pool_one_at_a_time_req_s": capP["throughput_req_s"],
                       "continuous_req_s": capB["throughput_req_s"],
                       "one_at_a_time_cell_updates_s": cap1["cell_updates_s"],
                       "continuous_cell_updates_s": capB["cell_updates_s"],
                       "continuous_mean_batch": capB["mean_batch"]}
    out["continuous"] = batched(arrivals())
    out["pool_one_at_a_time"] = batched(arrivals(), max_batch=1)
    for mode in ("one_at_a_time", "continuous", "pool_one_at_a_time"):
        log(f"[serving] {mode} at {rate:.1f} req/s: " + ", ".join(
            f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}" for k, v in out[mode].items()))
    out["gain"] = {"p99": out["one_at_a_time"]["p99_ms"] / out["continuous"]["p99_ms"],
                   "p50": out["one_at_a_time"]["p50_ms"] / out["continuous"]["p50_ms"],
                   "throughput_at_capacity": capB["throughput_req_s"] / cap1["throughput_req_s"],
                   "batching_alone_throughput_at_capacity": capB["throughput_req_s"] / capP["throughput_req_s"]}
    pool.free_images(ens)
    pool.close()
    out["config"] = dict(realisations=n_real, cells_lo=cells_lo, cells_hi=cells_hi, total_cells=total,
                         requests=n_req, rate_per_s=rate, budget_gb=budget_gb, pool_cells=capacity, load=load)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--realisations", type=int, default=256)
    ap.add_argument("--cells-lo", type=float, default=1e3)
    ap.add_argument("--cells-hi", type=float, default=1e6)
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--rate", type=float, default=None)
    ap.add_argument("--budget-gb", type=float, default=96.0)
    ap.add_argument("--load", type=float, default=0.8)
    a = ap.parse_args()
    res = benchmark(a.realisations, a.cells_lo, a.cells_hi, a.requests, a.rate, a.budget_gb,
                    log=lambda *x: print(*x, file=sys.stderr, flush=True), load=a.load)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
