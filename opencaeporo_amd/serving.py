"""Serving form: continuous batching of Newton-iteration requests from many
ensemble realisations on one GPU.

Each realisation is a reservoir model of its own size (its own grid, fluid
state and wells) driven through the reference's FIM Newton loop; a request
asks for one Newton iteration of one realisation (flash + kr/Pc + rock +
CalRes + AssembleMat on its resident state).  The server

* keeps a realisation's state resident in HBM between its iterations (one
  C-ABI context per realisation, `ocpg_ctx`, with its own stream);
* admits a realisation by device-memory budget: when the budget is full, the
  least recently used idle realisation is evicted (its primary variables and
  last-step state go to host memory) and readmitted on its next request;
* runs each iteration as one replay of a HIP graph captured at admission
  (`ocpg_capture_iteration`), so a request costs one launch;
* keeps admitting while others run: requests that arrive while the GPU is
  busy start as soon as their realisation is free and memory allows, and
  their kernels overlap with the ones already on the device (each context's
  stream runs independently).

`run()` replays a Poisson arrival sequence in real time and reports
throughput and the p50/p99 of request latency (arrival to completion), for
continuous batching and for one-at-a-time service of the same arrivals.

    python -m opencaeporo_amd.serving --realisations 64 --rate 200
"""
import argparse
import collections
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
        r = self.real[rid]
        ctx = self.pool.ctx
        if rid in self.active:
            ctx.pool_activate(self.base[rid], False)
            self.active.discard(rid)
        ctx.pool_download(self.base[rid], r.host, saved=True)
        ctx.pool_remove(self.base[rid])
        self.pool.release(self.base.pop(rid), r.nb)
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
        self.pool.ctx.pool_place(base, r.host, r.grid)
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


def benchmark(n_real=64, cells_lo=1e3, cells_hi=2e5, n_req=256, rate=None, budget_gb=16.0, seed=7,
              log=lambda *a: None, load=0.8):
    """load: the offered arrival rate as a fraction of the one-at-a-time
    capacity measured on the same ensemble (used when rate is None)."""
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    sat = fluids.make_satfunc()
    rock = fluids.make_rock()
    ens = make_ensemble(n_real, cells_lo, cells_hi, seed, eos, sat, rock)
    t0 = time.time()
    for r in ens:
        r.initialize()
    total = sum(r.nb for r in ens)
    log(f"[serving] {n_real} realisations, {total} cells, initial states in {time.time() - t0:.1f}s")
    # service time of one request on an idle GPU, to set the arrival rate
    probe = max(ens, key=lambda r: r.nb)
    probe.admit()
    probe.ctx.launch_iteration()
    probe.ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        probe.ctx.launch_iteration()
    probe.ctx.synchronize()
    largest_ms = (time.perf_counter() - t1) / 3 * 1e3
    probe.evict()
    log(f"[serving] largest realisation {probe.nb} cells: {largest_ms:.2f} ms per iteration")
    if rate is None:
        # offered load: a fraction of what one-at-a-time service sustains on
        # this ensemble (measured with every request queued at once)
        for r in ens:
            r.busy = False
        burst = [serving_request for serving_request in poisson_requests(n_req, 1e9, n_real, seed + 5)]
        srv = Server(ens, budget_gb * 1e9, max_inflight=1)
        srv.run(burst)
        cap = summarize(burst, {r.rid: r for r in ens})["throughput_req_s"]
        rate = load * cap
        log(f"[serving] one-at-a-time capacity {cap:.1f} requests/s; offered load {rate:.1f}/s")
    out = {}
    for mode, mi in (("continuous", None), ("one_at_a_time", 1)):
        for r in ens:
            if r.resident:
                r.evict()
            r.busy = False
        reqs = poisson_requests(n_req, rate, n_real, seed + 5)
        srv = Server(ens, budget_gb * 1e9, max_inflight=mi)
        srv.run(reqs)
        s = summarize(reqs, {r.rid: r for r in ens})
        s.update(admissions=srv.admissions, evictions=srv.evictions, max_concurrent=srv.max_concurrent)
        out[mode] = s
        log(f"[serving] {mode}: " + ", ".join(f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}"
                                              for k, v in s.items()))
    for r in ens:
        if r.resident:
            r.evict()
    out["config"] = dict(realisations=n_real, cells_lo=cells_lo, cells_hi=cells_hi, total_cells=total,
                         requests=n_req, rate_per_s=rate, budget_gb=budget_gb, load=load)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--realisations", type=int, default=64)
    ap.add_argument("--cells-lo", type=float, default=1e3)
    ap.add_argument("--cells-hi", type=float, default=2e5)
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=None)
    ap.add_argument("--budget-gb", type=float, default=16.0)
    a = ap.parse_args()
    res = benchmark(a.realisations, a.cells_lo, a.cells_hi, a.requests, a.rate, a.budget_gb,
                    log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
