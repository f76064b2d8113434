"""Serving form: the captured Newton iteration replays to the eager result,
and a realisation's result is the same served alone and in a mix of
concurrent requests with admissions and evictions."""
import numpy as np
import pytest

from opencaeporo_amd import fluids, lib, serving
from tests import gpu_util

pytestmark = pytest.mark.gpu


def models(n, lo, hi, seed):
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    return serving.make_ensemble(n, lo, hi, seed, eos, fluids.make_satfunc(), fluids.make_rock())


def system(r):
    IA, JA, val, rhs = r.ctx.get_bsr()
    return IA, JA, val, rhs


def test_graph_replay_matches_eager():
    gpu_util.say("[serving] graph replay vs eager")
    r = models(1, 3000, 3000, seed=3)[0]
    r.initialize()
    ctx = lib.Context(r.eos, r.sat, r.rock)
    ctx.set_cells(r.nb)
    ctx.upload(r.host)
    ctx.set_conn(r.grid)
    ctx.save_flash_inputs()
    # eager: the same calls one by one
    ctx.restore_flash_inputs()
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    ctx.calres(serving.DT, norms=False)
    ctx.assemble(serving.DT)
    ref = ctx.get_bsr()
    ctx.capture_iteration(serving.DT)
    for k in range(3):
        # eager work and synchronisations between the replays
        ctx.calkrpc_rock()
        ctx.synchronize()
        ctx.launch_iteration()
        while not ctx.iteration_done():
            pass
        got = ctx.get_bsr()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
    ctx.close()


def test_result_is_the_same_alone_and_in_the_mix():
    gpu_util.say("[serving] continuous batching vs alone")
    ens = models(6, 1000, 20000, seed=11)
    for r in ens:
        r.initialize()
    need = sum(r.bytes_needed() for r in ens)
    reqs = serving.poisson_requests(30, rate=2000.0, n_real=len(ens), seed=12)
    srv = serving.Server(ens, budget_bytes=0.6 * need)
    srv.run(reqs)
    assert all(q.end is not None for q in reqs)
    gpu_util.say(f"  admissions {srv.admissions} evictions {srv.evictions} max concurrent {srv.max_concurrent}")
    assert srv.evictions > 0 and srv.max_concurrent > 1
    for r in ens:
        if not r.resident:
            r.admit()
        r.ctx.launch_iteration()
        r.ctx.synchronize()
        mixed = system(r)
        r.evict()
        alone = models(6, 1000, 20000, seed=11)[r.rid]
        alone.initialize()
        alone.admit()
        alone.ctx.launch_iteration()
        alone.ctx.synchronize()
        ref = system(alone)
        alone.evict()
        for a, b in zip(mixed, ref):
            np.testing.assert_array_equal(a, b)


def test_reallocation_drops_the_captured_iteration():
    """A captured iteration holds the addresses it was captured with: after
    set_cells / set_conn reallocate, launching it is refused (ESTATE)."""
    gpu_util.say("[serving] captured graph dropped on reallocation")
    r = models(1, 2000, 2000, seed=4)[0]
    r.initialize()
    ctx = lib.Context(r.eos, r.sat, r.rock)
    for realloc in ("set_conn", "set_cells"):
        ctx.set_cells(r.nb)
        ctx.upload(r.host)
        ctx.set_conn(r.grid)
        ctx.save_flash_inputs()
        ctx.capture_iteration(serving.DT)
        if realloc == "set_conn":
            ctx.set_conn(r.grid)
        else:
            ctx.set_cells(r.nb)
        with pytest.raises(lib.OcpgError, match="code -4"):
            ctx.launch_iteration()
    ctx.close()


def _alone(r):
    """The realisation's iteration on a context of its own (the plain path)."""
    ctx = lib.Context(r.eos, r.sat, r.rock)
    ctx.set_cells(r.nb)
    ctx.upload(r.host)
    ctx.set_conn(r.grid)
    ctx.save_flash_inputs()
    ctx.restore_flash_inputs()
    ctx.flash_fim(stats=False)
    ctx.calkrpc_rock()
    ctx.calres(serving.DT, norms=False)
    ctx.assemble(serving.DT)
    out = ctx.get_bsr()
    ctx.close()
    return out


def test_pool_batches_give_each_realisation_its_own_result():
    """Continuous batching on one pooled context: realisations of uneven
    size placed at reserved ranges, admitted and evicted by the pool's
    capacity, batched into replays of one captured iteration over the active
    ones.  Each realisation's system equals, bit for bit, its iteration on a
    context of its own."""
    gpu_util.say("[serving] pooled batches vs alone")
    ens = models(8, 1000, 20000, seed=11)
    for r in ens:
        r.initialize()
    ref = {r.rid: _alone(r) for r in ens}
    total = sum(r.nb for r in ens)
    pool = serving.Pool(ens[0].eos, ens[0].sat, ens[0].rock, capacity=int(0.6 * total))
    pool.prime(ens)
    srv = serving.BatchServer(ens, pool)
    reqs = serving.poisson_requests(40, rate=3000.0, n_real=len(ens), seed=12)
    srv.run(reqs)
    assert all(q.end is not None and q.end >= q.start >= q.arrival for q in reqs)
    gpu_util.say(f"  batches {srv.batches} (largest {srv.max_batch_seen}), admissions {srv.admissions}, "
                 f"evictions {srv.evictions}")
    assert srv.max_batch_seen > 1 and srv.evictions > 0
    checked = set()
    for _ in range(len(ens)):
        for rid in list(srv.base):
            if rid in checked or rid not in srv.active:
                continue
            got = srv.system(rid)
            for a, b in zip(got, ref[rid]):
                np.testing.assert_array_equal(a, b)
            checked.add(rid)
        rest = [r.rid for r in ens if r.rid not in checked]
        if not rest:
            break
        srv.run([serving.Request(rid, 0.0) for rid in rest])
    assert checked == {r.rid for r in ens}
    srv.clear()
    pool.free_images(ens)
    pool.close()


def test_pool_image_moves_a_realisation():
    """A realisation placed from its cell state, exported, removed and
    imported at another base gives the same system; the state part of the
    image survives an export/import round trip."""
    gpu_util.say("[serving] pool image export / import")
    a, b = models(2, 3000, 6000, seed=21)
    for r in (a, b):
        r.initialize()
    pool = serving.Pool(a.eos, a.sat, a.rock, capacity=a.nb + b.nb + 100)
    pool.prime([a, b])
    ctx = pool.ctx
    ctx.pool_import(0, b.image)            # b first, so a lands at a new base
    ctx.pool_import(b.nb + 37, a.image)
    for base in (0, b.nb + 37):
        ctx.pool_activate(base, True)
    ctx.launch_iteration()
    ctx.synchronize()
    for base, r in ((0, b), (b.nb + 37, a)):
        got = ctx.pool_system(base, r.nb)
        for x, y in zip(got, _alone(r)):
            np.testing.assert_array_equal(x, y)
        back = r.host.copy()
        ctx.pool_download(base, back, saved=True)
        for f in serving.STATE_FIELDS:
            np.testing.assert_array_equal(getattr(back, f), getattr(r.host, f), err_msg=f)
    with pytest.raises(lib.OcpgError):
        ctx.pool_export(0, b.image, all=False)      # still switched on
    for base in (0, b.nb + 37):
        ctx.pool_activate(base, False)
    ctx.pool_export(0, b.image, all=False)
    ctx.pool_remove(0)
    ctx.pool_remove(b.nb + 37)
    ctx.pool_sync_copies()
    with pytest.raises(lib.OcpgError):
        ctx.reset_to_last_step()                    # a pool keeps no last step
    pool.free_images([a, b])
    pool.close()
