"""Serving: simulations of varying N and duration arrive over time and are
advanced by continuous batching on one GPU.

No reference counterpart (the reference runs one simulation per ``nbody``
generator).  Each request is the approx=True leapfrog of
``wendy_amd.nbody(x, v, m, dt=tick*h, nleap=tick, omega=...)`` run for
``steps`` leapfrog steps: every output interval advances a simulation by
min(tick, steps left) steps of size h (``tick`` for pool-resident requests,
``tick_large`` for large-path ones).

* Simulations with n <= ``split`` live in a ``Pool`` (wdy_pool_*): one device
  arena, one launch per tick for all of them, one workgroup per simulation.
* Larger ones run as their own ``Sim`` (the whole-GPU large path).  By
  default (``large='gang'``) every such simulation after its first output
  interval advances in one ``Gang``: one kernel launch per leapfrog step for
  all of them, on the gang's stream, driven by one host thread, so that many
  medium simulations fill the GPU the way one large one does.  The first
  interval (a full radix sort, which partitions the state for the warm
  kernel) runs on the simulation's own stream in a preparation worker, and
  finished simulations are read back and freed by a retire worker, so the
  gang thread only ever waits for the gang's own work.  ``large='threads'``
  drives each simulation through its intervals in its own worker thread
  instead.  Either way the large path's pace does not depend on the pool's
  ticks.

Admission is first come, first served within each kind (pool-resident and
large-path requests queue separately, so neither blocks the other): a request
waits while no slot, arena range or HBM budget is free, and joins the running
batch at the next tick once one is (requests that arrive while a tick runs
are uploaded during it).
A simulation's result does not depend on what else is being served.
"""
import collections
import concurrent.futures
import ctypes
import queue
import sys
import threading
import time

import numpy

from . import _lib
from .nbody import Gang, Sim

# HBM bytes per sheet: the pool arena (two buffer sets of key, velocity,
# index, mass) and an upper estimate for a large-path Sim (three state sets,
# sort scratch and warm-start metadata).
POOL_BYTES_PER_SHEET = 56
SIM_BYTES_PER_SHEET = 160


class Pool:
    """A ``wdy_pool`` handle: device arena + slot table for small
    simulations."""

    def __init__(self, capacity_sheets, max_sims=4096, max_n=1 << 16):
        self._h = None
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().wdy_pool_create(int(capacity_sheets), int(max_sims),
                                              int(max_n), ctypes.byref(h)))
        self._h = h
        self.capacity = int(capacity_sheets)
        self.max_sims = int(max_sims)
        self.max_n = int(max_n)
        self._keep = []   # host arrays of uploads still in flight
        self._n = {}

    def add(self, x, v, m, omega=None, twopiG=1.):
        """Upload one simulation (``m`` before the twopiG factor); returns its
        slot, or None when no slot or arena range is free."""
        x = _lib.f64(x)
        v = _lib.f64(v)
        m = _lib.f64(twopiG * numpy.asarray(m, dtype=numpy.float64))
        if not (len(x) == len(v) == len(m)):
            raise ValueError('x, v and m must have the same length')
        slot = ctypes.c_int(-1)
        rc = _lib.lib().wdy_pool_add(self._h, len(x), x, v, m, float(numpy.sum(m)),
                                     -1. if omega is None else float(omega) ** 2.,
                                     ctypes.byref(slot))
        if rc == _lib.WDY_ENOMEM:
            return None
        _lib.check(rc)
        self._keep.append((x, v, m))
        self._n[slot.value] = len(x)
        return slot.value

    def step(self, slots, h, nleap):
        """Enqueue one interval: slot slots[k] takes nleap[k] steps of h[k]."""
        slots = numpy.ascontiguousarray(slots, dtype=numpy.int32)
        h = numpy.ascontiguousarray(numpy.broadcast_to(h, slots.shape), dtype=numpy.float64)
        nleap = numpy.ascontiguousarray(numpy.broadcast_to(nleap, slots.shape),
                                        dtype=numpy.int32)
        _lib.check(_lib.lib().wdy_pool_step(self._h, len(slots), slots, h, nleap))

    def read(self, slot):
        """(x, v) of one slot in its original sheet order; blocks."""
        n = self._n[slot]
        x = numpy.empty(n)
        v = numpy.empty(n)
        _lib.check(_lib.lib().wdy_pool_read(self._h, int(slot),
                                            x.ctypes.data_as(ctypes.c_void_p),
                                            v.ctypes.data_as(ctypes.c_void_p)))
        self._keep.clear()
        return x, v

    def remove(self, slot):
        _lib.check(_lib.lib().wdy_pool_remove(self._h, int(slot)))
        del self._n[slot]

    def sync(self):
        """Wait for uploads and steps; returns the last step's device
        seconds."""
        sec = ctypes.c_double(0.)
        _lib.check(_lib.lib().wdy_pool_sync(self._h, ctypes.byref(sec)))
        self._keep.clear()
        return sec.value

    def stats(self):
        """(live slots, sheets in use, largest free arena range)."""
        live = ctypes.c_int(0)
        used = ctypes.c_int64(0)
        big = ctypes.c_int64(0)
        _lib.check(_lib.lib().wdy_pool_stats(self._h, ctypes.byref(live),
                                             ctypes.byref(used), ctypes.byref(big)))
        return live.value, used.value, big.value

    def close(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value and not _finalizing():
            _lib.lib().wdy_pool_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request:
    """One simulation to serve.  ``make`` returns (x, v, m) with m before
    twopiG (called once, at admission); ``n`` its size; ``arrival`` seconds
    after the start of ``Server.run``."""

    def __init__(self, rid, n, make, steps, h, omega=None, arrival=0., twopiG=1.):
        if steps < 1 or n < 1:
            raise ValueError('a request needs n >= 1 and steps >= 1')
        self.id = rid
        self.n = int(n)
        self.make = make
        self.steps = int(steps)
        self.h = float(h)
        self.omega = omega
        self.arrival = float(arrival)
        self.twopiG = twopiG
        self.ic = None   # (x, v, m) once made


class Result:
    """``where``: 'pool' or 'sim'; ``paths`` (large-path requests): the
    Sim's (warm steps, radix steps, redone intervals) and its warm-failure
    counts (capacity, overflow, stall)."""
    __slots__ = ('id', 'x', 'v', 'arrival', 'admitted', 'completed', 'n', 'steps',
                 'where', 'paths')

    @property
    def latency(self):
        return self.completed - self.arrival


class _Big:
    """A large-path simulation served through its own Sim."""

    def __init__(self, req, x, v, m):
        m = req.twopiG * numpy.asarray(m, dtype=numpy.float64)
        # the whole-GPU path whatever n (a Sim would pick its one-workgroup
        # path for the smallest n, a long kernel on a queue the gang may
        # share), with the tile shape for a shared GPU
        self.sim = Sim(x, v, m, numpy.sum(m),
                       -1. if req.omega is None else req.omega ** 2.,
                       path='large', shared=True)

    def advance(self, h, k):
        self.sim.advance(h, k)

    def settle(self):
        """Wait for (and check) the interval in flight."""
        self.sim.path_stats()

    def finish(self, n):
        x = numpy.empty(n)
        v = numpy.empty(n)
        self.sim.read(x, v)
        self.paths = self.sim.path_stats() + self.sim.warm_failures()
        self.sim.close()
        return x, v


class _GangAdvance:
    """The default gang: wendy_amd.Gang over the members' Sims."""

    MAX = Gang.MAX

    def __init__(self):
        self.gang = Gang()

    def advance(self, bigs, hs, ks):
        self.gang.advance([b.sim for b in bigs], hs, ks)

    def close(self):
        self.gang.close()


def _run_big(big, req, tick):
    """All output intervals of one large-path simulation (a worker thread;
    the library releases the GIL while the device works); returns (x, v,
    paths)."""
    left = req.steps
    while left:
        k = min(tick, left)
        big.advance(req.h, k)
        left -= k
    x, v = big.finish(req.n)
    return x, v, getattr(big, 'paths', None)


class _LargeThreads:
    """Large-path requests, one worker thread each."""

    def __init__(self, big_factory, tick, workers):
        self.big_factory = big_factory
        self.tick = tick
        self.pool = concurrent.futures.ThreadPoolExecutor(max(1, workers))
        self.live = {}   # future -> result

    def submit(self, r, res):
        x, v, m = r.make()
        b = self.big_factory(r, x, v, m)
        self.live[self.pool.submit(_run_big, b, r, self.tick)] = res

    def count(self):
        return len(self.live)

    def done(self, timeout=0.):
        """Completed (result, (x, v, paths)) pairs; waits up to ``timeout``
        seconds (None: until one completes) when none is ready."""
        fs = [f for f in self.live if f.done()]
        if not fs and self.live and (timeout is None or timeout > 0):
            fs, _ = concurrent.futures.wait(
                list(self.live), timeout=timeout,
                return_when=concurrent.futures.FIRST_COMPLETED)
        return [(self.live.pop(f), f.result()) for f in fs]

    def close(self):
        self.pool.shutdown(wait=True)


class _Member:
    __slots__ = ('big', 'req', 'res', 'left')

    def __init__(self, big, req, res, left):
        self.big, self.req, self.res, self.left = big, req, res, left


class _LargeGang:
    """Large-path requests in gangs.  A preparation worker makes each
    simulation and runs its first interval on its own stream; each of
    ``ngangs`` threads advances the prepared simulations given to it
    together (its own gang_factory() on its own stream, at most MAX members
    per call, in arrival order), so that while one gang's step is bound by
    a bucket's latency another's runs beside it; a retire worker reads back
    and frees finished simulations.  A prepared simulation joins the gang
    with the fewest sheets in flight.  Freeing a simulation's device memory
    waits for the whole device (hipFree synchronises it), ~30 ms beside busy
    gangs, so several retire workers free in parallel; with one, finished
    simulations queued up and held back admission."""

    def __init__(self, big_factory, gang_factory, tick, prep_workers=4,
                 ngangs=2, retire_workers=8):
        self.big_factory = big_factory
        self.tick = tick
        self.gangs = [gang_factory() for _ in range(max(1, ngangs))]
        self.max_members = getattr(self.gangs[0], 'MAX', 64)
        self.prep = concurrent.futures.ThreadPoolExecutor(max(1, prep_workers))
        self.retire = concurrent.futures.ThreadPoolExecutor(max(1, retire_workers))
        self.ready = [queue.Queue() for _ in self.gangs]  # None: stop
        self.load = [0] * len(self.gangs)   # sheets in flight per gang
        self.lock = threading.Lock()
        self.out = queue.Queue()        # (result, (x, v, paths) or exception)
        self.n_live = 0                 # submitted, not yet returned by done()
        self.gang_steps = 0             # gang calls (intervals) made
        self.threads = [threading.Thread(target=self._loop, args=(i,), daemon=True)
                        for i in range(len(self.gangs))]
        for t in self.threads:
            t.start()

    def submit(self, r, res):
        self.n_live += 1
        self.prep.submit(self._prepare, r, res)

    def count(self):
        return self.n_live

    def _finish(self, m, gi=None):
        try:
            x, v = m.big.finish(m.req.n)
            self.out.put((m.res, (x, v, getattr(m.big, 'paths', None))))
        except Exception as e:  # reported to the caller of done()
            self.out.put((m.res, e))

    def _prepare(self, r, res):
        try:
            x, v, m = r.make()
            b = self.big_factory(r, x, v, m)
            k = min(self.tick, r.steps)
            b.advance(r.h, k)
            settle = getattr(b, 'settle', None)
            if settle is not None:
                settle()
            mem = _Member(b, r, res, r.steps - k)
        except Exception as e:
            self.out.put((res, e))
            return
        if mem.left == 0:
            self._finish(mem)
            return
        with self.lock:
            gi = min(range(len(self.gangs)), key=lambda i: self.load[i])
            self.load[gi] += r.n
        self.ready[gi].put(mem)

    def _loop(self, gi):
        gang, ready = self.gangs[gi], self.ready[gi]
        active = []
        stop = False
        while True:
            if not active:
                if stop:
                    return
                item = ready.get()
                if item is None:
                    return
                active.append(item)
            while True:
                try:
                    item = ready.get_nowait()
                except queue.Empty:
                    break
                if item is None:
                    stop = True
                else:
                    active.append(item)
            members = active[:self.max_members]
            ks = [min(self.tick, m.left) for m in members]
            try:
                gang.advance([m.big for m in members],
                             [m.req.h for m in members], ks)
            except Exception as e:
                for m in active:
                    self.out.put((m.res, e))
                active = []
                continue
            with self.lock:
                self.gang_steps += 1
            for m, k in zip(members, ks):
                m.left -= k
            for m in members:
                if m.left == 0:
                    with self.lock:
                        self.load[gi] -= m.req.n
                    self.retire.submit(self._finish, m)
            active = [m for m in active if m.left > 0]

    def done(self, timeout=0.):
        got = []
        try:
            if timeout is None or timeout > 0:
                if self.n_live:
                    got.append(self.out.get(timeout=timeout))
            while True:
                got.append(self.out.get_nowait())
        except queue.Empty:
            pass
        out = []
        for res, val in got:
            self.n_live -= 1
            if isinstance(val, Exception):
                raise val
            out.append((res, val))
        return out

    def close(self):
        for q in self.ready:
            q.put(None)
        for t in self.threads:
            t.join()
        self.prep.shutdown(wait=True)
        self.retire.shutdown(wait=True)
        for g in self.gangs:
            close = getattr(g, 'close', None)
            if close is not None:
                close()


class Server:
    """Continuous-batching scheduler.

    ``budget_bytes``: HBM the served state may use (the pool arena counts in
    full); ``pool_sheets``/``max_sims``: the pool's arena and slot count;
    ``split``: largest n served from the pool; ``large``: 'gang' (the large
    path's simulations advance together, one launch per step) or 'threads'
    (one worker thread each); ``ngangs``: gangs (streams and driving
    threads) the large path spreads its simulations over;
    ``prep_workers``/``retire_workers``: threads that make and first-advance
    large-path simulations, and that read back and free finished ones;
    ``max_big``:
    large-path simulations in flight at once (default 64 per gang, 8 with
    threads); ``tick``: leapfrog
    steps per scheduling tick (one output interval) of the pool;
    ``tick_large``: steps per output interval of a large-path simulation
    (default ``tick``).
    ``pool_factory``/``big_factory``/``gang_factory`` exist so the
    scheduling logic can be exercised without a GPU.
    """

    def __init__(self, budget_bytes=200e9, pool_sheets=None, max_sims=2048,
                 split=1 << 16, max_big=None, tick=10, tick_large=None,
                 large='gang', prep_workers=4, ngangs=2, retire_workers=8,
                 pool_factory=None,
                 big_factory=None, gang_factory=None,
                 clock=time.perf_counter, sleep=time.sleep):
        if large not in ('gang', 'threads'):
            raise ValueError("large must be 'gang' or 'threads'")
        if pool_sheets is None:
            pool_sheets = max_sims * split
        self.pool_bytes = pool_sheets * POOL_BYTES_PER_SHEET
        if self.pool_bytes > budget_bytes:
            raise ValueError('the pool arena alone exceeds the HBM budget')
        self.budget = float(budget_bytes)
        self.split = int(split)
        self.large = large
        self.max_big = int(max_big if max_big is not None
                           else (64 * max(1, int(ngangs)) if large == 'gang'
                                 else 8))
        self.prep_workers = int(prep_workers)
        self.ngangs = int(ngangs)
        self.retire_workers = int(retire_workers)
        self.tick = int(tick)
        self.tick_large = int(tick if tick_large is None else tick_large)
        self.clock = clock
        self.sleep = sleep
        make_pool = pool_factory or (lambda: Pool(pool_sheets, max_sims, split))
        self.pool = make_pool()
        self.big_factory = big_factory or _Big
        self.gang_factory = gang_factory or _GangAdvance
        self.stats = {}

    def close(self):
        if self.pool is not None:
            self.pool.close()
        self.pool = None

    def run(self, requests, progress=None, progress_every=10.):
        """Serve every request; returns {id: Result}.  Latency counts from a
        request's scheduled arrival to the host having its final state.
        ``progress`` (optional) is called about every ``progress_every``
        seconds with a dict of the counts so far."""
        t0 = self.clock()
        now = lambda: self.clock() - t0
        last_report = [0.]
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        qs = collections.deque()   # arrived, waiting for the pool
        qb = collections.deque()   # arrived, waiting to run as a Sim
        small = {}   # slot -> [request, steps left, result]
        big_bytes = 0.
        results = {}
        ticks = stalls = 0
        work = 0
        pool_busy = 0.
        if self.large == 'gang':
            large = _LargeGang(self.big_factory, self.gang_factory,
                               self.tick_large, self.prep_workers, self.ngangs,
                               self.retire_workers)
        else:
            large = _LargeThreads(self.big_factory, self.tick_large, self.max_big)

        def arrive():
            t = now()
            while pending and pending[0].arrival <= t:
                r = pending.popleft()
                (qs if r.n <= self.split else qb).append(r)

        def result_for(r, where):
            res = Result()
            res.id, res.n, res.steps, res.arrival = r.id, r.n, r.steps, r.arrival
            res.where = where
            res.paths = None
            res.admitted = now()
            return res

        def admit():
            nonlocal big_bytes, stalls
            while qs:
                r = qs[0]
                if r.ic is None:
                    r.ic = r.make()
                slot = self.pool.add(*r.ic, omega=r.omega, twopiG=r.twopiG)
                if slot is None:
                    if not small:
                        raise MemoryError('request %r does not fit an empty pool'
                                          % (r.id,))
                    stalls += 1
                    break
                small[slot] = [r, r.steps, result_for(r, 'pool')]
                r.ic = None
                qs.popleft()
            while qb:
                r = qb[0]
                need = r.n * SIM_BYTES_PER_SHEET
                if (large.count() >= self.max_big
                        or self.pool_bytes + big_bytes + need > self.budget):
                    if not large.count():
                        raise MemoryError('request %r does not fit the HBM budget'
                                          % (r.id,))
                    stalls += 1
                    break
                large.submit(r, result_for(r, 'sim'))
                big_bytes += need
                qb.popleft()

        def retire_big(done):
            nonlocal big_bytes, work
            for res, (x, v, paths) in done:
                res.x, res.v, res.paths = x, v, paths
                res.completed = now()
                results[res.id] = res
                work += res.n * res.steps
                big_bytes -= res.n * SIM_BYTES_PER_SHEET

        def report():
            t = now()
            if progress is not None and t - last_report[0] >= progress_every:
                last_report[0] = t
                progress({'t_s': t, 'completed': len(results),
                          'waiting': len(pending) + len(qs) + len(qb),
                          'pool': len(small), 'large': large.count(),
                          'sheet_steps': work})

        try:
            while pending or qs or qb or small or large.count():
                report()
                arrive()
                retire_big(large.done())
                admit()
                if not small:
                    if not large.count():
                        if pending:
                            self.sleep(max(0., pending[0].arrival - now()))
                        continue
                    # only large-path work: wait for a completion or an arrival
                    wait = (max(0., pending[0].arrival - now()) if pending else None)
                    retire_big(large.done(timeout=wait))
                    continue
                ticks += 1
                slots = list(small)
                ks = [min(self.tick, small[s][1]) for s in slots]
                self.pool.step(slots, [small[s][0].h for s in slots], ks)
                # arrivals during the tick are uploaded now and join the next
                arrive()
                admit()
                pool_busy += self.pool.sync()
                for s, k in zip(slots, ks):
                    e = small[s]
                    e[1] -= k
                    if e[1] == 0:
                        res = e[2]
                        res.x, res.v = self.pool.read(s)
                        self.pool.remove(s)
                        res.completed = now()
                        results[res.id] = res
                        work += res.n * res.steps
                        del small[s]
        finally:
            large.close()
        wall = now()
        self.stats = {'ticks': ticks, 'admission_stalls': stalls, 'wall_s': wall,
                      'sheet_steps': work, 'pool_device_s': pool_busy}
        if self.large == 'gang':
            self.stats['gang_intervals'] = large.gang_steps
        return results


def latency_summary(results):
    """p50/p99 latency (seconds) and aggregate sheet-steps/s of a run."""
    lat = numpy.array([r.latency for r in results.values()])
    first = min(r.arrival for r in results.values())
    last = max(r.completed for r in results.values())
    work = sum(r.n * r.steps for r in results.values())
    return {'p50_s': float(numpy.percentile(lat, 50)),
            'p99_s': float(numpy.percentile(lat, 99)),
            'mean_s': float(lat.mean()), 'completed': len(lat),
            'sheet_steps_per_s': work / (last - first) if last > first else 0.}
