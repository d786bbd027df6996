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
* Larger ones run as their own ``Sim`` (the whole-GPU large path), each
  driven through all its intervals by a worker thread, so their launches
  overlap the pool's on the device and their pace does not depend on the
  pool's ticks.

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
import sys
import time

import numpy

from . import _lib
from .nbody import Sim

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
        self.sim = Sim(x, v, m, numpy.sum(m),
                       -1. if req.omega is None else req.omega ** 2.)

    def advance(self, h, k):
        self.sim.advance(h, k)

    def finish(self, n):
        x = numpy.empty(n)
        v = numpy.empty(n)
        self.sim.read(x, v)
        self.paths = self.sim.path_stats() + self.sim.warm_failures()
        self.sim.close()
        return x, v


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


class Server:
    """Continuous-batching scheduler.

    ``budget_bytes``: HBM the served state may use (the pool arena counts in
    full); ``pool_sheets``/``max_sims``: the pool's arena and slot count;
    ``split``: largest n served from the pool; ``max_big``: large-path
    simulations run at once; ``tick``: leapfrog steps per scheduling tick
    (one output interval) of the pool; ``tick_large``: steps per output
    interval of a large-path simulation (default ``tick``).
    ``pool_factory``/``big_factory`` exist so the scheduling logic can be
    exercised without a GPU.
    """

    def __init__(self, budget_bytes=200e9, pool_sheets=None, max_sims=2048,
                 split=1 << 16, max_big=8, tick=10, tick_large=None,
                 pool_factory=None,
                 big_factory=None, clock=time.perf_counter, sleep=time.sleep):
        if pool_sheets is None:
            pool_sheets = max_sims * split
        self.pool_bytes = pool_sheets * POOL_BYTES_PER_SHEET
        if self.pool_bytes > budget_bytes:
            raise ValueError('the pool arena alone exceeds the HBM budget')
        self.budget = float(budget_bytes)
        self.split = int(split)
        self.max_big = int(max_big)
        self.tick = int(tick)
        self.tick_large = int(tick if tick_large is None else tick_large)
        self.clock = clock
        self.sleep = sleep
        make_pool = pool_factory or (lambda: Pool(pool_sheets, max_sims, split))
        self.pool = make_pool()
        self.big_factory = big_factory or _Big
        self.stats = {}

    def close(self):
        if self.pool is not None:
            self.pool.close()
        self.pool = None

    def run(self, requests):
        """Serve every request; returns {id: Result}.  Latency counts from a
        request's scheduled arrival to the host having its final state."""
        t0 = self.clock()
        now = lambda: self.clock() - t0
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        qs = collections.deque()   # arrived, waiting for the pool
        qb = collections.deque()   # arrived, waiting to run as a Sim
        small = {}   # slot -> [request, steps left, result]
        big = {}     # future -> result
        big_bytes = 0.
        results = {}
        ticks = stalls = 0
        work = 0
        pool_busy = 0.
        workers = concurrent.futures.ThreadPoolExecutor(max(1, self.max_big))

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
                if (len(big) >= self.max_big
                        or self.pool_bytes + big_bytes + need > self.budget):
                    if not big:
                        raise MemoryError('request %r does not fit the HBM budget'
                                          % (r.id,))
                    stalls += 1
                    break
                res = result_for(r, 'sim')
                x, v, m = r.make()
                b = self.big_factory(r, x, v, m)
                big[workers.submit(_run_big, b, r, self.tick_large)] = res
                big_bytes += need
                qb.popleft()

        def retire_big(done):
            nonlocal big_bytes, work
            for f in done:
                res = big.pop(f)
                res.x, res.v, res.paths = f.result()
                res.completed = now()
                results[res.id] = res
                work += res.n * res.steps
                big_bytes -= res.n * SIM_BYTES_PER_SHEET

        try:
            while pending or qs or qb or small or big:
                arrive()
                retire_big([f for f in list(big) if f.done()])
                admit()
                if not small:
                    if not big:
                        if pending:
                            self.sleep(max(0., pending[0].arrival - now()))
                        continue
                    # only large-path work: wait for a completion or an arrival
                    wait = (max(0., pending[0].arrival - now()) if pending else None)
                    done, _ = concurrent.futures.wait(
                        list(big), timeout=wait,
                        return_when=concurrent.futures.FIRST_COMPLETED)
                    retire_big(done)
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
            workers.shutdown(wait=True)
        wall = now()
        self.stats = {'ticks': ticks, 'admission_stalls': stalls, 'wall_s': wall,
                      'sheet_steps': work, 'pool_device_s': pool_busy}
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
