"""The ``nbody`` generator and the diagnostics of wendy, on the GPU.

Drop-in for the reference's Python API (wendy/wendy.py:112-157,336-517):
same names, arguments, defaults, lazy errors and yield tuples, for the
approximate (``approx=True``) integration path.  The simulation state is
uploaded once and stays on the GPU between yields; each ``next()`` runs one
output interval of ``nleap`` leapfrog steps there and copies the de-sorted
positions and velocities back into the generator's own ``x``, ``v`` arrays.
"""
import copy
import ctypes
import sys
import warnings

import numpy

from . import _lib
from .forces import ExtForce

# wendy/wendy.py:102 -- every name maps to the same device sort
_sort_type_dict = {'quick': 0, 'merge': 1, 'tim': 2, 'qsort': 3, 'parallel': 4}


class Gang:
    """Several large-path simulations advanced together (a ``wdy_gang``):
    one kernel launch per leapfrog step for all members whose next interval
    is a warm-start one, on the gang's own stream.  ``advance(sims, h,
    nleap)`` gives each member the result of ``sim.advance(h[i], nleap[i])``
    bit for bit; ``h`` and ``nleap`` may be scalars.  At most 64 members per
    call.  Members made with ``Sim(..., shared=True)`` keep the tile shape
    that suits many simulations sharing the GPU."""

    MAX = 64

    def __init__(self):
        self._g = None
        lib = _lib.lib()
        self._destroy = lib.wdy_gang_destroy
        g = ctypes.c_void_p()
        _lib.check(lib.wdy_gang_create(ctypes.byref(g)))
        self._g = g

    def advance(self, sims, h, nleap):
        k = len(sims)
        if k == 0:
            return
        hs = numpy.ascontiguousarray(numpy.broadcast_to(
            numpy.asarray(h, dtype=numpy.float64), (k,)))
        ns = numpy.broadcast_to(numpy.asarray(nleap, dtype=numpy.int64), (k,))
        handles = (ctypes.c_void_p * k)(*[s._h.value for s in sims])
        cn = (ctypes.c_int * k)(*[int(q) for q in ns])
        _lib.check(_lib.lib().wdy_gang_advance(self._g, handles, k, hs, cn))

    def close(self, _finalizing=sys.is_finalizing):
        if self._g is not None and not _finalizing():
            self._destroy(self._g)
        self._g = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Sim:
    """One device-resident approx simulation (a ``wdy_sim`` handle).

    ``m`` is already multiplied by twopiG and ``totmass`` is its
    ``numpy.sum``, as wendy/wendy.py:371,383 prepare them.
    """

    def __init__(self, x, v, m, totmass, omega2, equal_mass=None, path=None,
                 warm=True, shared=False):
        """``path``: None (by size), 'small' (one persistent workgroup per
        output interval) or 'large' (whole-GPU steps).  ``warm=False`` keeps
        the large path on the full radix sort every step.  ``shared=True``:
        the simulation will run in a ``Gang`` beside others, so it keeps the
        throughput tile shape at every size (WDY_SHARED_GPU)."""
        self._h = None
        lib = _lib.lib()
        self._destroy = lib.wdy_destroy  # held for close() at interpreter exit
        x = _lib.f64(x)
        v = _lib.f64(v)
        m = _lib.f64(m)
        if not (len(x) == len(v) == len(m)):
            raise ValueError('x, v and m must have the same length')
        if equal_mass is None:
            equal_mass = bool(len(m) > 0 and numpy.all(m == m[0]))
        flags = _lib.WDY_EQUAL_MASS if equal_mass else 0
        if path == 'large':
            flags |= _lib.WDY_FORCE_LARGE
        elif path == 'small':
            flags |= _lib.WDY_FORCE_SMALL
        elif path is not None:
            raise ValueError("path must be None, 'small' or 'large'")
        if not warm:
            flags |= _lib.WDY_NO_WARM
        if shared:
            flags |= _lib.WDY_SHARED_GPU
        h = ctypes.c_void_p()
        _lib.check(lib.wdy_create(len(x), x, v, m, float(totmass),
                                  float(omega2), flags, ctypes.byref(h)))
        self._h = h
        self.n = len(x)
        self.equal_mass = equal_mass
        self._m = m
        self._totmass = float(totmass)
        self._omega2 = float(omega2)

    def advance(self, h, nleap):
        _lib.check(_lib.lib().wdy_advance(self._h, float(h), int(nleap)))

    def read(self, x_out=None, v_out=None):
        """De-sort into ``x_out``/``v_out`` (float64, length n, C-contiguous);
        returns the device seconds of the last ``advance``."""
        sec = ctypes.c_double(0.)
        for a in (x_out, v_out):
            if a is not None and (a.dtype != numpy.float64 or a.shape != (self.n,)
                                  or not a.flags['C_CONTIGUOUS']
                                  or not a.flags['WRITEABLE']):
                raise ValueError('output arrays must be writeable float64[n]')
        xp = None if x_out is None else x_out.ctypes.data_as(ctypes.c_void_p)
        vp = None if v_out is None else v_out.ctypes.data_as(ctypes.c_void_p)
        _lib.check(_lib.lib().wdy_read(self._h, xp, vp, ctypes.byref(sec)))
        return sec.value

    def read_and_advance(self, x_out, v_out, h, nleap):
        """De-sort the current state into ``x_out``/``v_out``, starting the
        next interval (h, nleap) on the device while the copy runs (see
        wdy_read_and_advance); returns the device seconds of the interval
        read.  Arrays from ``_lib.pinned_empty`` let the copy overlap."""
        for a in (x_out, v_out):
            if a is not None and (a.dtype != numpy.float64 or a.shape != (self.n,)
                                  or not a.flags['C_CONTIGUOUS']
                                  or not a.flags['WRITEABLE']):
                raise ValueError('output arrays must be writeable float64[n]')
        sec = ctypes.c_double(0.)
        xp = None if x_out is None else x_out.ctypes.data_as(ctypes.c_void_p)
        vp = None if v_out is None else v_out.ctypes.data_as(ctypes.c_void_p)
        _lib.check(_lib.lib().wdy_read_and_advance(self._h, xp, vp, float(h),
                                                   int(nleap), ctypes.byref(sec)))
        return sec.value

    def energy(self, twopiG=1., omega=None):
        """Total energy of the live device state (masses = m / twopiG)."""
        e = ctypes.c_double(0.)
        omega2 = -1. if omega is None else omega ** 2.
        _lib.check(_lib.lib().wdy_energy(self._h, float(twopiG), omega2,
                                         ctypes.byref(e)))
        return e.value

    def tune(self, bucket_target=0, window=6, shape=-1):
        """Warm-start tile shape, bucket target and window (large path);
        applies from the next output interval (see wdy_tune)."""
        _lib.check(_lib.lib().wdy_tune(self._h, int(shape), int(bucket_target),
                                       int(window)))

    def path_stats(self):
        """(warm-start steps, radix steps, redone intervals) so far."""
        a, b, c = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.check(_lib.lib().wdy_path_stats(self._h, ctypes.byref(a),
                                             ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, h, nleap, candidates=((8, 6), (9, 6), (0, 8)), trials=2):
        """Pick the warm-start tile shape by timing: each candidate (shape,
        window) runs ``trials`` warm output intervals of (h, nleap) on a
        temporary copy of the current state; the fastest is applied to this
        simulation (from its next interval).  Returns {candidate: seconds per
        interval}.  Large path only; the copy is made from the current
        positions and velocities and the masses given at construction."""
        import time
        x = numpy.empty(self.n)
        v = numpy.empty(self.n)
        self.read(x, v)
        m = numpy.array(self._m, copy=True)
        times = {}
        for shape, window in candidates:
            tmp = Sim(x, v, m, self._totmass, self._omega2,
                      equal_mass=self.equal_mass, path='large')
            try:
                tmp.tune(0, window, shape)
                tmp.advance(h, nleap)          # re-layout (cold) interval
                tmp.read()
                t0 = time.perf_counter()
                for _ in range(trials):
                    tmp.advance(h, nleap)
                tmp.read()
                dt = (time.perf_counter() - t0) / trials
                if tmp.path_stats()[2]:         # redone: not a fit
                    dt = float('inf')
                times[(shape, window)] = dt
            finally:
                tmp.close()
        best = min(times, key=times.get)
        self.tune(0, best[1], best[0])
        return times

    def warm_config(self):
        """(tile shape, bucket target, window) of the warm path; shape -1
        without one (see wdy_warm_config)."""
        c = [ctypes.c_int(0) for _ in range(3)]
        _lib.check(_lib.lib().wdy_warm_config(self._h, *[ctypes.byref(x) for x in c]))
        return tuple(x.value for x in c)

    def warm_failures(self):
        """(capacity, overflow, stall): warm intervals redone, by cause."""
        c = [ctypes.c_int64(0) for _ in range(3)]
        _lib.check(_lib.lib().wdy_warm_failures(self._h, *[ctypes.byref(x) for x in c]))
        return tuple(x.value for x in c)

    def set_timing(self, on=True):
        _lib.check(_lib.lib().wdy_set_timing(self._h, int(bool(on))))

    def kernel_time(self, kernel_class):
        """(summed milliseconds, launch count) of one kernel class."""
        ms = ctypes.c_double(0.)
        cnt = ctypes.c_int64(0)
        _lib.check(_lib.lib().wdy_kernel_time(self._h, int(kernel_class),
                                              ctypes.byref(ms),
                                              ctypes.byref(cnt)))
        return ms.value, cnt.value

    def set_ext_force(self, force, t0=0.):
        """Evaluate ``force`` at every step from the next ``advance``, with
        the clock set to ``t0``: a wendy_amd.forces.ExtForce is evaluated on
        the device inside the step kernels; any other callable ``f(x, t)`` is
        called on the host before every step, as the reference calls its
        ``ext_force`` (the simulation must be on the large path); None
        removes the force."""
        if force is not None and not isinstance(force, ExtForce):
            cb = _lib.ext_force_callback(force)
            _lib.check(_lib.lib().wdy_set_ext_force_host(
                self._h, ctypes.cast(cb, ctypes.c_void_p), float(t0)))
            self._ext_cb = cb  # the C side holds its address
            return
        kind = 0 if force is None else force.kind
        p = numpy.array(force.params() if force is not None else (0., 0., 0.),
                        dtype=numpy.float64)
        _lib.check(_lib.lib().wdy_set_ext_force(self._h, kind, p, float(t0)))
        self._ext_cb = None

    def time(self):
        """The clock t0 (advances only while an external force is set)."""
        t = ctypes.c_double(0.)
        _lib.check(_lib.lib().wdy_time(self._h, ctypes.byref(t)))
        return t.value

    def close(self, _finalizing=sys.is_finalizing):
        # at interpreter exit (a generator collected then) the process
        # releases the device memory itself; the HIP runtime may be gone
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value and not _finalizing():
            self._destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nbody(x, v, m, dt, t0=0., twopiG=1., omega=None, ext_force=None,
          approx=False, nleap=None, sort='quick',
          maxcoll=100000, warn_maxcoll=False,
          full_output=False):
    """Run a 1-D N-body simulation; generator of (x, v) at intervals dt.

    Signature and defaults of wendy.nbody (wendy/wendy.py:112-115).  With
    ``approx=True`` each output is ``nleap`` drift-kick-drift leapfrog steps
    of size dt/nleap with exact forces, run on the GPU.  ``sort`` accepts the
    reference's names ('quick', 'merge', 'tim', 'qsort', 'parallel'); all of
    them use the device sort, which orders exactly tied positions by original
    index.  Yields ``(x, v)`` or, with ``full_output``, ``(x, v, seconds)``
    where seconds is the device time of the interval.  The yielded arrays are
    the generator's own (in page-locked host memory) and are overwritten by
    the next ``next()``, as in the reference; unlike the reference, editing
    them does not feed back into the simulation (its state lives on the GPU),
    and the device computes the next interval while the caller works on the
    current one.

    ``ext_force`` is any callable ``f(x, t)``, evaluated at the clock
    ``t0``, which advances by dt/nleap after every force evaluation, as in
    the reference.  The forms of ``wendy_amd.forces`` (Harmonic, Isothermal,
    Pulse) are evaluated on the device inside the step kernels; any other
    callable is called on the host before every leapfrog step with the
    positions of that step, as the reference calls it (one device-host round
    trip per step; the state stays on the device).
    ``approx=False`` runs the exact event-driven solver on the CPU
    (libwendy_amd's wdy_exact_*), the reference's own machine-precision
    method, with its checks and its ``maxcoll``/``warn_maxcoll`` behaviour;
    ``full_output`` then yields ``(x, v, ncoll, seconds)``.
    """
    if approx:
        if nleap is None:
            raise ValueError('When approx is True, the number of leapfrog steps '
                             'nleap= per output time step needs to be set')
        for item in _nbody_approx(x, v, m, dt, nleap, t0=t0, sort=sort,
                                  omega=omega, ext_force=ext_force,
                                  twopiG=twopiG, full_output=full_output):
            yield item
    for item in _nbody_exact(x, v, m, dt, twopiG=twopiG, omega=omega,
                             ext_force=ext_force, maxcoll=maxcoll,
                             warn_maxcoll=warn_maxcoll, full_output=full_output):
        yield item


def _nbody_exact(x, v, m, dt, twopiG=1., omega=None, ext_force=None,
                 maxcoll=100000, warn_maxcoll=False, full_output=False):
    """Exact path (wendy/wendy.py:158-238) on the CPU: same checks and
    messages, same masses (x twopiG, / omega**2), same maxcoll handling."""
    if ext_force is not None:
        raise ValueError('External forces (ext_force=) are not supported for '
                         'the exact nbody calculation')
    if omega is not None and omega * dt > numpy.pi / 2.:
        raise ValueError('When omega is set, omega*dt needs to be less than '
                         'pi/2; please adjust dt')
    if numpy.any(m < numpy.median(m) * 10. ** -8.):
        raise ValueError('Tiny masses m much smaller than the median mass are '
                         'not supported, please remove these from the inputs')
    x = numpy.require(copy.copy(x), dtype=numpy.float64, requirements=['C', 'W'])
    v = numpy.require(copy.copy(v), dtype=numpy.float64, requirements=['C', 'W'])
    m = numpy.require(twopiG * copy.copy(m), dtype=numpy.float64,
                      requirements=['C', 'W'])
    if omega is not None:
        m /= omega ** 2.
    lib = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(lib.wdy_exact_create(len(x), x, v, m,
                                    -1. if omega is None else float(omega),
                                    ctypes.byref(h)))
    ncoll = 0
    c = ctypes.c_int64(0)
    hit = ctypes.c_int(0)
    sec = ctypes.c_double(0.)
    try:
        while True:
            _lib.check(lib.wdy_exact_advance(h, float(dt), int(maxcoll),
                                             ctypes.byref(c), ctypes.byref(hit),
                                             ctypes.byref(sec)))
            ncoll += c.value
            _lib.check(lib.wdy_exact_read(h, x.ctypes.data_as(ctypes.c_void_p),
                                          v.ctypes.data_as(ctypes.c_void_p)))
            if hit.value:
                if not warn_maxcoll:
                    raise RuntimeError('Maximum number of collisions per time '
                                       'step exceeded')
                warnings.warn('Maximum number of collisions per time step '
                              'exceeded', RuntimeWarning)
                _lib.check(lib.wdy_exact_reset(h))
            if full_output:
                yield (x, v, ncoll, sec.value)
            else:
                yield (x, v)
    finally:
        lib.wdy_exact_destroy(h)


def _nbody_approx(x, v, m, dt, nleap, t0=0., omega=None, ext_force=None,
                  sort='quick', twopiG=1., full_output=False):
    """Approx path (wendy/wendy.py:336-437) with device-resident state."""
    host_force = ext_force is not None and not isinstance(ext_force, ExtForce)
    _sort_type_dict[sort]  # unknown names raise KeyError, as the reference
    if omega is None:
        omega2 = -1.
    else:
        omega2 = omega ** 2.
    x = copy.copy(x)
    v = copy.copy(v)
    m = twopiG * copy.copy(m)
    x = numpy.require(x, dtype=numpy.float64, requirements=['C', 'W'])
    v = numpy.require(v, dtype=numpy.float64, requirements=['C', 'W'])
    m = numpy.require(m, dtype=numpy.float64, requirements=['C', 'W'])
    totmass = numpy.sum(m)
    # a host-evaluated force runs on the large path's per-step launches
    sim = Sim(x, v, m, totmass, omega2,
              path='large' if host_force else None, warm=not host_force)
    if ext_force is not None:
        sim.set_ext_force(ext_force, t0)
    dt_leap = dt / nleap
    # the yielded arrays live in page-locked host memory: each next() copies
    # the finished interval out while the following one already runs
    try:
        xo = _lib.pinned_empty(len(x))
        vo = _lib.pinned_empty(len(x))
    except (MemoryError, _lib.WendyError):  # no page-locked memory: pageable
        xo = numpy.empty(len(x))
        vo = numpy.empty(len(x))
    try:
        sim.advance(dt_leap, nleap)
        while True:
            seconds = sim.read_and_advance(xo, vo, dt_leap, nleap)
            if full_output:
                yield (xo, vo, seconds)
            else:
                yield (xo, vo)
    finally:
        sim.close()


def energy(x, v, m, twopiG=1., individual=False, omega=None):
    """wendy.energy (wendy/wendy.py:439-475), computed on the GPU.

    The total is ``sum m v^2/2 + twopiG sum_{i<j} m_i m_j |x_i - x_j|``
    (+ ``sum m omega^2 x^2/2``), evaluated from one device sort of x and the
    exclusive mass prefix as ``twopiG sum_i m_i x_i (2 C_i + m_i - M)``.
    """
    lib = _lib.lib()
    x = _lib.f64(x)
    v = _lib.f64(v)
    m = _lib.f64(m)
    omega2 = -1. if omega is None else omega ** 2.
    if individual:
        out = numpy.empty(len(x))
        _lib.check(lib.wdy_energy_individual_arrays(len(x), x, v, m,
                                                    float(twopiG), omega2, out))
        return out
    e = ctypes.c_double(0.)
    _lib.check(lib.wdy_energy_arrays(len(x), x, v, m, float(twopiG), omega2,
                                     ctypes.byref(e)))
    return e.value


def momentum(v, m):
    """wendy.momentum (wendy/wendy.py:477-491): sum m v, on the GPU."""
    v = _lib.f64(v)
    m = _lib.f64(m)
    p = ctypes.c_double(0.)
    _lib.check(_lib.lib().wdy_momentum_arrays(len(v), v, m, ctypes.byref(p)))
    return p.value


def potential(y, x, v, m, twopiG=1., omega=None):
    """wendy.potential (wendy/wendy.py:493-517) at points y, on the GPU."""
    y = _lib.f64(numpy.atleast_1d(y))
    x = _lib.f64(x)
    m = _lib.f64(m)
    out = numpy.empty(len(y))
    omega2 = -1. if omega is None else omega ** 2.
    _lib.check(_lib.lib().wdy_potential_arrays(len(y), y, len(x), x, m,
                                               float(twopiG), omega2, out))
    return out
