"""Batched ensembles: many independent simulations advanced together.

No reference counterpart (the reference runs one simulation per ``nbody``
generator); each simulation's arithmetic is that of ``wendy_amd.nbody`` with
``approx=True``.  One kernel launch advances every simulation by one output
interval (see wdy_batch_create in include/wendy_amd.h).
"""
import ctypes
import sys

import numpy

from . import _lib


class Batch:
    """An ensemble of independent simulations of uneven size on the GPU.

    ``xs``, ``vs``, ``ms`` are sequences of per-simulation arrays (masses as
    given to ``nbody``, before the ``twopiG`` factor); ``omegas`` a sequence of
    per-simulation harmonic frequencies (None entries: no harmonic term).
    """

    def __init__(self, xs, vs, ms, omegas=None, twopiG=1.):
        self._h = None
        lib = _lib.lib()
        if not (len(xs) == len(vs) == len(ms)):
            raise ValueError('xs, vs and ms must have the same length')
        if omegas is None:
            omegas = [None] * len(xs)
        sizes = numpy.array([len(a) for a in xs], dtype=numpy.int64)
        for a, b, c in zip(xs, vs, ms):
            if not (len(a) == len(b) == len(c)):
                raise ValueError('each simulation needs x, v, m of one length')
        self.offsets = numpy.zeros(len(xs) + 1, dtype=numpy.int64)
        self.offsets[1:] = numpy.cumsum(sizes)
        x = _lib.f64(numpy.concatenate([numpy.asarray(a, float) for a in xs]))
        v = _lib.f64(numpy.concatenate([numpy.asarray(a, float) for a in vs]))
        mk = [twopiG * numpy.asarray(a, dtype=numpy.float64) for a in ms]
        m = _lib.f64(numpy.concatenate(mk))
        tot = _lib.f64([numpy.sum(a) for a in mk])   # wendy/wendy.py:383
        om2 = _lib.f64([-1. if o is None else o ** 2. for o in omegas])
        h = ctypes.c_void_p()
        _lib.check(lib.wdy_batch_create(len(xs), self.offsets, x, v, m, tot, om2,
                                        ctypes.byref(h)))
        self._h = h
        self.n = int(self.offsets[-1])
        self.nsims = len(xs)

    def advance(self, h, nleap):
        """Enqueue one output interval (nleap steps of size h) of every
        simulation."""
        _lib.check(_lib.lib().wdy_batch_advance(self._h, float(h), int(nleap)))

    def read(self):
        """(list of x arrays, list of v arrays, device seconds of the last
        advance)."""
        x = numpy.empty(self.n)
        v = numpy.empty(self.n)
        sec = ctypes.c_double(0.)
        _lib.check(_lib.lib().wdy_batch_read(
            self._h, x.ctypes.data_as(ctypes.c_void_p),
            v.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sec)))
        o = self.offsets
        return ([x[o[k]:o[k + 1]] for k in range(self.nsims)],
                [v[o[k]:o[k + 1]] for k in range(self.nsims)], sec.value)

    def sync(self):
        sec = ctypes.c_double(0.)
        _lib.check(_lib.lib().wdy_batch_read(self._h, None, None,
                                             ctypes.byref(sec)))
        return sec.value

    def close(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        h, self._h = getattr(self, '_h', None), None
        if h is not None and h.value and not _finalizing():
            _lib.lib().wdy_batch_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
