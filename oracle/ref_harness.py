"""ctypes driver for the compiled reference approx step (test infrastructure).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Loads ``oracle/_ref/libwendy_ref*.so`` (built by ``oracle/Makefile`` from the
reference's own C sources) and drives ``_wendy_nbody_approx_onestep``
(wendy/wendy.c:385-418, prototype wendy/wendy.h:29-34),
preparing its arguments the way wendy/wendy.py:363-433 does; an external
force is passed through the plain-ctypes wrapper of wendy/wendy.py:406-418.  This does not
import the reference's Python package (which needs numba); it restates the
argument preparation in a few lines.
"""
import copy
import ctypes
import os

import numpy
from numpy.ctypeslib import ndpointer

from .nbody_oracle import SORT_TYPES

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, '_ref')

# struct array_w_index {int idx; double val;} -- 16 B, val at offset 8
# (wendy/wendy.h:12-16, ctypes mirror wendy/wendy.py:73-75)
XI_DTYPE = numpy.dtype([('idx', '<i4'), ('pad', '<i4'), ('val', '<f8')])

_F64 = ndpointer(dtype=numpy.float64, flags=('C_CONTIGUOUS', 'WRITEABLE'))
_XI = ndpointer(dtype=XI_DTYPE, flags=('C_CONTIGUOUS', 'WRITEABLE'))


def lib_path(variant=''):
    return os.path.join(REF_DIR, 'libwendy_ref%s.so' % variant)


def available(variant=''):
    return os.path.exists(lib_path(variant))


_libs = {}

_I32 = ndpointer(dtype=numpy.int32, flags=('C_CONTIGUOUS', 'WRITEABLE'))


def bind(lib):
    """Declares on ``lib`` the five entry points the reference's loader binds
    at import, with its argument and return types (wendy/wendy.py:38-101),
    and returns ``lib``.  Any library that exports them can be bound."""
    exact = [ctypes.c_int, _F64, _F64, _F64, _F64, _I32,
             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
             _F64, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    lib._wendy_nbody_onestep.restype = None
    lib._wendy_nbody_onestep.argtypes = exact
    lib._wendy_nbody_harm_onestep.restype = None
    lib._wendy_nbody_harm_onestep.argtypes = exact + [ctypes.c_double]
    f = lib._wendy_nbody_approx_onestep
    f.restype = None
    f.argtypes = [ctypes.c_int, _XI, _F64, _F64, _F64, _F64,
                  ctypes.c_double, ctypes.c_double, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                  ctypes.c_void_p, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_int),
                  ctypes.POINTER(ctypes.c_double), _F64]
    lib._solve_coll_quad.restype = ctypes.c_double
    lib._solve_coll_quad.argtypes = [ctypes.c_double] * 3
    lib._solve_coll_harm.restype = ctypes.c_double
    lib._solve_coll_harm.argtypes = [ctypes.c_double] * 4
    return lib


def load(variant=''):
    if variant not in _libs:
        _libs[variant] = bind(ctypes.CDLL(lib_path(variant)))
    return _libs[variant]


class RefApprox:
    """One approx simulation run by the compiled reference.

    Argument preparation follows wendy/wendy.py:363-387 and :422; each
    ``onestep()`` is one ``next()`` of the reference generator
    (wendy/wendy.py:424-437).
    """

    def __init__(self, x, v, m, dt, nleap, omega=None, twopiG=1.,
                 sort='quick', variant='', ext_force=None, t0=0., lib=None):
        self.lib = load(variant) if lib is None else lib
        self.ext_c = None
        if ext_force is not None:
            proto = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double),
                                     ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_double))

            def arrays(nn, xp, t, ap):
                if nn == 1:
                    return ext_force(xp.contents.value, t)
                xa = numpy.ctypeslib.as_array(xp, shape=(nn,))
                aa = numpy.ctypeslib.as_array(ap, shape=(nn,))
                aa[:] = ext_force(xa, t)
                return 0.
            self.ext_c = proto(arrays)
        self.omega2 = -1. if omega is None else omega ** 2.
        n = len(x)
        self.n = n
        self.x = numpy.require(copy.copy(x), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.v = numpy.require(copy.copy(v), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.m = numpy.require(twopiG * copy.copy(m), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.a = numpy.zeros(n)
        self.cumulmass = numpy.zeros(n)
        self.totmass = numpy.sum(self.m)
        self.xi = numpy.zeros(n, dtype=XI_DTYPE)
        self.xi['idx'] = numpy.arange(n, dtype=numpy.int32)
        self.xi['val'] = self.x
        self.t0 = ctypes.c_double(t0)
        self.err = ctypes.c_int(0)
        self.time_elapsed = ctypes.c_double(0.)
        self.h = dt / nleap
        self.nleap = int(nleap)
        self.sort_type = SORT_TYPES[sort]

    def onestep(self):
        self.lib._wendy_nbody_approx_onestep(
            self.n, self.xi, self.x, self.v, self.m, self.a, self.totmass,
            self.h, self.nleap, ctypes.byref(self.t0), self.omega2,
            None if self.ext_c is None else ctypes.cast(self.ext_c, ctypes.c_void_p),
            self.sort_type, ctypes.byref(self.err),
            ctypes.byref(self.time_elapsed), self.cumulmass)
        return self.x, self.v, self.time_elapsed.value


class RefExact:
    """The compiled reference's exact solver.

    Set-up restated from wendy/wendy.py:166-174 and :243-279 (argsort,
    running mass sums, crossing times through the library's own
    _solve_coll_quad/_solve_coll_harm); each ``onestep()`` is one
    ``_wendy_nbody_onestep`` / ``_wendy_nbody_harm_onestep`` call
    (wendy/wendy.c:53-315).
    """

    def __init__(self, x, v, m, dt, omega=None, twopiG=1., maxcoll=100000,
                 variant='', lib=None):
        self.lib = load(variant) if lib is None else lib
        self.omega = omega
        self.dt = dt
        self.maxcoll = maxcoll
        self.x = numpy.require(copy.copy(x), dtype=numpy.float64, requirements=['C', 'W'])
        self.v = numpy.require(copy.copy(v), dtype=numpy.float64, requirements=['C', 'W'])
        self.m = numpy.require(twopiG * copy.copy(m), dtype=numpy.float64,
                               requirements=['C', 'W'])
        if omega is not None:
            self.m /= omega ** 2.
        self.setup()
        self.ncoll = 0

    def setup(self):
        x, v, m = self.x, self.v, self.m
        sindx = numpy.argsort(x)
        below = numpy.cumsum(m[sindx])
        below[-1] = 0.
        below = numpy.roll(below, 1)
        above = numpy.cumsum(m[sindx][::-1])[::-1]
        above[0] = 0.
        above = numpy.roll(above, -1)
        a = (above - below)[numpy.argsort(sindx)]
        xs, vs, as_ = x[sindx], v[sindx], a[sindx]
        tc = []
        for k in range(len(x) - 1):
            if self.omega is None:
                tc.append(self.lib._solve_coll_quad(xs[k] - xs[k + 1], vs[k] - vs[k + 1],
                                                    (as_[k] - as_[k + 1]) / 2.))
            else:
                tc.append(self.lib._solve_coll_harm(xs[k] - xs[k + 1], vs[k] - vs[k + 1],
                                                    as_[k] - as_[k + 1], self.omega))
        self.tcoll = numpy.require(numpy.array(tc), dtype=numpy.float64,
                                   requirements=['C', 'W'])
        self.a = numpy.require(a, dtype=numpy.float64, requirements=['C', 'W'])
        self.sindx = numpy.require(sindx, dtype=numpy.int32, requirements=['C', 'W'])
        self.cindx = ctypes.c_int(int(numpy.argmin(self.tcoll)))
        self.next_tcoll = ctypes.c_double(self.tcoll[self.cindx.value])
        self.err = ctypes.c_int(0)

    def onestep(self):
        nc = ctypes.c_int(0)
        te = ctypes.c_double(0.)
        args = [len(self.x), self.x, self.v, self.a, self.m, self.sindx,
                ctypes.byref(self.cindx), ctypes.byref(self.next_tcoll), self.tcoll,
                self.dt, self.maxcoll, ctypes.byref(self.err), ctypes.byref(nc),
                ctypes.byref(te)]
        if self.omega is None:
            self.lib._wendy_nbody_onestep(*args)
        else:
            self.lib._wendy_nbody_harm_onestep(*args, self.omega)
        self.ncoll += nc.value
        return self.x, self.v, self.ncoll, self.err.value
