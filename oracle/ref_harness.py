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


def load(variant=''):
    if variant not in _libs:
        lib = ctypes.CDLL(lib_path(variant))
        f = lib._wendy_nbody_approx_onestep
        f.restype = None
        f.argtypes = [ctypes.c_int, _XI, _F64, _F64, _F64, _F64,
                      ctypes.c_double, ctypes.c_double, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                      ctypes.c_void_p, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double), _F64]
        _libs[variant] = lib
    return _libs[variant]


class RefApprox:
    """One approx simulation run by the compiled reference.

    Argument preparation follows wendy/wendy.py:363-387 and :422; each
    ``onestep()`` is one ``next()`` of the reference generator
    (wendy/wendy.py:424-437).
    """

    def __init__(self, x, v, m, dt, nleap, omega=None, twopiG=1.,
                 sort='quick', variant='', ext_force=None, t0=0.):
        self.lib = load(variant)
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
