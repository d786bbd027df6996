"""ctypes binding of libwendy_amd.so (the C ABI in include/wendy_amd.h).

The library is loaded by absolute path from this package directory, the way
the reference loads its C extension with ``ctypes.CDLL``
(wendy/wendy.py:13-37).  Unlike the reference, which only warns when its
extension is missing and then fails on first use, a missing library raises
ImportError-like ``RuntimeError`` immediately on first use: there is no CPU
fallback.
"""
import ctypes
import sys
import os

import numpy
from numpy.ctypeslib import ndpointer

HERE = os.path.dirname(os.path.abspath(__file__))
# WENDY_AMD_LIB names another build of the library (e.g. to compare builds)
LIB_PATH = os.environ.get('WENDY_AMD_LIB') or os.path.join(HERE, 'libwendy_amd.so')

WDY_OK = 0
WDY_EINVAL = -1
WDY_ENOMEM = -2
WDY_ERUNTIME = -3
WDY_EVERIFY = -4
WDY_ENODEV = -5
WDY_EUNSUPPORTED = -6
WDY_EQUAL_MASS = 1
WDY_FORCE_LARGE = 2
WDY_FORCE_SMALL = 4
WDY_NO_WARM = 8
WDY_SHARED_GPU = 16

# every C symbol include/wendy_amd.h declares
EXPORTED = ('wdy_abi_version', 'wdy_last_error', 'wdy_device_count',
            'wdy_create', 'wdy_advance', 'wdy_read', 'wdy_energy',
            'wdy_destroy', 'wdy_tune', 'wdy_path_stats',
            'wdy_gang_create', 'wdy_gang_destroy', 'wdy_gang_advance',
            'wdy_set_timing', 'wdy_kernel_time',
            'wdy_energy_arrays', 'wdy_energy_individual_arrays',
            'wdy_momentum_arrays', 'wdy_potential_arrays',
            'wdy_energy_device_noalloc',
            'wdy_set_ext_force', 'wdy_set_ext_force_host', 'wdy_time', 'wdy_warm_failures', 'wdy_warm_config',
            'wdy_read_and_advance', 'wdy_host_alloc', 'wdy_host_free',
            'wdy_copy_probe',
            'wdy_exact_create', 'wdy_exact_advance', 'wdy_exact_reset',
            'wdy_exact_read', 'wdy_exact_destroy',
            'wdy_batch_create', 'wdy_batch_advance', 'wdy_batch_read',
            'wdy_batch_destroy',
            'wdy_pool_create', 'wdy_pool_add', 'wdy_pool_step', 'wdy_pool_read',
            'wdy_pool_remove', 'wdy_pool_sync', 'wdy_pool_stats',
            'wdy_pool_destroy',
            '_wendy_nbody_approx_onestep', '_wendy_nbody_onestep',
            '_wendy_nbody_harm_onestep', '_solve_coll_quad',
            '_solve_coll_harm')

_F64 = ndpointer(dtype=numpy.float64, flags=('C_CONTIGUOUS',))
_F64W = ndpointer(dtype=numpy.float64, flags=('C_CONTIGUOUS', 'WRITEABLE'))
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)

_lib = None


class WendyError(RuntimeError):
    """A nonzero return code from libwendy_amd.so."""

    def __init__(self, code, msg):
        super().__init__('%s (code %d)' % (msg, code))
        self.code = code


def _declare(lib):
    S = ctypes.c_void_p
    lib.wdy_abi_version.restype = ctypes.c_int
    lib.wdy_abi_version.argtypes = []
    lib.wdy_last_error.restype = ctypes.c_char_p
    lib.wdy_last_error.argtypes = []
    lib.wdy_device_count.restype = ctypes.c_int
    lib.wdy_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.wdy_create.restype = ctypes.c_int
    lib.wdy_create.argtypes = [ctypes.c_int64, _F64, _F64, _F64,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(S)]
    lib.wdy_advance.restype = ctypes.c_int
    lib.wdy_advance.argtypes = [S, ctypes.c_double, ctypes.c_int]
    lib.wdy_read.restype = ctypes.c_int
    lib.wdy_read.argtypes = [S, S, S, _c_double_p]
    lib.wdy_energy.restype = ctypes.c_int
    lib.wdy_energy.argtypes = [S, ctypes.c_double, ctypes.c_double, _c_double_p]
    lib.wdy_destroy.restype = ctypes.c_int
    lib.wdy_destroy.argtypes = [S]
    lib.wdy_gang_create.restype = ctypes.c_int
    lib.wdy_gang_create.argtypes = [ctypes.POINTER(S)]
    lib.wdy_gang_destroy.restype = ctypes.c_int
    lib.wdy_gang_destroy.argtypes = [S]
    lib.wdy_gang_advance.restype = ctypes.c_int
    lib.wdy_gang_advance.argtypes = [S, ctypes.POINTER(S), ctypes.c_int,
                                     _F64, ctypes.POINTER(ctypes.c_int)]
    lib.wdy_tune.restype = ctypes.c_int
    lib.wdy_tune.argtypes = [S, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.wdy_path_stats.restype = ctypes.c_int
    lib.wdy_path_stats.argtypes = [S, _c_int64_p, _c_int64_p, _c_int64_p]
    lib.wdy_set_timing.restype = ctypes.c_int
    lib.wdy_set_timing.argtypes = [S, ctypes.c_int]
    lib.wdy_kernel_time.restype = ctypes.c_int
    lib.wdy_kernel_time.argtypes = [S, ctypes.c_int, _c_double_p, _c_int64_p]
    lib.wdy_energy_arrays.restype = ctypes.c_int
    lib.wdy_energy_arrays.argtypes = [ctypes.c_int64, _F64, _F64, _F64,
                                      ctypes.c_double, ctypes.c_double,
                                      _c_double_p]
    lib.wdy_energy_individual_arrays.restype = ctypes.c_int
    lib.wdy_energy_individual_arrays.argtypes = [
        ctypes.c_int64, _F64, _F64, _F64, ctypes.c_double, ctypes.c_double,
        _F64W]
    lib.wdy_energy_device_noalloc.restype = ctypes.c_int
    lib.wdy_energy_device_noalloc.argtypes = [ctypes.c_int64, S, S, S,
                                              ctypes.c_double, ctypes.c_double,
                                              S, S]
    lib.wdy_momentum_arrays.restype = ctypes.c_int
    lib.wdy_momentum_arrays.argtypes = [ctypes.c_int64, _F64, _F64,
                                        _c_double_p]
    _I64 = ndpointer(dtype=numpy.int64, flags=('C_CONTIGUOUS',))
    lib.wdy_batch_create.restype = ctypes.c_int
    lib.wdy_batch_create.argtypes = [ctypes.c_int64, _I64, _F64, _F64, _F64,
                                     _F64, _F64, ctypes.POINTER(S)]
    lib.wdy_batch_advance.restype = ctypes.c_int
    lib.wdy_batch_advance.argtypes = [S, ctypes.c_double, ctypes.c_int]
    lib.wdy_batch_read.restype = ctypes.c_int
    lib.wdy_batch_read.argtypes = [S, S, S, _c_double_p]
    lib.wdy_batch_destroy.restype = ctypes.c_int
    lib.wdy_batch_destroy.argtypes = [S]
    lib.wdy_exact_create.restype = ctypes.c_int
    lib.wdy_exact_create.argtypes = [ctypes.c_int64, _F64, _F64, _F64, ctypes.c_double,
                                     ctypes.POINTER(S)]
    lib.wdy_exact_advance.restype = ctypes.c_int
    lib.wdy_exact_advance.argtypes = [S, ctypes.c_double, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int), _c_double_p]
    lib.wdy_exact_reset.restype = ctypes.c_int
    lib.wdy_exact_reset.argtypes = [S]
    lib.wdy_exact_read.restype = ctypes.c_int
    lib.wdy_exact_read.argtypes = [S, S, S]
    lib.wdy_exact_destroy.restype = ctypes.c_int
    lib.wdy_exact_destroy.argtypes = [S]
    lib.wdy_read_and_advance.restype = ctypes.c_int
    lib.wdy_read_and_advance.argtypes = [S, S, S, ctypes.c_double, ctypes.c_int,
                                         _c_double_p]
    lib.wdy_copy_probe.restype = ctypes.c_int
    lib.wdy_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int, _c_double_p]
    lib.wdy_host_alloc.restype = ctypes.c_int
    lib.wdy_host_alloc.argtypes = [ctypes.c_int64, ctypes.POINTER(S)]
    lib.wdy_host_free.restype = ctypes.c_int
    lib.wdy_host_free.argtypes = [S]
    lib.wdy_warm_failures.restype = ctypes.c_int
    lib.wdy_warm_failures.argtypes = [S] + [ctypes.POINTER(ctypes.c_int64)] * 3
    lib.wdy_warm_config.restype = ctypes.c_int
    lib.wdy_warm_config.argtypes = [S] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.wdy_set_ext_force.restype = ctypes.c_int
    lib.wdy_set_ext_force.argtypes = [S, ctypes.c_int, _F64, ctypes.c_double]
    lib.wdy_set_ext_force_host.restype = ctypes.c_int
    lib.wdy_set_ext_force_host.argtypes = [S, ctypes.c_void_p, ctypes.c_double]
    lib.wdy_time.restype = ctypes.c_int
    lib.wdy_time.argtypes = [S, _c_double_p]
    _I32 = ndpointer(dtype=numpy.int32, flags=('C_CONTIGUOUS',))
    _c_int_p = ctypes.POINTER(ctypes.c_int)
    _c_i64_p = ctypes.POINTER(ctypes.c_int64)
    lib.wdy_pool_create.restype = ctypes.c_int
    lib.wdy_pool_create.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int64,
                                    ctypes.POINTER(S)]
    lib.wdy_pool_add.restype = ctypes.c_int
    lib.wdy_pool_add.argtypes = [S, ctypes.c_int64, _F64, _F64, _F64,
                                 ctypes.c_double, ctypes.c_double, _c_int_p]
    lib.wdy_pool_step.restype = ctypes.c_int
    lib.wdy_pool_step.argtypes = [S, ctypes.c_int, _I32, _F64, _I32]
    lib.wdy_pool_read.restype = ctypes.c_int
    lib.wdy_pool_read.argtypes = [S, ctypes.c_int, S, S]
    lib.wdy_pool_remove.restype = ctypes.c_int
    lib.wdy_pool_remove.argtypes = [S, ctypes.c_int]
    lib.wdy_pool_sync.restype = ctypes.c_int
    lib.wdy_pool_sync.argtypes = [S, _c_double_p]
    lib.wdy_pool_stats.restype = ctypes.c_int
    lib.wdy_pool_stats.argtypes = [S, _c_int_p, _c_i64_p, _c_i64_p]
    lib.wdy_pool_destroy.restype = ctypes.c_int
    lib.wdy_pool_destroy.argtypes = [S]
    lib.wdy_potential_arrays.restype = ctypes.c_int
    lib.wdy_potential_arrays.argtypes = [ctypes.c_int64, _F64, ctypes.c_int64,
                                         _F64, _F64, ctypes.c_double,
                                         ctypes.c_double, _F64W]


def lib():
    """The loaded library; raises RuntimeError if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                'libwendy_amd.so not found at %s: build it with '
                '`python -m wendy_amd.build` (there is no CPU fallback)'
                % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle)
        _lib = handle
    return _lib


def check(rc):
    """Map a WDY_E* return code to a Python exception."""
    if rc == WDY_OK:
        return
    msg = lib().wdy_last_error().decode(errors='replace')
    if rc == WDY_EINVAL:
        raise ValueError(msg)
    if rc == WDY_ENOMEM:
        raise MemoryError(msg)
    raise WendyError(rc, msg)


# the reference's callback type for an external force (wendy/wendy.py:406-410)
EXT_FORCE_CFUNC = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double),
                                   ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double))


def ext_force_callback(f):
    """Wraps a Python callable ``f(x, t)`` as the C callback the library
    calls before each step, with the reference's conventions
    (wendy/wendy.py:411-418): for one sheet it returns ``f(x, t)``, for N
    sheets it stores ``f(x, t)`` into a[:N]."""
    from numpy.ctypeslib import as_array

    def arrays(n, xp, t, ap):
        if n == 1:
            return float(f(xp.contents.value, t))
        xa = as_array(xp, shape=(n,))
        aa = as_array(ap, shape=(n,))
        aa[:] = f(xa, t)
        return 0.
    return EXT_FORCE_CFUNC(arrays)


def f64(a):
    return numpy.ascontiguousarray(a, dtype=numpy.float64)


class _HostBlock:
    """Page-locked host memory from wdy_host_alloc, freed when the last array
    viewing it is gone."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        self.ptr = None
        self._free = lib().wdy_host_free
        check(lib().wdy_host_alloc(int(nbytes), ctypes.byref(p)))
        self.ptr = p.value

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the memory itself
        try:
            if self.ptr and not _finalizing():
                self._free(self.ptr)
        except Exception:
            pass
        self.ptr = None
        self.ptr = None


def pinned_empty(n):
    """A float64[n] numpy array in page-locked host memory (device-to-host
    copies into it run at full speed and asynchronously)."""
    n = int(n)
    blk = _HostBlock(8 * max(n, 1))
    carr = (ctypes.c_double * max(n, 1)).from_address(blk.ptr)
    carr._owner = blk
    return numpy.ctypeslib.as_array(carr)[:n]
