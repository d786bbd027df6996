"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/wendy_amd.h declares, and the Python API reproduces the
reference's argument handling (lazy errors, error types, messages)."""
import ctypes
import os
import re

import numpy
import pytest

import wendy_amd
from wendy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'wendy_amd.h')


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    text = re.sub(r'^\s*#.*$', '', text, flags=re.M)  # preprocessor lines
    names = re.findall(r'\b(\w+)\s*\([^;{]*\)\s*;', text)
    return sorted(set(n for n in names
                      if n.startswith('wdy_') or n.startswith('_wendy')
                      or n.startswith('_solve_coll')))


def have_gpu():
    n = ctypes.c_int(0)
    return _lib.lib().wdy_device_count(ctypes.byref(n)) == 0 and n.value > 0


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), 'run python -m wendy_amd.build'
    lib = _lib.lib()
    assert lib.wdy_abi_version() == 1


def test_every_declared_symbol_is_exported():
    declared = header_functions()
    assert '_wendy_nbody_approx_onestep' in declared
    assert set(declared) == set(_lib.EXPORTED)
    handle = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(handle, name), name


def test_no_device_reports_enodev():
    if have_gpu():
        pytest.skip('a GPU is visible')
    lib = _lib.lib()
    x = numpy.zeros(4)
    h = ctypes.c_void_p()
    rc = lib.wdy_create(4, x, x, x + 1., 4., -1., 1, ctypes.byref(h))
    assert rc == _lib.WDY_ENODEV
    assert b'GPU' in lib.wdy_last_error()
    with pytest.raises(_lib.WendyError):
        wendy_amd.Sim(x, x, x + 1., 4., -1.)


def test_create_rejects_bad_arguments_before_touching_device():
    lib = _lib.lib()
    x = numpy.array([0., 1., numpy.nan])
    h = ctypes.c_void_p()
    assert lib.wdy_create(3, x, x, x, 1., -1., 0, ctypes.byref(h)) == _lib.WDY_EINVAL
    assert b'non-finite' in lib.wdy_last_error()
    y = numpy.array([0., 1., 2.])
    assert lib.wdy_create(0, y, y, y, 1., -1., 0, ctypes.byref(h)) == _lib.WDY_EINVAL
    assert lib.wdy_create(3, y, y, y, 1., -1., 64, ctypes.byref(h)) == _lib.WDY_EINVAL
    # WDY_EQUAL_MASS with unequal masses is refused
    assert lib.wdy_create(3, y, y, y, 1., -1., 1, ctypes.byref(h)) == _lib.WDY_EINVAL
    with pytest.raises(ValueError):
        _lib.check(_lib.WDY_EINVAL)


def test_nleap_error_is_lazy_with_reference_message():
    """tests/test_approx.py:187-196 of the reference."""
    x = numpy.array([-1., 1.])
    v = numpy.array([0., 0.])
    m = numpy.array([1., 1.])
    g = wendy_amd.nbody(x, v, m, 2, approx=True)
    with pytest.raises(ValueError) as excinfo:
        next(g)
    assert str(excinfo.value) == ('When approx is True, the number of leapfrog '
                                  'steps nleap= per output time step needs to '
                                  'be set')


def test_unknown_sort_raises_keyerror_lazily():
    g = wendy_amd.nbody(numpy.zeros(2), numpy.zeros(2), numpy.ones(2), 0.1,
                        approx=True, nleap=2, sort='bogus')
    with pytest.raises(KeyError):
        next(g)


def test_ext_force_callback_follows_reference_convention():
    """A Python ext_force is handed to the library as the reference's C
    callback (wendy/wendy.py:406-418): one sheet -> the return value; N
    sheets -> the values stored into a[:N], return 0."""
    calls = []

    def f(x, t):
        calls.append(t)
        return -2. * x + t

    cb = _lib.ext_force_callback(f)
    one = ctypes.c_double(1.5)
    assert cb(1, ctypes.pointer(one), 0.25, None) == -2.75
    x = numpy.array([0., 1., -3.])
    a = numpy.zeros(3)
    r = cb(3, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0.5,
           a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert r == 0.
    assert numpy.array_equal(a, -2. * x + 0.5)
    assert calls == [0.25, 0.5]


def test_compat_symbol_rejects_bad_sort_and_poisons_state():
    """An unknown sort_type (the reference silently skips its sort) is
    refused before the device is touched; err is set and, since the
    reference's caller never reads err, the handed-back x, v and xi[].val
    are NaN rather than an un-advanced state."""
    lib = ctypes.CDLL(_lib.LIB_PATH)
    f = lib._wendy_nbody_approx_onestep
    f.restype = None
    xi = numpy.zeros(2, dtype=[('idx', '<i4'), ('pad', '<i4'), ('val', '<f8')])
    xi['idx'] = [0, 1]
    x = numpy.array([0.5, -0.5])
    v = numpy.array([1., 2.])
    z = numpy.zeros(2)
    m = numpy.ones(2)
    err = ctypes.c_int(0)
    te = ctypes.c_double(0)
    t0 = ctypes.c_double(0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    f(ctypes.c_int(2), p(xi), p(x), p(v), p(m), p(z), ctypes.c_double(2.),
      ctypes.c_double(0.1), ctypes.c_int(2), ctypes.byref(t0),
      ctypes.c_double(-1.), None, ctypes.c_int(7), ctypes.byref(err),
      ctypes.byref(te), p(z))
    assert err.value == _lib.WDY_EINVAL
    assert numpy.all(numpy.isnan(x)) and numpy.all(numpy.isnan(v))
    assert numpy.all(numpy.isnan(xi['val']))
    assert numpy.array_equal(m, numpy.ones(2))


def test_batch_create_rejects_bad_arguments_before_touching_device():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    y = numpy.array([0., 1., 2., 3.])
    one = numpy.array([1., 1.])
    om = numpy.array([-1., -1.])
    ok = numpy.array([0, 2, 4], dtype=numpy.int64)
    for off in ([1, 2, 4], [0, 2, 2], [0, 3, 2]):
        rc = lib.wdy_batch_create(2, numpy.array(off, dtype=numpy.int64),
                                  y, y, y, one, om, ctypes.byref(h))
        assert rc == _lib.WDY_EINVAL, off
    bad = y.copy()
    bad[3] = numpy.inf
    assert lib.wdy_batch_create(2, ok, bad, y, y, one, om,
                                ctypes.byref(h)) == _lib.WDY_EINVAL
    neg = -y
    neg[0] = -1.
    assert lib.wdy_batch_create(2, ok, y, y, neg, one, om,
                                ctypes.byref(h)) == _lib.WDY_EINVAL
    assert lib.wdy_batch_create(2, ok, y, y, y, numpy.array([1., numpy.nan]),
                                om, ctypes.byref(h)) == _lib.WDY_EINVAL
    assert lib.wdy_batch_create(0, ok, y, y, y, one, om,
                                ctypes.byref(h)) == _lib.WDY_EINVAL
    assert h.value is None


def test_exact_create_rejects_bad_arguments():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    y = numpy.array([0., 1., 2.])
    one = numpy.ones(3)
    assert lib.wdy_exact_create(0, y, y, one, -1., ctypes.byref(h)) == _lib.WDY_EINVAL
    assert lib.wdy_exact_create(3, y, y, numpy.array([1., 0., 1.]), -1.,
                                ctypes.byref(h)) == _lib.WDY_EINVAL
    assert lib.wdy_exact_create(3, numpy.array([0., numpy.nan, 1.]), y, one, -1.,
                                ctypes.byref(h)) == _lib.WDY_EINVAL
    assert h.value is None
    # a valid one works without any GPU (the exact solver is host code)
    assert lib.wdy_exact_create(3, y, numpy.zeros(3), one, -1., ctypes.byref(h)) == 0
    n = ctypes.c_int64(-1)
    hit = ctypes.c_int(-1)
    sec = ctypes.c_double(-1.)
    assert lib.wdy_exact_advance(h, 0., 10, ctypes.byref(n), ctypes.byref(hit),
                                 ctypes.byref(sec)) == _lib.WDY_EINVAL
    assert lib.wdy_exact_advance(h, 0.1, 10, ctypes.byref(n), ctypes.byref(hit),
                                 ctypes.byref(sec)) == 0
    assert n.value == 0 and hit.value == 0
    assert lib.wdy_exact_destroy(h) == 0


def test_gpu_handles_report_no_device_here():
    """Without a GPU the pool and the page-locked allocator fail cleanly."""
    lib = _lib.lib()
    n = ctypes.c_int(0)
    if lib.wdy_device_count(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip('a GPU is visible')
    h = ctypes.c_void_p()
    assert lib.wdy_pool_create(1000, 4, 100, ctypes.byref(h)) == _lib.WDY_ENODEV
    assert lib.wdy_pool_create(10, 4, 100, ctypes.byref(h)) == _lib.WDY_EINVAL
    assert h.value is None
