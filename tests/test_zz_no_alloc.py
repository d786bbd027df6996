"""wdy_energy_device_noalloc, the allocation-free form of wdy_energy_arrays
(no reference counterpart), called through the C ABI on device arrays and
compared bit for bit with wdy_energy_arrays on the same host arrays: the two
run the same kernels in the same order, only the origin of their scratch
differs.  Calls follow one another on the same device, a smaller input after
a larger one and a larger one after that, so the second and third calls find
the static scratch holding the previous call's contents."""
import ctypes

import numpy
import pytest

import wendy_amd
from conftest import bits_equal
from oracle import nbody_oracle as O
from wendy_amd import _lib

pytestmark = pytest.mark.gpu


class _Hip:
    """Device memory and a stream from the HIP runtime libwendy_amd.so itself
    uses (the same library object: a second runtime in the process, such as
    the one bundled with torch, would not see the GPU after it)."""

    def __init__(self):
        path = None
        with open('/proc/self/maps') as f:
            for line in f:
                if 'libamdhip64.so' in line:
                    path = line.split()[-1]
                    break
        assert path, 'libwendy_amd.so has not loaded the HIP runtime'
        self.lib = ctypes.CDLL(path)
        self.ptrs = []

    def check(self, rc):
        assert rc == 0, 'HIP runtime call failed: %d' % rc

    def upload(self, a):
        a = numpy.ascontiguousarray(a, dtype=numpy.float64)
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)))
        self.ptrs.append(p)
        self.check(self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(a.nbytes), 1))  # H2D
        return p

    def download(self, p, n):
        out = numpy.empty(n)
        self.check(self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p,
                                      ctypes.c_size_t(out.nbytes), 2))  # D2H
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def _noalloc(hip, x, v, m, twopiG, omega, stream):
    xd, vd, md = hip.upload(x), hip.upload(v), hip.upload(m)
    ed = hip.upload(numpy.array([numpy.nan]))
    omega2 = -1. if omega is None else omega ** 2.
    rc = _lib.lib().wdy_energy_device_noalloc(len(x), xd, vd, md, float(twopiG),
                                              float(omega2), ed, stream)
    assert rc == 0, _lib.lib().wdy_last_error()
    hip.check(hip.lib.hipStreamSynchronize(stream))
    e = float(hip.download(ed, 1)[0])
    hip.free()
    return e


@pytest.fixture
def hip(gpu_lib):
    h = _Hip()
    s = ctypes.c_void_p()
    h.check(h.lib.hipStreamCreate(ctypes.byref(s)))
    h.stream = s
    yield h
    h.free()
    h.lib.hipStreamDestroy(s)


@pytest.mark.parametrize('order', [(1 << 20, 3000), (50000, 1 << 20)])
def test_noalloc_energy_equals_energy_arrays(hip, order):
    print('noalloc energy', order, flush=True)
    for k, n in enumerate(order + (77777,)):
        x, v, m = O.sech2_disk(n, seed=40 + k, mass=('lognormal', 0.4))
        twopiG, omega = (1., None) if k % 2 == 0 else (2.5, 0.3)
        want = wendy_amd.energy(x, v, m, twopiG=twopiG, omega=omega)
        # on a stream of the caller's, then on the default stream
        got = _noalloc(hip, x, v, m, twopiG, omega, hip.stream if k != 1 else None)
        assert bits_equal(got, want), (n, got, want)


def test_noalloc_energy_equal_masses_and_ties(hip):
    n = 4096
    x, v, m = O.sech2_disk(n, seed=5)
    x[::7] = x[0]                      # a large group of exact ties
    want = wendy_amd.energy(x, v, m, omega=0.1)
    assert bits_equal(_noalloc(hip, x, v, m, 1., 0.1, hip.stream), want)


def test_noalloc_rejects_sizes_beyond_its_scratch(hip):
    n = _lib.lib().wdy_energy_device_noalloc
    p = hip.upload(numpy.zeros(4))
    assert n((1 << 20) + 1, p, p, p, 1., -1., p, None) == _lib.WDY_EINVAL
    assert b'static scratch' in _lib.lib().wdy_last_error()
    assert n(0, p, p, p, 1., -1., p, None) == _lib.WDY_EINVAL
    assert n(4, p, p, p, 0., -1., p, None) == _lib.WDY_EINVAL
