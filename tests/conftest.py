import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs a GPU and the built libwendy_amd.so')


def load_golden(name):
    with numpy.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    d['omega'] = None if numpy.isnan(d['omega']) else float(d['omega'])
    d['dt'] = float(d['dt'])
    d['nleap'] = int(d['nleap'])
    d['twopiG'] = float(d['twopiG'])
    d['sort'] = str(d['sort'])
    return d


GOLDEN_CASES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith('.npz'))


def rel_inf(a, b):
    """max|a-b| / max|b| (infinity-norm relative difference)."""
    scale = numpy.max(numpy.abs(b))
    return float(numpy.max(numpy.abs(a - b)) / (scale if scale > 0 else 1.))


def bits_equal(a, b):
    """Bitwise equality with -0.0 and +0.0 identified (the device stores
    positions with the two zeros folded together)."""
    a = numpy.asarray(a, dtype=numpy.float64) + 0.0
    b = numpy.asarray(b, dtype=numpy.float64) + 0.0
    return numpy.array_equal(a.view(numpy.uint64), b.view(numpy.uint64))


@pytest.fixture(scope='session')
def gpu_lib():
    """The built library on a GPU box; fails loudly (never skips) if it is
    missing or no GPU is visible."""
    from wendy_amd import _lib
    lib = _lib.lib()
    import ctypes
    n = ctypes.c_int(0)
    rc = lib.wdy_device_count(ctypes.byref(n))
    assert rc == 0 and n.value >= 1, 'no GPU visible to libwendy_amd.so'
    return lib
