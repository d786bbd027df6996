"""The reference's loader bound to this library, driving the approximate step
on the GPU exactly as wendy/wendy.py:363-433 drives it (argtypes of
wendy/wendy.py:76-93, xi records, m * twopiG, pairwise totmass, the omega2
sentinel, and an external force as the reference's plain-ctypes callback,
wendy/wendy.py:406-418), against the compiled reference itself.

Bars: bit for bit (positions, velocities, the sorted xi records, the clock
t0) with equal power-of-two masses, where the reference's running mass sum
is exact; 1e-10 relative (infinity norm) otherwise.  The force functions are
numpy expressions evaluated on the host in both, so they contribute the same
bits."""
import ctypes

import numpy
import pytest

from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O
from oracle import ref_harness as R
from wendy_amd import _lib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not R.available(),
                                 reason='oracle/_ref not built')]


@pytest.fixture(scope='module')
def ours(gpu_lib):
    return R.bind(ctypes.CDLL(_lib.LIB_PATH))


def _compare(ours, x, v, m, dt, nleap, nout, omega=None, ext_force=None,
             t0=0., sort='parallel'):
    exact = bool(numpy.all(m == m[0]) and numpy.frexp(m[0])[0] == 0.5)
    mine = R.RefApprox(x, v, m, dt, nleap, omega=omega, sort=sort,
                       ext_force=ext_force, t0=t0, lib=ours)
    ref = R.RefApprox(x, v, m, dt, nleap, omega=omega, sort=sort,
                      ext_force=ext_force, t0=t0)
    for k in range(nout):
        mx, mv, mt = mine.onestep()
        rx, rv, _ = ref.onestep()
        assert mine.err.value == 0, k
        assert mt > 0.
        if exact:
            assert bits_equal(mx, rx) and bits_equal(mv, rv), k
            assert bits_equal(mine.xi['val'], ref.xi['val']), k
        else:
            assert rel_inf(mx, rx) <= 1e-10 and rel_inf(mv, rv) <= 1e-10, k
            assert rel_inf(mine.xi['val'], ref.xi['val']) <= 1e-10, k
        # tie-free inputs: the sorted order is unique
        assert numpy.array_equal(mine.xi['idx'], ref.xi['idx']), k
        assert bits_equal(mine.t0.value, ref.t0.value), k
    return mine, ref


@pytest.mark.parametrize('n,omega', [(4096, None), (4096, 0.3), (1000, None)])
def test_approx_step_without_force(ours, n, omega):
    print('loader approx', n, omega, flush=True)
    x, v, m = O.sech2_disk(n, seed=21)
    _compare(ours, x, v, m, 0.02, 8, 4, omega=omega)


@pytest.mark.parametrize('n,mass', [(3, 'three'), (8, 'lognormal'),
                                    (256, 'equal'), (101, 'lognormal'),
                                    (2000, 'equal')])
def test_approx_step_with_host_force(ours, n, mass):
    """A time-dependent Python force through the reference's callback: above
    10 sheets one call per step on the de-sorted positions, at or below one
    call per sheet in sorted order; t0 advances by dt per force
    evaluation."""
    print('loader ext', n, mass, flush=True)
    if mass == 'three':
        x, v, m = (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
                   numpy.array([1., 2., 3.]))
    else:
        x, v, m = O.sech2_disk(n, seed=n, mass='equal' if mass == 'equal'
                               else ('lognormal', 0.3))

    def force(y, t):
        return -1.21 * y + 0.25 * numpy.sin(3. * t)

    mine, ref = _compare(ours, x, v, m, 0.02, 6, 4, ext_force=force, t0=0.7)
    assert mine.t0.value != 0.7          # the clock moved, as the reference's
    # a holds the last force evaluation (E + F) in sheet order
    assert rel_inf(mine.a, ref.a) <= 1e-12


def test_compat_step_rejects_nothing_the_reference_accepts(ours):
    """Every sort name of the reference's table runs (all map to the device
    sort)."""
    x, v, m = O.sech2_disk(512, seed=3)
    for sort in ('quick', 'merge', 'tim', 'qsort', 'parallel'):
        _compare(ours, x, v, m, 0.02, 4, 1, sort=sort)
