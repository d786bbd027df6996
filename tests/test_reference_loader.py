"""The reference's own loader can bind this library: all five entry points it
looks up at import (wendy/wendy.py:38-101) are exported with the reference's
argument and return types, and the four host-side ones (the exact solver's
steps and its two crossing-time roots) return the compiled reference's
results bit for bit, call after call, when driven exactly as the reference's
Python drives them (set-up of wendy/wendy.py:243-279, loop of :176-205).
Runs without a GPU: the approximate step, which needs one, is exercised
through the same binding in tests/test_gpu_reference_loader.py."""
import ctypes

import numpy
import pytest

from conftest import bits_equal
from oracle import nbody_oracle as O
from oracle import ref_harness as R
from wendy_amd import _lib

pytestmark = pytest.mark.skipif(not R.available(),
                                reason='oracle/_ref not built here')


@pytest.fixture(scope='module')
def ours():
    # a fresh handle, so that the argtypes set here do not leak into the
    # package's own binding
    return R.bind(ctypes.CDLL(_lib.LIB_PATH))


def test_loader_binds_every_symbol(ours):
    for name in ('_wendy_nbody_onestep', '_wendy_nbody_harm_onestep',
                 '_wendy_nbody_approx_onestep', '_solve_coll_quad',
                 '_solve_coll_harm'):
        assert getattr(ours, name).argtypes is not None, name


def test_crossing_roots_match_reference(ours):
    """wendy/wendy.c:33-52 on approaching pairs, harmonic and free."""
    ref = R.load()
    rs = numpy.random.RandomState(5)
    for _ in range(2000):
        c0 = -abs(rs.normal())          # lower sheet below: x_l - x_u < 0
        c1 = rs.normal() * 3.
        c2 = rs.normal()
        assert bits_equal(ours._solve_coll_quad(c0, c1, c2 / 2.),
                          ref._solve_coll_quad(c0, c1, c2 / 2.))
        w = abs(rs.normal()) + 0.1
        assert bits_equal(ours._solve_coll_harm(c0, c1, c2, w),
                          ref._solve_coll_harm(c0, c1, c2, w))


def _three(masses):
    return (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
            numpy.array(masses))


def _disk(n, seed, omega):
    rs = numpy.random.RandomState(seed)
    x = numpy.arctanh(2. * rs.uniform(size=n) - 1) * 2.
    v = rs.normal(size=n)
    v -= numpy.mean(v)
    m = numpy.ones(n) / n * (1. + 0.1 * (2. * rs.uniform(size=n) - 1))
    return x, v, m


CASES = [('three_equal', None), ('three_unequal', 1.1), ('disk101', None),
         ('disk101', 0.9), ('disk500', None)]


def _case(name):
    if name == 'three_equal':
        return _three([1., 1., 1.])
    if name == 'three_unequal':
        return _three([1., 2., 3.])
    if name == 'disk101':
        return _disk(101, 2, None)
    return O.sech2_disk(500, seed=4, mass=('lognormal', 0.3))


@pytest.mark.parametrize('name,omega', CASES)
def test_exact_steps_match_reference(ours, name, omega):
    """The whole caller-held state after every call -- positions,
    velocities, forces (or equilibrium offsets), rank order, crossing times,
    the earliest crossing and the crossing count -- equals the reference's."""
    x, v, m = _case(name)
    dt = 0.05
    mine = R.RefExact(x, v, m, dt, omega=omega, lib=ours)
    ref = R.RefExact(x, v, m, dt, omega=omega)
    assert bits_equal(mine.tcoll, ref.tcoll)
    total = 0
    for k in range(40):
        mx, mv, mn, merr = mine.onestep()
        rx, rv, rn, rerr = ref.onestep()
        assert (merr, mn) == (rerr, rn), k
        assert bits_equal(mx, rx) and bits_equal(mv, rv), k
        assert bits_equal(mine.a, ref.a), k
        assert numpy.array_equal(mine.sindx, ref.sindx), k
        assert bits_equal(mine.tcoll, ref.tcoll), k
        assert mine.cindx.value == ref.cindx.value, k
        assert bits_equal(mine.next_tcoll.value, ref.next_tcoll.value), k
        total = mn
    assert total > 0


def test_maxcoll_sets_err_like_reference(ours):
    """err = -2 once maxcoll crossings were processed in one call; the rest
    of that interval's crossings are skipped in both."""
    x, v, m = _disk(101, 2, None)
    mine = R.RefExact(x, v, m, 0.5, maxcoll=7, lib=ours)
    ref = R.RefExact(x, v, m, 0.5, maxcoll=7)
    mx, mv, mn, merr = mine.onestep()
    rx, rv, rn, rerr = ref.onestep()
    assert merr == rerr == -2 and mn == rn == 7
    assert bits_equal(mx, rx) and bits_equal(mv, rv)
    assert bits_equal(mine.tcoll, ref.tcoll)


def test_harm_step_refuses_nonpositive_omega(ours):
    x, v, m = _three([1., 2., 3.])
    st = R.RefExact(x, v, m, 0.05, omega=1.1, lib=ours)
    err = ctypes.c_int(0)
    nc = ctypes.c_int(0)
    te = ctypes.c_double(0.)
    x0 = st.x.copy()
    ours._wendy_nbody_harm_onestep(3, st.x, st.v, st.a, st.m, st.sindx,
                                   ctypes.byref(st.cindx),
                                   ctypes.byref(st.next_tcoll), st.tcoll, 0.05,
                                   100, ctypes.byref(err), ctypes.byref(nc),
                                   ctypes.byref(te), 0.)
    assert err.value == _lib.WDY_EINVAL
    assert numpy.array_equal(st.x, x0)
