"""The exact event-driven solver (wendy_amd.nbody(approx=False), CPU code in
libwendy_amd.so) against the compiled reference's exact solver and against
the reference's own tests/test_wendy.py and tests/test_approx_harm.py
checks.  Runs without a GPU."""
import math
import warnings

import numpy
import pytest

import wendy_amd
from conftest import bits_equal
from oracle import nbody_oracle as O
from oracle import ref_harness as R


def _disk101(rs):
    # tests/test_wendy.py:34-45 of the reference (own RandomState)
    N, totmass, sigma = 101, 1., 1.
    zh = 2. * sigma ** 2. / totmass
    x = numpy.arctanh(2. * rs.uniform(size=N) - 1) * zh
    v = rs.normal(size=N) * sigma
    v -= numpy.mean(v)
    m = numpy.ones_like(x) / N * (1. + 0.1 * (2. * rs.uniform(size=N) - 1))
    return x, v, m


def _three(masses):
    return (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
            numpy.array(masses))


CASES = [('three_equal', None), ('three_unequal', None), ('disk101', None),
         ('three_unequal', 1.1), ('disk101', 1.1), ('disk1000', None)]


def _case(name):
    if name == 'three_equal':
        return _three([1., 1., 1.])
    if name == 'three_unequal':
        return _three([1., 2., 3.])
    if name == 'disk101':
        return _disk101(numpy.random.RandomState(2))
    x, v, m = O.sech2_disk(1000, seed=9, mass=('lognormal', 0.3))
    return x, v, m


@pytest.mark.skipif(not R.available(), reason='oracle/_ref not built here')
@pytest.mark.parametrize('name,omega', CASES)
def test_exact_matches_reference_solver(name, omega):
    """Same crossings, same counts, same trajectories as the compiled
    reference, output after output (bit for bit: the arithmetic is the
    reference's and no two crossings are simultaneous here)."""
    x, v, m = _case(name)
    dt = 0.05
    ref = R.RefExact(x, v, m, dt, omega=omega)
    g = wendy_amd.nbody(x, v, m, dt, omega=omega, full_output=True)
    for k in range(40):
        rx, rv, rn, err = ref.onestep()
        gx, gv, gn, _ = next(g)
        assert err == 0
        assert gn == rn, (k, gn, rn)
        assert bits_equal(gx, rx) and bits_equal(gv, rv), k


@pytest.mark.parametrize('masses', [[1., 1., 1.], [1., 2., 3.]])
def test_energy_conservation_three(masses):
    """tests/test_wendy.py:6-32."""
    x, v, m = _three(masses)
    g = wendy_amd.nbody(x, v, m, 0.05)
    E = O.energy(x, v, m)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(O.energy(tx, tv, m) - E) < 1e-10


def test_energy_conservation_disk():
    """tests/test_wendy.py:34-51."""
    x, v, m = _disk101(numpy.random.RandomState(2))
    g = wendy_amd.nbody(x, v, m, 0.05)
    E = O.energy(x, v, m)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(O.energy(tx, tv, m) - E) < 1e-10


def test_energy_conservation_harmonic():
    """tests/test_approx_harm.py's exact-solver counterpart: with omega the
    energy includes the external potential."""
    x, v, m = _disk101(numpy.random.RandomState(5))
    g = wendy_amd.nbody(x, v, m, 0.05, omega=1.1)
    E = O.energy(x, v, m, omega=1.1)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(O.energy(tx, tv, m, omega=1.1) - E) < 1e-10


def test_momentum_conservation():
    """tests/test_wendy.py:67-79."""
    x, v, m = _three([1., 2., 3.])
    g = wendy_amd.nbody(x, v, m, 0.05)
    p = O.momentum(v, m)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(O.momentum(tv, m) - p) < 1e-10


def test_count_ncoll():
    """tests/test_wendy.py:106-123: two sheets from rest at +-1 first cross
    at t = sqrt(2), then every 2 sqrt(2)."""
    g = wendy_amd.nbody(numpy.array([-1., 1.]), numpy.array([0., 0.]),
                        numpy.array([1., 1.]), 1, full_output=True)
    counts = [next(g)[2] for _ in range(5)]
    assert counts == [0, 1, 1, 1, 2]


def test_free_fall_closed_form():
    """Before the first crossing the two sheets fall on x = -+(1 - t^2/2)."""
    g = wendy_amd.nbody(numpy.array([-1., 1.]), numpy.array([0., 0.]),
                        numpy.array([1., 1.]), 0.25)
    for k in range(1, 5):
        tx, tv = next(g)
        t = 0.25 * k
        assert abs(tx[0] + (1. - t * t / 2.)) < 1e-15 and abs(tv[0] - t) < 1e-15


def test_maxcoll_error_and_warning():
    """tests/test_wendy.py:125-148."""
    x, v, m = numpy.array([-1., 1.]), numpy.array([0., 0.]), numpy.array([1., 1.])
    g = wendy_amd.nbody(x, v, m, 2, maxcoll=0, full_output=True)
    with pytest.raises(RuntimeError) as e:
        next(g)
    assert str(e.value) == 'Maximum number of collisions per time step exceeded'
    g = wendy_amd.nbody(x, v, m, 2, maxcoll=0, full_output=True, warn_maxcoll=True)
    with pytest.warns(RuntimeWarning) as rec:
        next(g)
    assert len(rec) == 1
    assert rec[0].message.args[0] == 'Maximum number of collisions per time step exceeded'


def test_input_checks():
    """tests/test_wendy.py:150-165 and the other checks of
    wendy/wendy.py:158-165, with the reference's messages."""
    x, v = numpy.array([-1., 1.]), numpy.array([0., 0.])
    g = wendy_amd.nbody(x, v, numpy.array([1., 0.]), 2)
    with pytest.raises(ValueError) as e:
        next(g)
    assert str(e.value) == ('Tiny masses m much smaller than the median mass are '
                            'not supported, please remove these from the inputs')
    g = wendy_amd.nbody(x, v, numpy.array([1., 1.]), 2, omega=1.)
    with pytest.raises(ValueError):
        next(g)
    g = wendy_amd.nbody(x, v, numpy.array([1., 1.]), 2,
                        ext_force=wendy_amd.forces.Harmonic(1.))
    with pytest.raises(ValueError):
        next(g)
