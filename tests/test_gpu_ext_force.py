"""External forces on the device (wendy_amd.forces; the reference's
ext_force, wendy/wendy.py:389-420, wendy/wendy.c:362-370) against the oracle,
whose external-force step and clock equal the compiled reference bit for bit
(tests/test_oracle.py).  Bars: bitwise for the harmonic form with
power-of-two equal masses (E = -p0 x and F are exact mirrors), else
||.||_inf relative 1e-10 (the device's tanh and exp may differ from numpy's by
an ulp).  Arbitrary Python callables are called on the host before every
step (the reference's own callback convention) and are checked against the
same oracle.  Plus the reference's own tests/test_approx_external.py: its
long runs with the device form of its harmonic lambda, and its lambda and
force class themselves on the host-callback path over fewer outputs."""
import math

import numpy
import pytest

import wendy_amd
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O
from wendy_amd import _lib, forces

pytestmark = pytest.mark.gpu

TOL = 1e-10


def menu():
    return {'harmonic': forces.Harmonic(1.1), 'isothermal': forces.Isothermal(-0.7, 0.5),
            'pulse': forces.Pulse(0.3, 0.012, 0.005)}


def run(x, v, m, force, dt, nleap, nout, path=None, warm=True, t0=0.003):
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path=path, warm=warm)
    sim.set_ext_force(force, t0)
    st = O.ApproxState(x, v, m, dt, nleap, ext_force=force, t0=t0)
    xo = numpy.empty(len(x))
    vo = numpy.empty(len(x))
    exact = (isinstance(force, forces.Harmonic) and numpy.all(m == m[0])
             and numpy.frexp(m[0])[0] == 0.5)
    try:
        for k in range(nout):
            sim.advance(dt / nleap, nleap)
            sim.read(xo, vo)
            wx, wv = st.onestep()
            if exact:
                assert bits_equal(xo, wx) and bits_equal(vo, wv), k
            else:
                assert rel_inf(xo, wx) <= TOL, (k, rel_inf(xo, wx))
                assert rel_inf(vo, wv) <= TOL, (k, rel_inf(vo, wv))
            assert sim.time() == st.t     # the clock, bit for bit
        return sim.path_stats()
    finally:
        sim.close()


@pytest.mark.parametrize('name', ['harmonic', 'isothermal', 'pulse'])
@pytest.mark.parametrize('n,mass', [(3, 'three'), (101, 'lognormal'), (4096, 'equal')])
def test_small_path(gpu_lib, name, n, mass):
    print('ext small', name, n, flush=True)
    if mass == 'three':
        x, v, m = (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
                   numpy.array([1., 2., 3.]))
    else:
        x, v, m = O.sech2_disk(n, seed=n, mass='equal' if mass == 'equal'
                               else ('lognormal', 0.1))
    run(x, v, m, menu()[name], 0.01, 20, 4)


@pytest.mark.parametrize('name', ['harmonic', 'isothermal', 'pulse'])
@pytest.mark.parametrize('warm', [True, False])
def test_large_path(gpu_lib, name, warm):
    print('ext large', name, warm, flush=True)
    x, v, m = O.sech2_disk(40000, seed=3, mass=('lognormal', 0.5))
    warm_steps, cold, redo = run(x, v, m, menu()[name], 0.01, 5, 4, path='large',
                                 warm=warm)
    assert warm_steps == (15 if warm else 0)


def test_large_path_equal_masses_bitwise(gpu_lib):
    x, v, m = O.sech2_disk(1 << 15, seed=8)
    run(x, v, m, forces.Harmonic(0.3), 0.01, 5, 4, path='large')


def test_pulse_momentum_on_device(gpu_lib):
    """Total momentum changes by M h sum_k g(t_k), t_k = t0 + k h."""
    f = forces.Pulse(0.5, 0.07, 0.03)
    x, v, m = O.sech2_disk(30000, seed=6)
    h, nleap, t0 = 0.01, 5, 0.02
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1.)
    sim.set_ext_force(f, t0)
    try:
        for _ in range(3):
            sim.advance(h, nleap)
        vo = numpy.empty(len(x))
        sim.read(None, vo)
    finally:
        sim.close()
    ts, t = [], t0
    for _ in range(15):
        ts.append(t)
        t = t + h
    dp = math.fsum(m * vo) - math.fsum(m * v)
    assert abs(dp - math.fsum(m) * h * math.fsum(f(0., tk) for tk in ts)) <= 1e-12


def test_nbody_accepts_menu_and_callables(gpu_lib):
    """A device form and an arbitrary Python callable (called on the host
    before every step, as the reference calls its ext_force) both match the
    oracle, whose callback path equals the compiled reference's bit for
    bit; with power-of-two equal masses the callable's run is bitwise."""
    x, v, m = O.sech2_disk(64, seed=2)
    f = lambda y, t: -1.3 * y + 0.2 * numpy.cos(t)
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=10, t0=0.4,
                        ext_force=f)
    st = O.ApproxState(x, v, m, 0.05, 10, ext_force=f, t0=0.4)
    for _ in range(3):
        gx, gv = next(g)
        wx, wv = st.onestep()
        assert bits_equal(gx, wx) and bits_equal(gv, wv)
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=10, t0=0.5,
                        ext_force=forces.Pulse(1., 0.5, 0.1))
    st = O.ApproxState(x, v, m, 0.05, 10, ext_force=forces.Pulse(1., 0.5, 0.1), t0=0.5)
    for _ in range(3):
        gx, gv = next(g)
        wx, wv = st.onestep()
        assert rel_inf(gx, wx) <= TOL and rel_inf(gv, wv) <= TOL


@pytest.mark.parametrize('n,mass', [(3, 'three'), (101, 'lognormal'),
                                    (20000, 'lognormal')])
def test_host_callable_against_oracle(gpu_lib, n, mass):
    """Host-evaluated forces on the large path, sheet by sheet (n <= 10)
    and on the whole array, against the oracle to 1e-10 (the mass prefix
    is reassociated); the clock bit for bit."""
    print('ext host', n, mass, flush=True)
    if mass == 'three':
        x, v, m = (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
                   numpy.array([1., 2., 3.]))
    else:
        x, v, m = O.sech2_disk(n, seed=n, mass=('lognormal', 0.2))

    class Tide:  # a callable object, as in the reference's tests
        def __call__(self, y, t):
            return 0.4 * numpy.tanh(y) * numpy.exp(-t)

    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large', warm=False)
    f = Tide()
    sim.set_ext_force(f, 0.1)
    st = O.ApproxState(x, v, m, 0.01 * 7, 7, ext_force=f, t0=0.1)
    xo = numpy.empty(len(x))
    vo = numpy.empty(len(x))
    try:
        for k in range(3):
            sim.advance(0.01, 7)
            sim.read(xo, vo)
            wx, wv = st.onestep()
            assert rel_inf(xo, wx) <= TOL and rel_inf(vo, wv) <= TOL, k
            assert sim.time() == st.t
    finally:
        sim.close()


def test_host_callable_refused_on_small_path(gpu_lib):
    x, v, m = O.sech2_disk(64, seed=2)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='small')
    try:
        with pytest.raises(_lib.WendyError):
            sim.set_ext_force(lambda y, t: -y, 0.)
    finally:
        sim.close()


# ---- the reference's tests/test_approx_external.py ---------------------------
def _disk101(rs):
    N, totmass, sigma = 101, 1., 1.
    zh = 2. * sigma ** 2. / totmass
    x = numpy.arctanh(2. * rs.uniform(size=N) - 1) * zh
    v = rs.normal(size=N) * sigma
    v -= numpy.mean(v)
    m = numpy.ones_like(x) / N * (1. + 0.1 * (2. * rs.uniform(size=N) - 1))
    return x, v, m


@pytest.mark.parametrize('masses,nleap,tol', [([1., 1., 1.], 100000, 1e-6),
                                              ([1., 2., 3.], 10000, 1e-5)])
def test_reference_energy_conservation_three(gpu_lib, masses, nleap, tol):
    """tests/test_approx_external.py:5-43."""
    x = numpy.array([-1.1, 0.1, 1.3])
    v = numpy.array([3., 2., -5.])
    m = numpy.array(masses)
    omega = 1.1
    g = wendy_amd.nbody(x, v, m, 0.05, ext_force=forces.Harmonic(omega), approx=True,
                        nleap=nleap)
    E = wendy_amd.energy(x, v, m, omega=omega)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(wendy_amd.energy(tx, tv, m, omega=omega) - E) / E < tol


def test_reference_energy_conservation_sech2disk(gpu_lib):
    """tests/test_approx_external.py:45-97."""
    x, v, m = _disk101(numpy.random.RandomState(2))
    omega = 1.1
    g = wendy_amd.nbody(x, v, m, 0.05, ext_force=forces.Harmonic(omega), approx=True,
                        nleap=1000)
    E = wendy_amd.energy(x, v, m, omega=omega)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(wendy_amd.energy(tx, tv, m, omega=omega) - E) / E < 1e-6


def test_reference_external_against_exact(gpu_lib):
    """tests/test_approx_external.py:99-113: the approximate integration with
    the harmonic force given as ext_force agrees with the exact solver run
    with omega to 1e-5 over 100 outputs (nleap = 2000)."""
    x, v, m = _disk101(numpy.random.RandomState(3))
    omega = 1.1
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=2000,
                        ext_force=forces.Harmonic(omega))
    ge = wendy_amd.nbody(x, v, m, 0.05, omega=omega)
    for _ in range(100):
        tx, tv = next(g)
        txe, tve = next(ge)
        assert numpy.all(numpy.fabs(tx - txe) < 1e-5)
        assert numpy.all(numpy.fabs(tv - tve) < 1e-5)


@pytest.mark.parametrize('wclass', [False, True])
def test_reference_external_lambda_energy_sech2disk(gpu_lib, wclass):
    """tests/test_approx_external.py:45-97 with the reference's own lambda
    and force class (host-evaluated: one device-host round trip per step),
    over 10 of its 100 outputs."""
    x, v, m = _disk101(numpy.random.RandomState(2))
    omega = 1.1

    class Eforce(object):
        def __init__(self, omega):
            self._omega2 = omega ** 2.

        def __call__(self, x, t):
            return -self._omega2 * x
    f = Eforce(omega) if wclass else (lambda x, t: -omega ** 2. * x)
    g = wendy_amd.nbody(x, v, m, 0.05, ext_force=f, approx=True, nleap=1000)
    E = wendy_amd.energy(x, v, m, omega=omega)
    for _ in range(10):
        tx, tv = next(g)
        assert abs(wendy_amd.energy(tx, tv, m, omega=omega) - E) / E < 1e-6
