"""The reference's own approx-path tests (tests/test_approx.py,
tests/test_approx_harm.py), run against wendy_amd on the GPU with the same
inputs and bars; plus the diagnostics (energy, momentum, potential) against
the oracle and closed forms, and the reference-signature C entry point."""
import ctypes

import numpy
import pytest

import wendy_amd
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O
from oracle import ref_harness as R
from wendy_amd import _lib

pytestmark = pytest.mark.gpu


def _disk(rs, n=101):
    # tests/test_approx.py:34-44 (own RandomState instead of the global seed)
    zh = 2.
    x = numpy.arctanh(2. * rs.uniform(size=n) - 1) * zh
    v = rs.normal(size=n)
    v -= numpy.mean(v)
    m = numpy.ones_like(x) / n * (1. + 0.1 * (2. * rs.uniform(size=n) - 1))
    return x, v, m


def _energy_conserved(x, v, m, dt, nleap, nout, omega=None, **kw):
    g = wendy_amd.nbody(x, v, m, dt, approx=True, nleap=nleap, omega=omega, **kw)
    E = wendy_amd.energy(x, v, m, omega=omega)
    for _ in range(nout):
        tx, tv = next(g)[:2]
        assert abs(wendy_amd.energy(tx, tv, m, omega=omega) - E) / E < 1e-6


@pytest.mark.parametrize('masses', [[1., 1., 1.], [1., 2., 3.]])
@pytest.mark.parametrize('omega', [None, 1.1])
def test_energy_conservation_three_body(gpu_lib, masses, omega):
    """tests/test_approx.py:6-32, tests/test_approx_harm.py:5-33 (nleap=1e5,
    all 100 outputs)."""
    _energy_conserved(numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
                      numpy.array(masses), 0.05, 100000, 100, omega=omega)


@pytest.mark.parametrize('sort', ['quick', 'merge', 'tim', 'qsort', 'parallel'])
def test_energy_conservation_sech2disk(gpu_lib, sort):
    """tests/test_approx.py:34-127 (one test per sort kind)."""
    rs = numpy.random.RandomState(2)
    x, v, m = _disk(rs)
    _energy_conserved(x, v, m, 0.05, 1000, 100, sort=sort)


def test_energy_conservation_sech2disk_51(gpu_lib):
    """tests/test_approx.py:129-149."""
    x, v, m = _disk(numpy.random.RandomState(3), n=51)
    _energy_conserved(x, v, m, 0.05, 1000, 100, sort='parallel')


def test_energy_conservation_sech2disk_harmonic(gpu_lib):
    """tests/test_approx_harm.py:35-53."""
    x, v, m = _disk(numpy.random.RandomState(4))
    _energy_conserved(x, v, m, 0.05, 1000, 100, omega=1.1)


def test_momentum_conservation(gpu_lib):
    """tests/test_approx.py:151-163."""
    x = numpy.array([-1.1, 0.1, 1.3])
    v = numpy.array([3., 2., -5.])
    m = numpy.array([1., 2., 3.])
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=1000)
    p = wendy_amd.momentum(v, m)
    for _ in range(100):
        tx, tv = next(g)
        assert abs(wendy_amd.momentum(tv, m) - p) < 1e-10


def test_tracer_masses(gpu_lib):
    """tests/test_approx.py:165-185."""
    x, v, m = _disk(numpy.random.RandomState(5))
    m[101 // 2:] = 0.
    m *= 2.
    _energy_conserved(x, v, m, 0.05, 1000, 100)


def test_time(gpu_lib):
    """tests/test_approx.py:198-211."""
    x, v, m = _disk(numpy.random.RandomState(6))
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=1000, full_output=True)
    tx, tv, time_elapsed = next(g)
    assert 0. < time_elapsed < 1.


def test_samex(gpu_lib):
    """tests/test_approx.py:234-248: two sheets land on the same x at the
    first force evaluation."""
    x = numpy.array([-1.1, -2. * 0.05 / 2. / 10000, 0.3 * 0.05 / 2. / 10000, 1.3])
    v = numpy.array([3., 2., -.3, -5.])
    m = numpy.array([1., 1., 1., 1.])
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=10000)
    E = wendy_amd.energy(x, v, m)
    tx, tv = next(g)
    assert abs(wendy_amd.energy(tx, tv, m) - E) / E < 1e-6
    # and it agrees with the oracle, which breaks the tie by index
    st = O.ApproxState(x, v, m, 0.05, 10000)
    ox, ov = st.onestep()
    assert bits_equal(tx, ox) and bits_equal(tv, ov)


def test_yields_same_arrays_and_inputs_untouched(gpu_lib):
    x, v, m = _disk(numpy.random.RandomState(7))
    x0, v0 = x.copy(), v.copy()
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=10)
    a = next(g)
    b = next(g)
    assert a[0] is b[0] and a[1] is b[1]
    assert numpy.array_equal(x, x0) and numpy.array_equal(v, v0)


# ---- diagnostics ------------------------------------------------------------

@pytest.mark.parametrize('omega', [None, 0.8])
@pytest.mark.parametrize('n', [1, 3, 50, 5000])
def test_energy_matches_oracle(gpu_lib, n, omega):
    rs = numpy.random.RandomState(n)
    x = rs.normal(size=n)
    v = rs.normal(size=n)
    m = rs.uniform(0.1, 2., size=n)
    e = wendy_amd.energy(x, v, m, twopiG=1.3, omega=omega)
    ref = O.energy(x, v, m, twopiG=1.3, omega=omega)
    assert abs(e - ref) <= 1e-12 * abs(ref)
    if n <= 50:
        bf = O.brute_force_energy(x, v, m, twopiG=1.3, omega=omega)
        assert abs(e - bf) <= 1e-12 * abs(bf)


def test_energy_individual_and_potential(gpu_lib):
    """tests/test_wendy.py:81-104 closed forms, and vs the oracle."""
    x = numpy.array([-1.1, 0.1, 1.3])
    v = numpy.array([3., 2., -5.])
    m = numpy.array([1., 2., 3.])
    E = wendy_amd.energy(x, v, m, individual=True)
    assert abs(E[0] - m[0] * v[0] ** 2 / 2 - m[0] * (m[1] * abs(x[0] - x[1])
                                                     + m[2] * abs(x[0] - x[2]))) < 1e-10
    assert abs(E[1] - m[1] * v[1] ** 2 / 2 - m[1] * (m[0] * abs(x[0] - x[1])
                                                     + m[2] * abs(x[2] - x[1]))) < 1e-10
    assert abs(E[2] - m[2] * v[2] ** 2 / 2 - m[2] * (m[0] * abs(x[0] - x[2])
                                                     + m[1] * abs(x[2] - x[1]))) < 1e-10
    y = numpy.arange(-2., 2.5, 0.5)
    p = wendy_amd.potential(y, x, v, m)
    for ty, tp in zip(y, p):
        assert abs(tp - numpy.sum(m * numpy.fabs(x - ty))) < 1e-10
    rs = numpy.random.RandomState(0)
    x = rs.normal(size=700)
    v = rs.normal(size=700)
    m = rs.uniform(size=700)
    numpy.testing.assert_allclose(
        wendy_amd.energy(x, v, m, twopiG=2., individual=True, omega=0.4),
        O.energy(x, v, m, twopiG=2., individual=True, omega=0.4), rtol=1e-12)
    y = rs.normal(size=33)
    numpy.testing.assert_allclose(wendy_amd.potential(y, x, v, m, omega=0.4),
                                  O.potential(y, x, v, m, omega=0.4), rtol=1e-12)


def test_momentum_matches_oracle(gpu_lib):
    rs = numpy.random.RandomState(1)
    v = rs.normal(size=12345)
    m = rs.uniform(size=12345)
    assert abs(wendy_amd.momentum(v, m) - O.momentum(v, m)) < 1e-11


@pytest.mark.parametrize('path', ['small', 'large'])
@pytest.mark.parametrize('mass', ['equal', 'lognormal'])
def test_sim_energy_of_live_state(gpu_lib, path, mass):
    x, v, m = O.sech2_disk(3000, seed=1, mass='equal' if mass == 'equal'
                           else ('lognormal', 0.5))
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.09, path=path)
    xo = numpy.empty(3000)
    vo = numpy.empty(3000)
    for _ in range(2):
        sim.advance(0.002, 5)
        sim.read(xo, vo)
        e = sim.energy(omega=0.3)
        ref = O.energy(xo, vo, m, omega=0.3)
        assert abs(e - ref) <= 1e-12 * abs(ref)
    sim.close()


# ---- reference-signature C entry point --------------------------------------

def _compat_call(lib, xi, x, v, m, a, cum, totmass, dt, nleap, omega2):
    err = ctypes.c_int(0)
    te = ctypes.c_double(0.)
    t0 = ctypes.c_double(0.)
    P = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)
    lib._wendy_nbody_approx_onestep(
        ctypes.c_int(len(x)), P(xi), P(x), P(v), P(m), P(a),
        ctypes.c_double(totmass), ctypes.c_double(dt), ctypes.c_int(nleap),
        ctypes.byref(t0), ctypes.c_double(omega2), None, ctypes.c_int(4),
        ctypes.byref(err), ctypes.byref(te), P(cum))
    return err.value, te.value


@pytest.mark.parametrize('n', [101, 3000])
def test_compat_entry_point_matches_reference_state(gpu_lib, n):
    """The reference-signature symbol, driven exactly like the reference's
    ctypes call (wendy/wendy.py:425-433), reproduces the oracle's xi, x, v,
    a and cumulmass over consecutive calls."""
    lib = ctypes.CDLL(_lib.LIB_PATH)
    lib._wendy_nbody_approx_onestep.restype = None
    x0, v0, m0 = O.sech2_disk(n, seed=4)
    dt, nleap, omega2 = 0.02, 5, 0.2 ** 2.
    x = x0.copy()
    v = v0.copy()
    m = m0.copy()
    a = numpy.zeros(n)
    cum = numpy.zeros(n)
    xi = numpy.zeros(n, dtype=R.XI_DTYPE)
    xi['idx'] = numpy.arange(n, dtype=numpy.int32)
    xi['val'] = x0
    st = O.ApproxState(x0, v0, m0, dt * nleap, nleap, omega=0.2)
    for _ in range(3):
        err, te = _compat_call(lib, xi, x, v, m, a, cum, numpy.sum(m), dt,
                               nleap, omega2)
        assert err == 0 and te > 0
        ox, ov = st.onestep()
        exact = numpy.frexp(m0[0])[0] == 0.5
        if exact:
            assert bits_equal(x, ox) and bits_equal(v, ov)
        else:
            assert rel_inf(x, ox) <= 1e-10 and rel_inf(v, ov) <= 1e-10
        assert numpy.array_equal(xi['idx'], st.idx)
        assert rel_inf(xi['val'], st.val) <= 1e-10
    # a and cumulmass hold the last force evaluation: recompute it from the
    # sorted positions before the final half drift
    h = dt
    val_force = (xi['val'] - (h / 2.) * v[xi['idx']])
    ms = m[xi['idx']]
    a_sorted = O.force_sorted(val_force, ms, numpy.sum(m), omega2)
    assert rel_inf(a[xi['idx']], a_sorted) <= 1e-9
    assert rel_inf(cum, O.exclusive_mass_prefix(ms)) <= 1e-12


@pytest.mark.parametrize('omega', [None, 1.1])
def test_approx_against_exact(gpu_lib, omega):
    """tests/test_approx.py:213-232 and tests/test_approx_harm.py's
    counterpart: the GPU leapfrog (nleap = 2000) and the exact event-driven
    solver (CPU) agree to 1e-5 over 100 outputs."""
    x, v, m = _disk(numpy.random.RandomState(11))
    g = wendy_amd.nbody(x, v, m, 0.05, approx=True, nleap=2000, omega=omega)
    ge = wendy_amd.nbody(x, v, m, 0.05, omega=omega)
    for _ in range(100):
        tx, tv = next(g)
        txe, tve = next(ge)
        assert numpy.all(numpy.fabs(tx - txe) < 1e-5)
        assert numpy.all(numpy.fabs(tv - tve) < 1e-5)


@pytest.mark.skipif(not R.available(), reason='oracle/_ref not built')
@pytest.mark.parametrize('mass,omega', [('equal', 0.1), ('lognormal', None)])
def test_trajectory_and_energy_drift_against_reference(gpu_lib, mass, omega):
    """The acceptance bars against the compiled reference itself: after 100
    leapfrog steps positions and velocities agree to 1e-10 relative, and the
    relative energy drift is within 10% of the reference's own drift (plus a
    1e-14 floor)."""
    n = 1 << 14
    x, v, m = O.sech2_disk(n, seed=17, mass='equal' if mass == 'equal'
                           else ('lognormal', 0.5))
    dt, nleap = 0.05, 10
    ref = R.RefApprox(x, v, m, dt, nleap, omega=omega, sort='quick')
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1. if omega is None else omega ** 2.,
                        path='large')
    xo = numpy.empty(n)
    vo = numpy.empty(n)
    E0 = O.energy(x, v, m, omega=omega)
    try:
        for _ in range(10):
            rx, rv, _ = ref.onestep()
            sim.advance(dt / nleap, nleap)
        sim.read(xo, vo)
    finally:
        sim.close()
    assert rel_inf(xo, rx) <= 1e-10 and rel_inf(vo, rv) <= 1e-10
    d_ref = abs((O.energy(rx, rv, m, omega=omega) - E0) / E0)
    d_gpu = abs((O.energy(xo, vo, m, omega=omega) - E0) / E0)
    assert abs(d_gpu - d_ref) <= 0.1 * d_ref + 1e-14, (d_gpu, d_ref)
