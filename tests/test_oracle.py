"""Pins the CPU oracle (oracle/nbody_oracle.py) before anything is checked
against it: golden fixtures written from the compiled reference, the compiled
reference itself where oracle/_ref is built, closed forms and brute force."""
import math

import numpy
import pytest

from conftest import GOLDEN_CASES, bits_equal, load_golden
from oracle import nbody_oracle as O
from oracle import ref_harness as R


@pytest.mark.parametrize('case', GOLDEN_CASES)
def test_oracle_matches_golden_bitwise(case):
    """Fixtures come from the reference C (tests/golden/make_golden.py); the
    inputs are tie-free, so the restatement must match bit for bit."""
    g = load_golden(case)
    st = O.ApproxState(g['x0'], g['v0'], g['m'], g['dt'], g['nleap'],
                       omega=g['omega'], twopiG=g['twopiG'])
    outs = list(g['outputs'])
    for k in range(1, max(outs) + 1):
        x, v = st.onestep()
        if k in outs:
            j = outs.index(k)
            assert bits_equal(x, g['x_out'][j]), (case, k)
            assert bits_equal(v, g['v_out'][j]), (case, k)


@pytest.mark.skipif(not R.available(), reason='oracle/_ref not built here')
@pytest.mark.parametrize('sort', ['quick', 'merge', 'tim', 'qsort', 'parallel'])
@pytest.mark.parametrize('omega', [None, 0.7])
def test_oracle_matches_reference_library(sort, omega):
    x, v, m = O.sech2_disk(257, seed=5, mass=('lognormal', 0.4))
    ref = R.RefApprox(x, v, m, 0.02, 7, omega=omega, sort=sort,
                      variant='_small')
    st = O.ApproxState(x, v, m, 0.02, 7, omega=omega)
    for _ in range(5):
        rx, rv, _ = ref.onestep()
        ox, ov = st.onestep()
        assert bits_equal(rx, ox) and bits_equal(rv, ov)


def test_canonical_order_brute_force():
    rs = numpy.random.RandomState(1)
    for _ in range(50):
        n = rs.randint(1, 12)
        val = rs.choice([-1.5, -0.0, 0.0, 0.25, 2.0, 7.0], size=n)
        idx = rs.permutation(n).astype(numpy.int32)
        order = O.canonical_order(val, idx)
        # brute force: rank = #{j: (val_j, idx_j) < (val_i, idx_i)}
        for pos, i in enumerate(order):
            r = sum(1 for j in range(n)
                    if val[j] < val[i] or (val[j] == val[i] and idx[j] < idx[i]))
            assert r == pos


def test_presorted_order_equals_canonical_order():
    rs = numpy.random.RandomState(2)
    for n in (1, 2, 7, 500, 20000):
        idx = rs.permutation(n).astype(numpy.int32)
        val = rs.normal(size=n)
        near = numpy.sort(val) + 1e-3 * rs.normal(size=n)   # a few swaps
        ties = rs.choice([-1.0, -0.0, 0.0, 3.0], size=n)
        for a in (val, near, ties):
            assert numpy.array_equal(O.canonical_order_presorted(a, idx),
                                     O.canonical_order(a, idx))


def test_presorted_order_gives_the_same_trajectory():
    x, v, m = O.sech2_disk(3000, seed=5, mass=('lognormal', 0.5))
    x[7] = x[8]                      # an exact tie at the start
    a = O.ApproxState(x, v, m, 0.01, 4, omega=0.2)
    b = O.ApproxState(x, v, m, 0.01, 4, omega=0.2, presorted_order=True)
    for _ in range(3):
        ax, av = a.onestep()
        bx, bv = b.onestep()
        assert bits_equal(ax, bx) and bits_equal(av, bv)


@pytest.mark.parametrize('omega', [None, 1.3])
def test_force_matches_brute_force_pair_sum(omega):
    rs = numpy.random.RandomState(4)
    n = 61
    x = rs.normal(size=n)
    x[10] = x[20]  # an exact tie, split by index
    m = rs.uniform(0., 2., size=n)
    idx = numpy.arange(n, dtype=numpy.int32)
    order = O.canonical_order(x, idx)
    omega2 = -1. if omega is None else omega ** 2.
    a_sorted = O.force_sorted(x[order], m[order], numpy.sum(m), omega2)
    a = numpy.empty(n)
    a[order] = a_sorted
    bf = O.brute_force_accel(x, m, omega=omega)
    numpy.testing.assert_allclose(a, bf, rtol=0, atol=1e-13)


def test_free_fall_parabola():
    """Two sheets at +-1 at rest: constant force -+m until they cross at
    t = sqrt(2) (tests/test_wendy.py:106-123 uses the same setup).  DKD
    leapfrog is exact for constant acceleration: x = +-(1 - t^2/2)."""
    x = numpy.array([-1., 1.])
    v = numpy.zeros(2)
    m = numpy.ones(2)
    dt, nleap = 0.125, 4
    st = O.ApproxState(x, v, m, dt, nleap)
    for k in range(1, 11):  # t = 1.25 < sqrt(2)
        xo, vo = st.onestep()
        t = k * dt
        numpy.testing.assert_allclose(xo, [-(1 - t * t / 2), 1 - t * t / 2],
                                      rtol=0, atol=1e-14)
        numpy.testing.assert_allclose(vo, [t, -t], rtol=0, atol=1e-14)


def test_harmonic_tracer_rotation():
    """Massless sheets in omega^2 x^2/2: the DKD map per step is the linear
    map M = [[1 - (wh)^2/2, h (1 - (wh)^2/4)], [-w^2 h, 1 - (wh)^2/2]]
    (a rotation by theta with cos(theta) = 1 - (wh)^2/2)."""
    w, dt, nleap = 1.1, 0.3, 6
    h = dt / nleap
    x = numpy.array([0.3, -1.2, 2.0])
    v = numpy.array([0.5, 0.1, -0.7])
    m = numpy.zeros(3)
    st = O.ApproxState(x, v, m, dt, nleap, omega=w)
    a = 1 - (w * h) ** 2 / 2
    M = numpy.array([[a, h * (1 - (w * h) ** 2 / 4)], [-w * w * h, a]])
    Mn = numpy.linalg.matrix_power(M, nleap)
    state = numpy.vstack([x, v])
    for _ in range(5):
        xo, vo = st.onestep()
        state = Mn @ state
        numpy.testing.assert_allclose(xo, state[0], rtol=0, atol=1e-13)
        numpy.testing.assert_allclose(vo, state[1], rtol=0, atol=1e-13)
    # the map conserves the modified energy w^2 x^2 (1 - (wh)^2/4)^-1 + v^2
    # up to rounding; check det(M) = 1
    assert abs(numpy.linalg.det(M) - 1) < 1e-14


def test_momentum_conserved_by_oracle():
    x, v, m = O.sech2_disk(200, seed=2, mass=('lognormal', 0.5))
    p0 = math.fsum(m * v)
    st = O.ApproxState(x, v, m, 0.05, 20)
    for _ in range(5):
        _, vo = st.onestep()
    assert abs(math.fsum(m * vo) - p0) < 1e-13


@pytest.mark.parametrize('omega', [None, 0.9])
def test_energy_matches_brute_force(omega):
    rs = numpy.random.RandomState(8)
    x = rs.normal(size=40)
    v = rs.normal(size=40)
    m = rs.uniform(0.1, 1., size=40)
    e = O.energy(x, v, m, twopiG=1.7, omega=omega)
    bf = O.brute_force_energy(x, v, m, twopiG=1.7, omega=omega)
    assert abs(e - bf) < 1e-12 * abs(bf)


def test_energy_individual_and_potential_closed_forms():
    """The reference's hand-expanded checks (tests/test_wendy.py:81-104)."""
    x = numpy.array([-1.1, 0.1, 1.3])
    v = numpy.array([3., 2., -5.])
    m = numpy.array([1., 2., 3.])
    E = O.energy(x, v, m, individual=True)
    assert abs(E[0] - m[0] * v[0] ** 2 / 2 - m[0] * (m[1] * abs(x[0] - x[1])
                                                     + m[2] * abs(x[0] - x[2]))) < 1e-10
    assert abs(E[1] - m[1] * v[1] ** 2 / 2 - m[1] * (m[0] * abs(x[0] - x[1])
                                                     + m[2] * abs(x[2] - x[1]))) < 1e-10
    assert abs(E[2] - m[2] * v[2] ** 2 / 2 - m[2] * (m[0] * abs(x[0] - x[2])
                                                     + m[1] * abs(x[2] - x[1]))) < 1e-10
    y = numpy.arange(-2., 2.5, 0.5)
    p = O.potential(y, x, v, m)
    for ty, tp in zip(y, p):
        assert abs(tp - numpy.sum(m * numpy.fabs(x - ty))) < 1e-10


def test_planted_mistakes_are_caught():
    """Each plausible mistake in the force fails the brute-force check."""
    rs = numpy.random.RandomState(3)
    x = rs.normal(size=30)
    m = rs.uniform(0.5, 1.5, size=30)
    order = O.canonical_order(x, numpy.arange(30, dtype=numpy.int32))
    ms = m[order]
    bf = O.brute_force_accel(x, m)[order]
    tot = numpy.sum(m)
    C = O.exclusive_mass_prefix(ms)
    mistakes = {
        'dropped self-mass term': tot - 2. * C,
        'inclusive prefix': tot - 2. * (C + ms) - ms,
        'sign flipped': -(tot - 2. * C - ms),
        'unsorted masses': tot - 2. * O.exclusive_mass_prefix(m) - m,
    }
    for name, f in mistakes.items():
        assert numpy.max(numpy.abs(f - bf)) > 1e-3, name


# ---- external force (wendy/wendy.c:362-370, clock at :401-408) ----------------
def _ext_forces():
    from wendy_amd import forces
    return [forces.Harmonic(1.1), forces.Isothermal(-0.7, 0.5),
            forces.Pulse(0.3, 0.12, 0.05)]


def _three():
    # tests/test_approx_external.py:7-10,27-30 of the reference
    return (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
            numpy.array([1., 2., 3.]))


@pytest.mark.skipif(not R.available(), reason='oracle/_ref not built here')
@pytest.mark.parametrize('k', [0, 1, 2])
@pytest.mark.parametrize('n', [3, 101])
def test_oracle_ext_force_matches_reference_library(k, n):
    """The oracle's external-force step and clock equal the compiled
    reference driven through the plain ctypes callback, bit for bit, on both
    sides of EXTERNAL_SWITCH (per-sheet scalar calls vs one array call)."""
    f = _ext_forces()[k]
    if n == 3:
        x, v, m = _three()
    else:
        x, v, m = O.sech2_disk(n, seed=4, mass=('lognormal', 0.1))
    ref = R.RefApprox(x, v, m, 0.05, 40, ext_force=f, t0=0.01)
    ora = O.ApproxState(x, v, m, 0.05, 40, ext_force=f, t0=0.01)
    for _ in range(5):
        rx, rv, _ = ref.onestep()
        ox, ov = ora.onestep()
        assert bits_equal(rx, ox) and bits_equal(rv, ov)
    assert ref.t0.value == ora.t


def test_ext_harmonic_equals_omega_for_tracers():
    """Massless sheets in an external harmonic force move exactly as with the
    built-in omega term: -omega**2 x + 0 == 0.0 + (0 - omega2 x)."""
    from wendy_amd import forces
    x, v, _ = O.sech2_disk(50, seed=5)
    m = numpy.zeros(50)
    a = O.ApproxState(x, v, m, 0.1, 7, ext_force=forces.Harmonic(0.8))
    b = O.ApproxState(x, v, m, 0.1, 7, omega=0.8)
    for _ in range(4):
        ax, av = a.onestep()
        bx, bv = b.onestep()
        assert bits_equal(ax, bx) and bits_equal(av, bv)


def test_ext_pulse_momentum_follows_the_clock():
    """A uniform acceleration g(t) changes the total momentum by
    M h sum_k g(t_k) with t_k = t0 + k h (the force time of step k): the
    self-gravity pairs cancel.  A clock off by half a step fails this."""
    from wendy_amd import forces
    f = forces.Pulse(0.5, 0.07, 0.03)
    x, v, m = O.sech2_disk(200, seed=6)
    h, nleap, t0 = 0.01, 5, 0.02
    st = O.ApproxState(x, v, m, h * nleap, nleap, ext_force=f, t0=t0)
    p0 = math.fsum(m * v)
    ts = []
    t = t0
    for _ in range(3 * nleap):
        ts.append(t)
        t = t + h
    for _ in range(3):
        st.onestep()
    dp = math.fsum(m * st.v) - p0
    want = math.fsum(m) * h * math.fsum(f(0., tk) for tk in ts)
    assert abs(dp - want) <= 1e-12
    shifted = math.fsum(m) * h * math.fsum(f(0., tk + h / 2.) for tk in ts)
    assert abs(dp - shifted) > 1e-6
