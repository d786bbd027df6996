"""The large path's warm-start step (one kernel per leapfrog step) against the
oracle, including its corner paths, reached at small N by shrinking its
limits (as the reference's CI shrinks the parallel-sort thresholds,
.github/workflows/build.yml:30): tiny buckets, a zero window (every sheet that
changes bucket travels through the overflow list), and forced failures that
make the library redo an interval on the radix path.

Bars as in test_gpu_parity: bitwise for power-of-two equal masses, else
||.||_inf relative 1e-10."""
import numpy
import pytest

import wendy_amd
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


def make(n, seed, mass='equal', vshift=None):
    return O.sech2_disk(n, seed=seed, mass='equal' if mass == 'equal'
                        else ('lognormal', 0.5), vshift=vshift)


def exact_for(m):
    return bool(numpy.all(m == m[0]) and numpy.frexp(m[0])[0] == 0.5)


def check(got, want, exact, what):
    if exact:
        assert bits_equal(got[0], want[0]), (what, rel_inf(got[0], want[0]))
        assert bits_equal(got[1], want[1]), (what, rel_inf(got[1], want[1]))
    else:
        assert rel_inf(got[0], want[0]) <= TOL, (what, rel_inf(got[0], want[0]))
        assert rel_inf(got[1], want[1]) <= TOL, (what, rel_inf(got[1], want[1]))


def run_both(x, v, m, dt, nleap, nout, omega=None, tune=None, warm=True,
             hs=None):
    """Run the large path and the oracle side by side; returns the sim's
    path statistics.  ``hs`` optionally gives each output's leapfrog h."""
    omega2 = -1. if omega is None else omega ** 2.
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), omega2, path='large', warm=warm)
    if tune:
        sim.tune(*tune)
    st = O.ApproxState(x, v, m, dt, nleap, omega=omega)
    xo = numpy.empty(len(x))
    vo = numpy.empty(len(x))
    exact = exact_for(m)
    try:
        for k in range(nout):
            h = dt / nleap if hs is None else hs[k]
            st.h = h
            sim.advance(h, nleap)
            sim.read(xo, vo)
            check((xo, vo), st.onestep(), exact, 'output %d' % (k + 1))
        return sim.path_stats()
    finally:
        sim.close()


@pytest.mark.parametrize('n,mass,omega', [
    (32768, 'equal', 0.1),
    (20000, 'equal', None),
    (30000, 'lognormal', None),
    (25000, 'lognormal', 0.3),
])
def test_warm_matches_oracle(gpu_lib, n, mass, omega):
    print('warm', n, mass, omega, flush=True)
    x, v, m = make(n, n, mass)
    warm, cold, redo = run_both(x, v, m, 0.01, 5, 4, omega=omega)
    assert warm == 15 and cold == 5 and redo == 0


def test_warm_bulk_flow(gpu_lib):
    """The large unequal-mass workload's shape: +0.3 bulk velocity on x > 0."""
    x, v, m = make(40000, 7, 'lognormal', vshift=0.3)
    warm, cold, redo = run_both(x, v, m, 0.005, 10, 3)
    assert warm == 20 and redo == 0


@pytest.mark.parametrize('tune', [(64, 2), (32, 1), (100, 5)])
def test_warm_many_small_buckets(gpu_lib, tune):
    x, v, m = make(6000, 1, 'lognormal')
    warm, _, _ = run_both(x, v, m, 0.02, 4, 4, tune=tune)
    assert warm > 0


@pytest.mark.parametrize('mass', ['equal', 'lognormal'])
def test_warm_zero_window_uses_overflow_list(gpu_lib, mass):
    """Window 0: every sheet that changes bucket is a far mover."""
    x, v, m = make(8192, 2, mass)
    warm, _, redo = run_both(x, v, m, 0.02, 4, 4, tune=(64, 0))
    assert warm > 0


def test_warm_redo_after_overflow(gpu_lib):
    """Far movers beyond the overflow list's capacity make the warm interval
    fail on the device; the library rewinds, redoes it on the radix path,
    widens the window and goes on -- with unchanged results."""
    n = 1 << 18
    x, v, m = make(n, 3)
    warm, cold, redo = run_both(x, v, m, 0.02, 5, 5, tune=(32, 0))
    assert redo >= 1


def test_warm_capacity_limit_bucket(gpu_lib):
    """Buckets targeted at the LDS capacity: any that outgrows it fails the
    interval, which is redone; results are unchanged either way."""
    x, v, m = make(1 << 17, 4)
    warm, cold, redo = run_both(x, v, m, 0.02, 5, 4, tune=(4096, 8, 0))
    assert warm + cold == 20


def test_warm_h_change_between_outputs(gpu_lib):
    """A different h invalidates the anticipated half drift: that interval
    runs on the radix path, then the warm path resumes."""
    x, v, m = make(20000, 5)
    hs = [0.002, 0.002, 0.001, 0.001, 0.002]
    warm, cold, redo = run_both(x, v, m, 0.01, 5, 5, hs=hs)
    assert cold == 15 and warm == 10


def test_warm_equals_radix_path(gpu_lib):
    """Equal masses: the warm kernel and the radix path compute the same
    ranks, so their trajectories agree bit for bit even for masses that are
    not powers of two."""
    x, v, m = make(50000, 6)
    outs = []
    for warm in (True, False):
        sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.01, path='large', warm=warm)
        xo = numpy.empty(len(x))
        vo = numpy.empty(len(x))
        for _ in range(3):
            sim.advance(0.001, 6)
        sim.read(xo, vo)
        outs.append((xo, vo))
        sim.close()
    assert bits_equal(outs[0][0], outs[1][0]) and bits_equal(outs[0][1], outs[1][1])


def test_warm_unequal_deterministic(gpu_lib):
    x, v, m = make(60000, 8, 'lognormal')
    res = []
    for _ in range(2):
        sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large')
        xo = numpy.empty(len(x))
        vo = numpy.empty(len(x))
        for _ in range(3):
            sim.advance(0.002, 5)
        sim.read(xo, vo)
        res.append((xo, vo))
        assert sim.path_stats()[0] == 10
        sim.close()
    assert bits_equal(res[0][0], res[1][0]) and bits_equal(res[0][1], res[1][1])


@pytest.mark.parametrize('shape', [0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize('mass', ['equal', 'lognormal'])
def test_warm_every_shape(gpu_lib, shape, mass):
    """Every tile shape (threads, bucket capacity, position bins) gives the
    oracle's trajectory."""
    print('shape', shape, mass, flush=True)
    x, v, m = make(40000, 10 + shape, mass)
    warm, cold, redo = run_both(x, v, m, 0.01, 5, 3, omega=0.1, tune=(0, 8, shape))
    assert warm == 10 and cold == 5 and redo == 0


def test_warm_window_limit_per_shape(gpu_lib):
    x, v, m = make(20000, 9)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large')
    try:
        sim.tune(0, 31, 6)
        with pytest.raises(ValueError):
            sim.tune(0, 32, 6)
        with pytest.raises(ValueError):
            sim.tune(0, 64, 3)
        sim.tune(0, 64, 0)
    finally:
        sim.close()


@pytest.mark.parametrize('mass', ['equal', 'lognormal'])
@pytest.mark.parametrize('pipe', ['0', '1'])
def test_warm_kernel_forms_agree(gpu_lib, monkeypatch, mass, pipe):
    """The one-workgroup-per-bucket kernel (WDY_WARM_PIPE=0) and the
    persistent pipelined kernel (the default) both give the oracle's
    trajectory, including through a redo and small buckets."""
    monkeypatch.setenv('WDY_WARM_PIPE', pipe)
    print('pipe', pipe, mass, flush=True)
    x, v, m = make(50000, 21, mass)
    warm, cold, redo = run_both(x, v, m, 0.01, 5, 3, omega=0.1)
    assert warm == 10 and redo == 0
    x, v, m = make(9000, 22, mass)
    warm, _, _ = run_both(x, v, m, 0.02, 4, 3, tune=(64, 2))
    assert warm > 0


def test_first_warm_interval_after_fitting_succeeds(gpu_lib):
    """The window fitted to the sampled displacement (here wider than the
    default) is applied before the first partition, so the first warm
    interval does not fail."""
    x, v, m = make(1 << 20, 30)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.01, path='large')
    try:
        for _ in range(3):
            sim.advance(0.01, 5)   # ~1e4 ranks per step for the fastest sheets
        sim.read()
        assert sim.path_stats() == (10, 5, 0)
        assert sim.warm_failures() == (0, 0, 0)
    finally:
        sim.close()


@pytest.mark.parametrize('h,shape', [(0.001, 8), (0.06, 8), (0.15, 7), (0.4, 0)])
def test_window_fitting_ladder(gpu_lib, h, shape):
    """The window fitted at the first interval from the sampled displacement
    (N = 2^18, starting from the throughput shape a shared GPU uses): the
    default 1536-sheet buckets while the window stays within 48 buckets
    either side, then 2048-sheet buckets up to 60, then 4096-sheet ones; in
    every case both intervals equal the oracle bit for bit (equal masses
    2^-18)."""
    print('fitting ladder h', h, flush=True)
    x, v, m = O.sech2_disk(1 << 18, seed=3)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large', shared=True)
    st = O.ApproxState(x, v, m, 2 * h, 2, presorted_order=True)
    xo, vo = numpy.empty(len(x)), numpy.empty(len(x))
    try:
        sim.advance(h, 2)
        sim.read(xo, vo)
        sh, T, W = sim.warm_config()
        assert sh == shape, (h, sh, T, W)
        assert 6 <= W <= {8: 48, 7: 60, 0: 64}[shape], (h, sh, T, W)
        wx, wv = st.onestep()
        assert bits_equal(xo, wx) and bits_equal(vo, wv)
        sim.advance(h, 2)   # warm, or redone on the radix path
        sim.read(xo, vo)
        wx, wv = st.onestep()
        assert bits_equal(xo, wx) and bits_equal(vo, wv)
        assert sum(sim.path_stats()[:2]) == 4  # warm + radix steps
    finally:
        sim.close()


@pytest.mark.parametrize('lg,mass,shape', [
    (14, 'equal', 3), (16, 'lognormal', 3), (18, 'equal', 1),
    (20, 'equal', 5), (20, 'lognormal', 6), (21, 'lognormal', 6),
    (21, 'equal', 0), (22, 'lognormal', 7)])
def test_initial_shape_by_size(gpu_lib, lg, mass, shape):
    """A simulation that has the GPU to itself starts from the tile shape
    measured fastest for its size (larger workgroups while its buckets do not
    fill the GPU); a shared one (a gang member) from the default; both give
    the oracle's trajectory."""
    print('initial shape 2^%d %s' % (lg, mass), flush=True)
    x, v, m = make(1 << lg, 40 + lg, mass)
    exact = exact_for(m)
    for shared, want in ((False, shape), (True, 8)):
        sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.1 ** 2., path='large',
                            shared=shared)
        st = O.ApproxState(x, v, m, 0.004, 4, omega=0.1)
        xo, vo = numpy.empty(len(x)), numpy.empty(len(x))
        try:
            for k in range(2):
                sim.advance(0.001, 4)
                sim.read(xo, vo)
                check((xo, vo), st.onestep(), exact, 'output %d' % (k + 1))
            assert sim.warm_config()[0] == want, (shared, sim.warm_config())
            assert sim.path_stats()[0] == 4   # the second interval ran warm
        finally:
            sim.close()


def test_autotune_picks_a_shape_and_keeps_results(gpu_lib):
    """Sim.autotune times candidate tile shapes on a copy of the state; the
    simulation itself continues unchanged (its trajectory still matches the
    oracle)."""
    x, v, m = make(60000, 40)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.01, path='large')
    st = O.ApproxState(x, v, m, 0.005, 5, omega=0.1)
    xo = numpy.empty(len(x))
    vo = numpy.empty(len(x))
    try:
        sim.advance(0.001, 5)
        times = sim.autotune(0.001, 5, candidates=((2, 8), (0, 8)), trials=1)
        assert set(times) == {(2, 8), (0, 8)}
        assert all(t > 0 for t in times.values())
        sim.read(xo, vo)
        wx, wv = st.onestep()
        assert bits_equal(xo, wx) or rel_inf(xo, wx) <= TOL
        for _ in range(2):
            sim.advance(0.001, 5)
            sim.read(xo, vo)
            wx, wv = st.onestep()
            assert rel_inf(xo, wx) <= TOL and rel_inf(vo, wv) <= TOL
    finally:
        sim.close()


def test_warm_large_groups_of_exact_ties(gpu_lib):
    """Massless sheets on a coarse grid of positions and velocities: groups
    of hundreds of exactly tied sheets persist through every step (equal
    (x, v), equal force), spanning bucket boundaries. The warm path sorts
    them by (position, index) and matches the oracle bit for bit."""
    rng = numpy.random.default_rng(5)
    n = 40000
    x = numpy.round(rng.standard_normal(n) * 8.) / 8.
    v = numpy.round(rng.standard_normal(n) * 4.) / 4.
    m = numpy.zeros(n)
    warm, cold, redo = run_both(x, v, m, 0.02, 5, 4, omega=0.7)
    assert warm + cold == 20


def test_warm_cold_collapse(gpu_lib):
    """A uniform slab released from rest collapses: velocities grow from zero
    to order one within a few outputs, so the window sized at the start is
    soon too narrow. Whatever the warm path does (widen, move to larger
    buckets, redo on the radix path), the trajectory stays the oracle's."""
    rng = numpy.random.default_rng(11)
    n = 1 << 16
    x = rng.uniform(-1., 1., n)
    v = numpy.zeros(n)
    m = numpy.full(n, 1. / n)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large')
    st = O.ApproxState(x, v, m, 0.2, 10)
    xo = numpy.empty(n)
    vo = numpy.empty(n)
    try:
        for k in range(8):    # through the first collapse (t ~ 1.4)
            sim.advance(0.02, 10)
            sim.read(xo, vo)
            wx, wv = st.onestep()
            assert bits_equal(xo, wx) and bits_equal(vo, wv), k
        warm, cold, redo = sim.path_stats()
        print('collapse paths', warm, cold, redo, sim.warm_failures(), flush=True)
        assert warm + cold == 80
    finally:
        sim.close()


def test_advance_returns_before_the_interval_finishes(gpu_lib):
    """wdy_advance only enqueues a warm interval: the host call returns in a
    fraction of the interval's device time (the error word is checked by the
    next call on the handle)."""
    import time
    n = 1 << 23
    x, v, m = make(n, 12)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), 0.01, path='large')
    try:
        h, nleap = 0.005 / 40, 40
        for _ in range(2):                  # cold interval, then a warm one
            sim.advance(h, nleap)
            sim.read()
        t0 = time.perf_counter()
        sim.advance(h, nleap)
        host = time.perf_counter() - t0
        dev = sim.read()
        warm, cold, redo = sim.path_stats()
        print('advance host %.3f ms, interval %.3f ms' % (host * 1e3, dev * 1e3),
              flush=True)
        assert redo == 0 and warm == 2 * nleap
        assert dev > 2e-3 and host < 0.5 * dev
    finally:
        sim.close()


@pytest.mark.parametrize('reader', ['advance', 'read_and_advance'])
def test_warm_failure_found_by_the_next_call(gpu_lib, reader):
    """Intervals whose warm steps fail on the device (window 0 and tiny
    buckets: the overflow list runs over) are enqueued without waiting; the
    next call on the handle -- another advance, or the generator's
    read_and_advance -- finds the failure, rewinds and redoes the interval on
    the radix path before its own work.  Outputs stay the oracle's."""
    n = 1 << 18
    x, v, m = make(n, 3)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large')
    sim.tune(32, 0)
    dt, nleap = 0.02, 5
    st = O.ApproxState(x, v, m, dt, nleap)
    xo = numpy.empty(n)
    vo = numpy.empty(n)
    try:
        sim.advance(dt / nleap, nleap)      # cold: builds the warm layout
        want = [st.onestep()]
        if reader == 'advance':
            for _ in range(3):              # each settles the one before
                sim.advance(dt / nleap, nleap)
                want.append(st.onestep())
            sim.read(xo, vo)
            check((xo, vo), want[-1], True, 'after 4 outputs')
        else:
            for k in range(4):
                sim.read_and_advance(xo, vo, dt / nleap, nleap)
                check((xo, vo), want[k], True, 'output %d' % (k + 1))
                want.append(st.onestep())
            sim.read(xo, vo)
            check((xo, vo), want[4], True, 'output 5')
        warm, cold, redo = sim.path_stats()
        print('deferred redo', reader, warm, cold, redo, sim.warm_failures(),
              flush=True)
        assert redo >= 1
    finally:
        sim.close()
