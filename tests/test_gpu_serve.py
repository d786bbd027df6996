"""Serving (wendy_amd.serve): a simulation's result is the same alone and in
the mix.  Requests of both kinds (pool-resident n <= split, large-path n >
split), unequal and equal masses, with and without omega, arrive staggered
while others run; each result is compared bitwise with the same request
served alone, and with the oracle (the same sequence of output intervals) at
||.||_inf relative 1e-10."""
import numpy
import pytest

import wendy_amd
from wendy_amd import serve
from conftest import rel_inf
from oracle import nbody_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


def requests(seed=0):
    rng = numpy.random.default_rng(seed)
    specs = [(300, 23, None, 'lognormal'), (5000, 30, 0.1, 'lognormal'),
             (4096, 17, None, 'equal'), (1, 5, None, 'equal'),
             (20000, 12, 0.05, 'lognormal'), (40000, 21, None, 'equal'),
             (777, 40, 0.2, 'lognormal'), (9000, 9, None, 'lognormal')]
    reqs = []
    for i, (n, steps, omega, mass) in enumerate(specs):
        x, v, m = O.sech2_disk(n, seed=100 + i, zh=rng.uniform(0.5, 2.),
                               mass='equal' if mass == 'equal' else ('lognormal', 0.3))
        reqs.append(serve.Request(i, n, lambda a=(x, v, m): a, steps, 0.01,
                                  omega=omega, arrival=0.002 * i))
    return reqs


def oracle_result(r, tick):
    x, v, m = r.make()
    st = O.ApproxState(x, v, m, r.h * tick, tick, omega=r.omega)
    left = r.steps
    while left:
        k = min(tick, left)
        st.nleap = k
        st.h = r.h
        out = st.onestep()
        left -= k
    return out


def test_alone_equals_in_the_mix(gpu_lib):
    print('serve mix', flush=True)
    reqs = requests()
    # split 8192: the two largest requests take the large path
    s = serve.Server(budget_bytes=4e9, pool_sheets=1 << 20, max_sims=64,
                     split=8192, max_big=2, tick=7)
    try:
        mix = s.run(reqs)
        assert sorted(mix) == list(range(len(reqs)))
        assert {mix[i].where for i in mix} == {'pool', 'sim'}
        for r in reqs:
            alone = s.run([serve.Request(r.id, r.n, r.make, r.steps, r.h,
                                         omega=r.omega)])[r.id]
            got = mix[r.id]
            assert numpy.array_equal(got.x, alone.x), r.id
            assert numpy.array_equal(got.v, alone.v), r.id
            wx, wv = oracle_result(r, 7)
            assert rel_inf(got.x, wx) <= TOL, (r.id, rel_inf(got.x, wx))
            assert rel_inf(got.v, wv) <= TOL, (r.id, rel_inf(got.v, wv))
    finally:
        s.close()


def test_pool_slots_are_reused_and_isolated(gpu_lib):
    """A pool with room for two simulations serves six: ranges and slots
    freed by finished simulations are reused without disturbing the one
    still running."""
    print('serve reuse', flush=True)
    reqs = requests(1)[:6]
    s = serve.Server(budget_bytes=1e9, pool_sheets=12000, max_sims=2,
                     split=8192, max_big=1, tick=5)
    try:
        res = s.run(reqs)
        assert s.stats['admission_stalls'] > 0
        for r in reqs:
            wx, wv = oracle_result(r, 5)
            assert rel_inf(res[r.id].x, wx) <= TOL, r.id
            assert rel_inf(res[r.id].v, wv) <= TOL, r.id
    finally:
        s.close()


def test_pool_rejects_bad_calls(gpu_lib):
    p = serve.Pool(1000, max_sims=2, max_n=500)
    try:
        x = numpy.zeros(600)
        with pytest.raises(ValueError):
            p.add(x, x, x + 1.)            # n above max_n
        a = p.add(numpy.arange(400.), numpy.zeros(400), numpy.ones(400))
        b = p.add(numpy.arange(400.), numpy.zeros(400), numpy.ones(400))
        assert {a, b} == {0, 1}
        assert p.add(numpy.arange(10.), numpy.zeros(10), numpy.ones(10)) is None
        p.remove(a)
        assert p.add(numpy.arange(300.), numpy.zeros(300), numpy.ones(300)) == a
        with pytest.raises(ValueError):
            p.step([b, b], 0.01, 3)         # slot listed twice
        with pytest.raises(ValueError):
            p.step([5], 0.01, 3)            # not a live slot
        p.step([a, b], [0.01, 0.02], [3, 0])
        p.sync()
        live, used, big = p.stats()
        assert live == 2 and used == 700 and big == 200
    finally:
        p.close()
