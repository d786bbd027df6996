"""The serving load itself (Poisson arrivals at 200/s; N log-uniform in
[2^10, 2^20]; sech^2 disks with z_h ~ U(0.5, 2), sigma = sqrt(z_h);
lognormal(0, 0.3) masses; omega ~ U(0, 0.2); h = 0.01), drawn from the same
generator the benchmark uses (bench.serve_workload), over a short arrival
window and with the step counts capped so the run stays short; both kinds of
request are present: pool-resident (n <= 2^14) and large-path (up to 2^20,
advancing together in gangs beside the pool's launches).

Each result is compared bit for bit with the same request served alone (a
fresh server holding only it, a large-path request in a Sim of its own),
each request of n <= 2^14 with the oracle
(the same sequence of output intervals) at ||.||_inf relative 1e-10, and
every result's energy with its initial energy."""
import numpy
import pytest

import bench
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O
from wendy_amd import serve
import wendy_amd

pytestmark = pytest.mark.gpu

STEP_CAP = 50
TICK = 10


def capped(reqs):
    return [serve.Request(r.id, r.n, r.make, min(r.steps, STEP_CAP), r.h,
                          omega=r.omega, arrival=r.arrival) for r in reqs]


SPLIT = bench.SERVE_SPLIT


def server():
    # the benchmark's configuration (bench.bench_serve): large-path requests
    # advance together in gangs
    return serve.Server(budget_bytes=200e9, max_sims=2048, split=SPLIT,
                        tick=TICK, large='gang', ngangs=2)


def oracle_result(r):
    x, v, m = r.make()
    st = O.ApproxState(x, v, m, r.h * TICK, TICK, omega=r.omega)
    left = r.steps
    while left:
        k = min(TICK, left)
        st.nleap = k
        st.h = r.h
        out = st.onestep()
        left -= k
    return out


def test_serving_load_alone_equals_mix_and_oracle(gpu_lib):
    reqs = capped(bench.serve_workload(0.3, 200., seed=3))
    ns = numpy.array([r.n for r in reqs])
    print('serve load: %d requests, n %d..%d, %d above 2^16'
          % (len(reqs), ns.min(), ns.max(), int((ns > 1 << 16).sum())), flush=True)
    assert len(reqs) >= 40 and (ns > 1 << 16).sum() >= 5 and (ns <= 1 << 14).sum() >= 5
    s = server()
    try:
        mix = s.run(reqs)
    finally:
        s.close()
    assert sorted(mix) == sorted(r.id for r in reqs)
    print('mix served', s.stats, flush=True)
    for r in reqs:
        got = mix[r.id]
        assert got.where == ('pool' if r.n <= SPLIT else 'sim')
        if got.where == 'sim':
            warm, cold, redo, cap, ovf, stall = got.paths
            # the first interval sorts from scratch, later ones run warm
            # (an interval the device rejects is redone on the radix path)
            assert warm + cold == r.steps, got.paths
            assert stall == 0 and (r.steps <= TICK or warm > 0), got.paths
        # alone: a fresh server (a small pool suffices) holding only it, a
        # large-path request in a Sim of its own
        alone_s = serve.Server(budget_bytes=200e9, pool_sheets=1 << 17,
                               max_sims=4, split=SPLIT, max_big=1, tick=TICK,
                               large='threads')
        try:
            alone = alone_s.run([serve.Request(r.id, r.n, r.make, r.steps, r.h,
                                               omega=r.omega)])[r.id]
        finally:
            alone_s.close()
        assert bits_equal(got.x, alone.x) and bits_equal(got.v, alone.v), r.id
        x0, v0, m0 = r.make()
        e0 = wendy_amd.energy(x0, v0, m0, omega=r.omega)
        e1 = wendy_amd.energy(got.x, got.v, m0, omega=r.omega)
        assert abs(e1 - e0) <= 1e-4 * abs(e0), (r.id, r.n, e0, e1)
        if r.n <= 1 << 14:
            ox, ov = oracle_result(r)
            assert rel_inf(got.x, ox) <= 1e-10 and rel_inf(got.v, ov) <= 1e-10, r.id
