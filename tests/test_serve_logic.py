"""Scheduling logic of wendy_amd.serve.Server, on the CPU with stand-in pool
and large-simulation objects that record what they are asked to do: every
request gets exactly its steps in ticks of min(tick, steps left), requests
are admitted first come first served against slots and the HBM budget, and
arrivals during a tick join the next one."""
import numpy
import pytest

from wendy_amd import serve


class FakePool:
    def __init__(self, capacity, max_sims):
        self.capacity, self.max_sims = capacity, max_sims
        self.live = {}           # slot -> [n, [k, ...], h]
        self.used = 0
        self.max_live = 0
        self.steps = []

    def add(self, x, v, m, omega=None, twopiG=1.):
        n = len(x)
        if len(self.live) >= self.max_sims or self.used + n > self.capacity:
            return None
        slot = min(set(range(self.max_sims)) - set(self.live))
        self.live[slot] = [n, [], None, x.copy()]
        self.used += n
        self.max_live = max(self.max_live, len(self.live))
        return slot

    def step(self, slots, h, nleap):
        assert len(set(slots)) == len(slots)
        self.steps.append(len(slots))
        for s, hh, k in zip(slots, h, nleap):
            self.live[s][1].append(int(k))
            self.live[s][2] = hh

    def sync(self):
        return 0.

    def read(self, slot):
        n, ks, h, x = self.live[slot]
        return x + sum(ks), numpy.full(n, float(len(ks)))

    def remove(self, slot):
        self.used -= self.live[slot][0]
        del self.live[slot]

    def close(self):
        pass


class FakeBig:
    made = []

    def __init__(self, req, x, v, m):
        self.x = x.copy()
        self.ks = []
        FakeBig.made.append(self)

    def advance(self, h, k):
        self.ks.append(k)

    def finish(self, n):
        return self.x + sum(self.ks), numpy.full(n, float(len(self.ks)))


class FakeGang:
    """Advances its members one interval each per call, as wdy_gang_advance
    does (at most MAX members per call)."""
    MAX = 3
    calls = []

    def advance(self, bigs, hs, ks):
        assert 1 <= len(bigs) <= self.MAX and len(set(map(id, bigs))) == len(bigs)
        FakeGang.calls.append(len(bigs))
        for b, h, k in zip(bigs, hs, ks):
            b.advance(h, k)

    def close(self):
        pass


def req(i, n, steps, arrival=0.):
    return serve.Request(i, n, lambda n=n, i=i: (numpy.full(n, float(i)),
                                                 numpy.zeros(n), numpy.ones(n)),
                         steps, 0.01, arrival=arrival)


@pytest.fixture(params=['gang', 'threads'])
def large(request):
    return request.param


def make_server(cap=10_000, slots=4, large='gang', **kw):
    pools = []

    def pf():
        pools.append(FakePool(cap, slots))
        return pools[0]
    FakeBig.made = []
    FakeGang.calls = []
    s = serve.Server(budget_bytes=kw.pop('budget', 1e12), pool_sheets=cap,
                     max_sims=slots, pool_factory=pf, big_factory=FakeBig,
                     gang_factory=FakeGang, large=large, **kw)
    return s, pools[0]


def test_every_request_gets_its_steps_in_ticks(large):
    s, pool = make_server(split=100, tick=10, large=large)
    reqs = [req(0, 50, 25), req(1, 100, 10), req(2, 500, 7), req(3, 101, 31),
            req(4, 1, 1)]
    res = s.run(reqs)
    assert sorted(res) == [0, 1, 2, 3, 4]
    for r in reqs:
        out = res[r.id]
        nt = -(-r.steps // 10)
        # x advanced by the total step count, in ceil(steps/tick) intervals
        assert numpy.all(out.x == r.id + r.steps)
        assert numpy.all(out.v == nt)
        assert out.where == ('pool' if r.n <= 100 else 'sim')
        assert out.completed >= out.admitted >= out.arrival
    for b in FakeBig.made:
        assert all(k == 10 for k in b.ks[:-1]) and 1 <= b.ks[-1] <= 10
    assert s.stats['sheet_steps'] == sum(r.n * r.steps for r in reqs)


def test_slots_limit_admission_first_come_first_served():
    s, pool = make_server(slots=2, split=1000, tick=5)
    reqs = [req(i, 10, 5 * (i + 1)) for i in range(6)]
    res = s.run(reqs)
    assert len(res) == 6 and pool.max_live == 2
    assert s.stats['admission_stalls'] > 0
    adm = [res[i].admitted for i in range(6)]
    assert adm == sorted(adm)            # admitted in arrival order
    assert not pool.live and pool.used == 0


def test_arena_capacity_limits_admission():
    s, pool = make_server(cap=100, slots=8, split=1000, tick=3)
    res = s.run([req(i, 60, 4) for i in range(3)])
    assert len(res) == 3 and pool.max_live == 1


def test_budget_limits_large_simulations(large):
    per = serve.SIM_BYTES_PER_SHEET
    pool_bytes = 100 * serve.POOL_BYTES_PER_SHEET
    s, _ = make_server(cap=100, split=10, budget=pool_bytes + 2.5 * 1000 * per,
                       max_big=8, tick=4, large=large)
    res = s.run([req(i, 1000, 8) for i in range(5)])
    assert len(res) == 5 and s.stats['admission_stalls'] > 0
    # at most two 1000-sheet simulations fit beside the arena at once
    spans = sorted((r.admitted, r.completed) for r in res.values())
    for i in range(len(spans)):
        live = sum(1 for a, c in spans if a <= spans[i][0] < c)
        assert live <= 2


def test_request_larger_than_budget_is_refused(large):
    s, _ = make_server(cap=100, split=10, budget=100 * serve.POOL_BYTES_PER_SHEET + 10,
                       large=large)
    with pytest.raises(MemoryError):
        s.run([req(0, 1000, 3)])


def test_late_arrivals_wait_for_their_time():
    s, pool = make_server(split=1000, tick=2)
    res = s.run([req(0, 10, 4, arrival=0.0), req(1, 10, 2, arrival=0.05)])
    assert res[1].admitted >= 0.05
    assert res[1].latency >= 0.


def test_latency_summary():
    s, _ = make_server(split=1000, tick=2)
    res = s.run([req(i, 10, 2 + i) for i in range(10)])
    summ = serve.latency_summary(res)
    assert summ['completed'] == 10 and summ['p99_s'] >= summ['p50_s'] >= 0.


def test_gang_members_capped_per_call():
    """More large-path simulations than one gang call takes: they advance in
    arrival order, at most MAX per call, each by exactly its steps."""
    s, _ = make_server(split=10, tick=3, max_big=10, large='gang')
    reqs = [req(i, 100 + i, 7 + 2 * i) for i in range(8)]
    res = s.run(reqs)
    assert sorted(res) == list(range(8))
    for r in reqs:
        assert numpy.all(res[r.id].x == r.id + r.steps)
        assert res[r.id].where == 'sim'
    assert FakeGang.calls and max(FakeGang.calls) <= FakeGang.MAX
    assert s.stats['gang_intervals'] == len(FakeGang.calls)
