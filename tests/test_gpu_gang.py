"""Gangs (wdy_gang_*): several large-path simulations advanced by one kernel
launch per leapfrog step.  Each member's result must be the result of
advancing it alone, bit for bit, whatever else is in the gang: the members
run the same arithmetic with their own epochs, look-back words and error
words, and a member whose warm interval fails is rewound and redone on the
radix path by its own next call, as when it runs alone."""
import numpy
import pytest

import wendy_amd
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O

pytestmark = pytest.mark.gpu

# n, seed, masses, omega, h, nleap, tune (bucket target, window)
MEMBERS = [
    (16384, 21, 'equal', 0.1, 0.002, 4, None),
    (70000, 22, 'lognormal', None, 0.004, 3, None),
    (150000, 23, 'lognormal', 0.2, 0.002, 5, None),
    (5000, 24, 'equal', None, 0.003, 2, None),
    (1 << 18, 3, 'equal', None, 0.004, 5, (32, 0)),  # fails and is redone
]


def ic(n, seed, mass):
    return O.sech2_disk(n, seed=seed, mass='equal' if mass == 'equal'
                        else ('lognormal', 0.5))


def make(spec, shared=True):
    n, seed, mass, omega, _, _, tune = spec
    x, v, m = ic(n, seed, mass)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1. if omega is None else omega ** 2.,
                        path='large', shared=shared)
    if tune:
        sim.tune(*tune)
    return sim


def read(sim):
    x = numpy.empty(sim.n)
    v = numpy.empty(sim.n)
    sim.read(x, v)
    return x, v


def same(a, b, what):
    xa, va = read(a)
    xb, vb = read(b)
    assert bits_equal(xa, xb) and bits_equal(va, vb), (what, rel_inf(xa, xb),
                                                       rel_inf(va, vb))


def test_gang_equals_members_alone(gpu_lib):
    gang = wendy_amd.Gang()
    A = [make(s) for s in MEMBERS]
    B = [make(s) for s in MEMBERS]
    hs = [s[4] for s in MEMBERS]
    ns = [s[5] for s in MEMBERS]
    try:
        for out in range(4):
            gang.advance(A, hs, ns)
            for b, h, k in zip(B, hs, ns):
                b.advance(h, k)
            for i, (a, b) in enumerate(zip(A, B)):
                same(a, b, 'member %d output %d' % (i, out + 1))
        for i, (a, b) in enumerate(zip(A, B)):
            pa, pb = a.path_stats(), b.path_stats()
            assert pa == pb, (i, pa, pb)
            if MEMBERS[i][6] is None:
                assert pa[0] > 0, (i, pa)  # warm steps ran (in the gang)
        assert A[-1].path_stats()[2] >= 1  # the tuned member was redone
    finally:
        for s in A + B:
            s.close()
        gang.close()


def test_gang_member_matches_oracle(gpu_lib):
    """Power-of-two equal masses: exact against the oracle."""
    spec = MEMBERS[0]
    n, seed, mass, omega, h, nleap, _ = spec
    x, v, m = ic(n, seed, mass)
    st = O.ApproxState(x, v, m, h * nleap, nleap, omega=omega)
    gang = wendy_amd.Gang()
    sims = [make(spec), make(MEMBERS[1])]
    try:
        for out in range(3):
            gang.advance(sims, [h, MEMBERS[1][4]], [nleap, MEMBERS[1][5]])
            xo, vo = read(sims[0])
            wx, wv = st.onestep()
            assert bits_equal(xo, wx) and bits_equal(vo, wv), out
    finally:
        for s in sims:
            s.close()
        gang.close()


def test_gang_membership_changes(gpu_lib):
    """Members join, leave, are read between intervals and are closed while
    the gang goes on (the serving pattern)."""
    specs = [(40000, 31, 'lognormal', 0.15, 0.002, 3, None),
             (90000, 32, 'equal', None, 0.002, 4, None),
             (60000, 33, 'lognormal', None, 0.003, 2, None)]
    gang = wendy_amd.Gang()
    A = [make(s) for s in specs]
    B = [make(s) for s in specs]
    plan = [[0, 1], [0, 1, 2], [1, 2], [2], [1, 2]]
    try:
        for out, members in enumerate(plan):
            gang.advance([A[i] for i in members], [specs[i][4] for i in members],
                         [specs[i][5] for i in members])
            for i in members:
                B[i].advance(specs[i][4], specs[i][5])
            for i in members:
                same(A[i], B[i], 'member %d output %d' % (i, out + 1))
            if out == 1:
                A[0].close()   # leaves for good; the others go on
    finally:
        for s in A + B:
            s.close()
        gang.close()


def test_gang_of_64(gpu_lib):
    """The most members one call takes (one ballot over their ticket
    starts), each a few buckets, with different step counts."""
    k = wendy_amd.Gang.MAX
    specs = [(2000 + 37 * i, 100 + i, 'lognormal' if i % 3 else 'equal',
              None if i % 2 else 0.1, 0.002, 1 + i % 4, None) for i in range(k)]
    gang = wendy_amd.Gang()
    A = [make(s) for s in specs]
    B = [make(s) for s in specs]
    try:
        for out in range(3):
            gang.advance(A, [s[4] for s in specs], [s[5] for s in specs])
            for b, s in zip(B, specs):
                b.advance(s[4], s[5])
        for i in range(k):
            same(A[i], B[i], 'member %d' % i)
            assert A[i].path_stats() == B[i].path_stats()
    finally:
        for s in A + B:
            s.close()
        gang.close()


def test_gang_argument_errors(gpu_lib):
    gang = wendy_amd.Gang()
    sims = [make(MEMBERS[3]), make(MEMBERS[0])]
    try:
        with pytest.raises(ValueError):
            gang.advance([sims[0], sims[0]], 0.002, 2)
        with pytest.raises(ValueError):
            gang.advance(sims, 0.002, 0)
        with pytest.raises(ValueError):
            gang.advance(sims, float('nan'), 2)
        big = [sims[0]] * (wendy_amd.Gang.MAX + 1)
        with pytest.raises(ValueError):
            gang.advance(big, 0.002, 2)
        gang.advance(sims, 0.002, 2)   # still usable
        for s in sims:
            read(s)
    finally:
        for s in sims:
            s.close()
        gang.close()


def test_gang_members_on_other_shapes_run_alone(gpu_lib):
    """Members whose tile shape has no gang form (a simulation made for a GPU
    of its own starts from the shape fastest for its size) run their
    intervals on their own streams: same results as alone, no gang steps
    needed."""
    specs = [MEMBERS[1], MEMBERS[2]]
    gang = wendy_amd.Gang()
    A = [make(s, shared=False) for s in specs] + [make(MEMBERS[0])]
    B = [make(s, shared=False) for s in specs] + [make(MEMBERS[0])]
    hs = [s[4] for s in specs] + [MEMBERS[0][4]]
    ns = [s[5] for s in specs] + [MEMBERS[0][5]]
    try:
        for out in range(3):
            gang.advance(A, hs, ns)
            for b, h, k in zip(B, hs, ns):
                b.advance(h, k)
            for i, (a, b) in enumerate(zip(A, B)):
                same(a, b, 'member %d output %d' % (i, out + 1))
        for a, b in zip(A, B):
            assert a.path_stats() == b.path_stats() and a.path_stats()[0] > 0
    finally:
        for s in A + B:
            s.close()
        gang.close()
