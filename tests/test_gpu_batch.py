"""Batched ensembles (wdy_batch_*): every simulation of a batch must give what
it gives alone -- the oracle's result, to the same bars as the single
simulation path (bitwise for power-of-two equal masses, else ||.||_inf
relative 1e-10) -- whatever its size and whatever else shares the launch.

Sizes cover both device regimes: n <= 2048 (the simulation stays in one
workgroup's LDS for the interval) and n > 2048 (arena-resident: one pass of
a sliding sorted window per step, redone by a block odd-even merge sort when
a sheet moved too far -- always on the first step, from the unsorted input
order), with partial first windows, n = 1 and exactly tied positions."""
import numpy
import pytest

import wendy_amd
from conftest import bits_equal, rel_inf
from oracle import nbody_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


def exact_for(m):
    return bool(numpy.all(m == m[0]) and numpy.frexp(m[0])[0] == 0.5)


def ensemble(specs, seed=0):
    """specs: (n, mass kind, omega) -> lists x, v, m (m without twopiG), omega."""
    xs, vs, ms, oms = [], [], [], []
    for k, (n, mass, omega) in enumerate(specs):
        if n < 8:
            rng = numpy.random.default_rng(seed + k)
            x = rng.standard_normal(n)
            v = rng.standard_normal(n)
            m = numpy.full(n, 1. / 4.) if mass == 'equal' else rng.uniform(0.5, 2., n)
        else:
            x, v, m = O.sech2_disk(n, seed=seed + k, mass='equal' if mass == 'equal'
                                   else ('lognormal', 0.5))
        if mass == 'ties':
            x = numpy.round(x, 1)   # many exactly tied positions
        xs.append(x)
        vs.append(v)
        ms.append(m)
        oms.append(omega)
    return xs, vs, ms, oms


def run_batch_vs_oracle(specs, dt, nleap, nout, hs=None, seed=0):
    xs, vs, ms, oms = ensemble(specs, seed)
    b = wendy_amd.Batch(xs, vs, ms, oms)
    states = [O.ApproxState(x, v, m, dt, nleap, omega=om)
              for x, v, m, om in zip(xs, vs, ms, oms)]
    try:
        for k in range(nout):
            h = dt / nleap if hs is None else hs[k]
            b.advance(h, nleap)
            gx, gv, sec = b.read()
            assert sec > 0.
            for j, st in enumerate(states):
                st.h = h
                wx, wv = st.onestep()
                what = 'sim %d (n=%d %s) output %d' % (j, len(xs[j]), specs[j][1], k + 1)
                if exact_for(ms[j]):
                    assert bits_equal(gx[j], wx), (what, rel_inf(gx[j], wx))
                    assert bits_equal(gv[j], wv), (what, rel_inf(gv[j], wv))
                else:
                    assert rel_inf(gx[j], wx) <= TOL, (what, rel_inf(gx[j], wx))
                    assert rel_inf(gv[j], wv) <= TOL, (what, rel_inf(gv[j], wv))
    finally:
        b.close()


def test_batch_lds_resident_sizes(gpu_lib):
    print('batch lds sizes', flush=True)
    specs = [(1, 'equal', None), (2, 'unequal', 0.1), (3, 'equal', 0.2),
             (64, 'equal', None), (100, 'lognormal', None), (777, 'lognormal', 0.05),
             (1024, 'equal', 0.1), (2000, 'ties', None), (2047, 'lognormal', None),
             (2048, 'equal', None)]
    run_batch_vs_oracle(specs, 0.01, 5, 3)


def test_batch_arena_resident_sizes(gpu_lib):
    print('batch arena sizes', flush=True)
    specs = [(2049, 'equal', None), (2500, 'ties', None), (3585, 'lognormal', 0.1),
             (4096, 'equal', None), (4097, 'equal', None), (6000, 'lognormal', 0.1),
             (8192, 'equal', 0.05), (10000, 'ties', None), (20000, 'lognormal', None),
             (33000, 'equal', None)]
    run_batch_vs_oracle(specs, 0.01, 4, 3, seed=11)


def test_batch_large_steps_redo_mid_run(gpu_lib):
    """Steps long enough that sheets move past the sliding window's
    tolerance (over a thousand ranks per step at n = 20000) after some
    outputs of short steps: the window's check fails mid-run, the step is
    redone by the full sort, and the results still equal the oracle's; then
    short steps again."""
    specs = [(20000, 'equal', None), (9000, 'lognormal', 0.1), (300, 'equal', None)]
    dt = 0.02
    run_batch_vs_oracle(specs, dt, 4, 5, hs=[0.005, 0.005, 0.3, 0.3, 0.005])


def test_batch_mixed_with_h_and_nleap_change(gpu_lib):
    """Sizes on both sides of 2048 in one launch; h changes between outputs
    and so does the parity of the step count (the current buffer set flips
    on the device)."""
    print('batch mixed', flush=True)
    specs = [(300, 'equal', None), (5000, 'lognormal', 0.1), (4096, 'lognormal', None),
             (2048, 'equal', 0.1), (2049, 'lognormal', None),
             (12345, 'equal', 0.2), (17, 'lognormal', None)]
    xs, vs, ms, oms = ensemble(specs, 5)
    b = wendy_amd.Batch(xs, vs, ms, oms)
    states = [O.ApproxState(x, v, m, 0.01, 1, omega=om)
              for x, v, m, om in zip(xs, vs, ms, oms)]
    try:
        for h, nleap in [(0.002, 5), (0.001, 4), (0.003, 3), (0.002, 2)]:
            b.advance(h, nleap)
            gx, gv, _ = b.read()
            for j, st in enumerate(states):
                st.h = h
                st.nleap = nleap
                wx, wv = st.onestep()
                tol = 0. if exact_for(ms[j]) else TOL
                assert rel_inf(gx[j], wx) <= tol, (j, h, nleap)
                assert rel_inf(gv[j], wv) <= tol, (j, h, nleap)
    finally:
        b.close()


def test_batch_member_equals_single_simulation(gpu_lib):
    """A simulation alone (wendy_amd.Sim) and inside a batch give the same
    bits (power-of-two equal masses: both are exact)."""
    print('batch vs single', flush=True)
    specs = [(2048, 'equal', 0.1), (9000, 'equal', None), (50, 'equal', None)]
    xs, vs, ms, oms = ensemble(specs, 21)
    b = wendy_amd.Batch(xs, vs, ms, oms)
    sims = [wendy_amd.Sim(x, v, m, numpy.sum(m), -1. if om is None else om ** 2.)
            for x, v, m, om in zip(xs, vs, ms, oms)]
    try:
        for _ in range(3):
            b.advance(0.002, 5)
            gx, gv, _ = b.read()
            for j, s in enumerate(sims):
                s.advance(0.002, 5)
                sx = numpy.empty(len(xs[j]))
                sv = numpy.empty(len(xs[j]))
                s.read(sx, sv)
                assert bits_equal(gx[j], sx) and bits_equal(gv[j], sv), j
    finally:
        b.close()
        for s in sims:
            s.close()


def test_batch_many_small_simulations(gpu_lib):
    """A 1000-simulation ensemble of uneven small sizes (one workgroup each)."""
    print('batch many', flush=True)
    rng = numpy.random.default_rng(3)
    sizes = rng.integers(1, 600, 1000)
    specs = [(int(n), 'lognormal' if k % 3 else 'equal', 0.1 if k % 2 else None)
             for k, n in enumerate(sizes)]
    run_batch_vs_oracle(specs, 0.01, 3, 2, seed=100)
