"""Write the golden fixtures for the approx path from the compiled reference.

Run in a container where the reference sources exist and ``make -C oracle``
has built ``oracle/_ref/libwendy_ref*.so``::

    python tests/golden/make_golden.py

Each case runs the reference's ``_wendy_nbody_approx_onestep``
(wendy/wendy.c:385-418) through ``oracle/ref_harness.py`` and stores the
inputs, the parameters and the (x, v) yielded after selected outputs in
``tests/golden/<case>.npz``.  The fixtures are data only (inputs and expected
outputs); they let the CPU test suite pin the oracle where the reference
library is not built, and let the GPU tests compare against the reference's
own numbers.
"""
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import nbody_oracle as O  # noqa: E402
from oracle import ref_harness as R  # noqa: E402


def _three(masses):
    return (numpy.array([-1.1, 0.1, 1.3]), numpy.array([3., 2., -5.]),
            numpy.array(masses, dtype=float))


def _disk101(seed):
    # the N=101 sech^2 disk of tests/test_approx.py:34-44 (own RNG stream)
    rs = numpy.random.RandomState(seed)
    n = 101
    zh = 2.
    x = numpy.arctanh(2. * rs.uniform(size=n) - 1) * zh
    v = rs.normal(size=n)
    v -= numpy.mean(v)
    m = numpy.ones_like(x) / n * (1. + 0.1 * (2. * rs.uniform(size=n) - 1))
    return x, v, m


def cases():
    out = {}
    x, v, m = _three([1., 1., 1.])
    out['three_equal'] = dict(x=x, v=v, m=m, dt=0.05, nleap=1000, omega=None,
                              sort='quick', outputs=[1, 2, 5, 10])
    x, v, m = _three([1., 2., 3.])
    out['three_unequal_harm'] = dict(x=x, v=v, m=m, dt=0.05, nleap=1000,
                                     omega=1.1, sort='quick',
                                     outputs=[1, 2, 5, 10])
    x, v, m = _disk101(11)
    out['disk101_parallel'] = dict(x=x, v=v, m=m, dt=0.05, nleap=100,
                                   omega=None, sort='parallel', variant='_small',
                                   outputs=[1, 5, 10])
    x, v, m = _disk101(12)
    m[101 // 2:] = 0.
    m *= 2.
    out['disk101_tracers'] = dict(x=x, v=v, m=m, dt=0.05, nleap=100,
                                  omega=None, sort='merge', outputs=[1, 5, 10])
    x, v, m = _disk101(13)
    out['disk101_harm_tim'] = dict(x=x, v=v, m=m, dt=0.05, nleap=100,
                                   omega=1.1, sort='tim', outputs=[1, 5, 10])
    # the CPU parity workload (N=1e4 equal masses, dt=0.01, nleap=10), seed 0
    x, v, m = O.sech2_disk(10000, seed=0)
    out['cpu_workload_n10000'] = dict(x=x, v=v, m=m, dt=0.01, nleap=10,
                                      omega=None, sort='parallel',
                                      outputs=[1, 10])
    # the single-GPU workload's shape at N=2^12: equal masses 2^-12,
    # omega=0.1, dt=0.005, nleap=20
    x, v, m = O.sech2_disk(4096, seed=0)
    out['gpu_workload_n4096'] = dict(x=x, v=v, m=m, dt=0.005, nleap=20,
                                     omega=0.1, sort='parallel',
                                     outputs=[1, 10])
    # the large unequal-mass workload's shape at N=2^12: lognormal(0, 0.5)
    # masses, bulk offset +0.3 on x > 0, dt=0.005, nleap=10
    x, v, m = O.sech2_disk(4096, seed=0, mass=('lognormal', 0.5), vshift=0.3)
    out['unequal_workload_n4096'] = dict(x=x, v=v, m=m, dt=0.005, nleap=10,
                                         omega=None, sort='qsort',
                                         outputs=[1, 10])
    return out


def main():
    for name, c in cases().items():
        ref = R.RefApprox(c['x'], c['v'], c['m'], c['dt'], c['nleap'],
                          omega=c['omega'], sort=c['sort'],
                          variant=c.get('variant', ''))
        xs, vs = [], []
        for k in range(1, max(c['outputs']) + 1):
            x, v, _ = ref.onestep()
            if k in c['outputs']:
                xs.append(x.copy())
                vs.append(v.copy())
        numpy.savez_compressed(
            os.path.join(HERE, name + '.npz'),
            x0=c['x'], v0=c['v'], m=c['m'], dt=c['dt'], nleap=c['nleap'],
            omega=numpy.nan if c['omega'] is None else c['omega'],
            twopiG=1., sort=c['sort'], outputs=numpy.array(c['outputs']),
            x_out=numpy.array(xs), v_out=numpy.array(vs))
        print('wrote', name)


if __name__ == '__main__':
    main()
