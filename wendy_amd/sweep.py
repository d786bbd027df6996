"""Measured check of the fixed tuning choices across simulation sizes.

``python -m wendy_amd.sweep`` times one output interval of the sech^2 disk
(z_h = sigma = 1, omega = 0.1, dt = 0.005, nleap = 20; equal masses and
lognormal(0, 0.5) masses) at each N of ``--sizes``:

* the automatic choice: the path chosen by size, and on the large path the
  tile shape, bucket target and window the window fitting picks;
* every alternative: the other path where both apply (N up to the
  one-workgroup path's limit), and on the large path each other tile shape
  in ``--shapes`` with the window that reaches as many ranks as the
  automatic one (an interval that a shape had to redo is not a fit).

It prints one markdown row per (N, masses): the automatic choice's time, the
best alternative and the automatic choice's margin over it.  Each time is
the median over ``--trials`` intervals after one discarded warm interval.
"""
import argparse
import math
import sys
import time

import numpy

from . import _lib
from .nbody import Sim

def disk(n, seed, lognormal):
    rs = numpy.random.RandomState(seed)
    u = rs.uniform(size=n)
    u[u == 0.] = 0.5
    x = numpy.arctanh(2. * u - 1.)
    v = rs.normal(size=n)
    if lognormal:
        m = rs.lognormal(0., 0.5, size=n)
        m /= m.sum()
    else:
        m = numpy.full(n, 1. / n)
    return x, v, m


def time_sim(x, v, m, omega2, h, nleap, trials, path=None, tune=None):
    """Median seconds per interval (inf if the warm path had to redo one),
    and the warm configuration used."""
    sim = Sim(x, v, m, numpy.sum(m), omega2, path=path)
    try:
        if tune is not None:
            sim.tune(*tune)
        sim.advance(h, nleap)    # first (radix) interval
        sim.advance(h, nleap)    # first warm interval, discarded
        sim.read()
        ts = []
        for _ in range(trials):
            t0 = time.perf_counter()
            sim.advance(h, nleap)
            sim.read()
            ts.append(time.perf_counter() - t0)
        redo = sim.path_stats()[2]
        cfg = sim.warm_config()
        return (float('inf') if redo else float(numpy.median(ts))), cfg
    finally:
        sim.close()


def shape_cap(shape):
    return {0: 4096, 1: 2048, 2: 1024, 3: 1024, 4: 2048, 5: 3072, 6: 3072,
            7: 2048, 8: 1536, 9: 1792}[shape]


def target(cap):
    room = max(cap // 5, 256)
    return cap - room if cap > 2 * room else cap // 2


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--sizes', default='8,10,12,14,16,18,20,22,24',
                    help='log2 N values')
    ap.add_argument('--shapes', default='8,9,7,2,0')
    ap.add_argument('--trials', type=int, default=3)
    ap.add_argument('--all', action='store_true',
                    help='also print every alternative\'s time')
    args = ap.parse_args(argv)
    dt, nleap, omega2 = 0.005, 20, 0.01
    h = dt / nleap
    print('| N | masses | automatic choice | its time (ms) | best alternative | '
          'its time (ms) | automatic / best |')
    print('|---|---|---|---|---|---|---|')
    for lg in [int(q) for q in args.sizes.split(',')]:
        n = 1 << lg
        for lognormal in (False, True):
            x, v, m = disk(n, lg, lognormal)
            auto, cfg = time_sim(x, v, m, omega2, h, nleap, args.trials)
            small = cfg[0] < 0   # the size picked the one-workgroup path
            alts = {}
            if small:
                alts['large path'] = time_sim(x, v, m, omega2, h, nleap,
                                              args.trials, path='large')[0]
                label = 'small path'
            else:
                shape, T, W = cfg
                label = 'shape %d, target %d, window %d' % (shape, T, W)
                reach = max(1, T * W)
                for s in [int(q) for q in args.shapes.split(',')]:
                    if s == shape:
                        continue
                    Ts = target(shape_cap(s))
                    Ws = max(1, math.ceil(reach / Ts))
                    try:
                        alts['shape %d, window %d' % (s, Ws)] = time_sim(
                            x, v, m, omega2, h, nleap, args.trials,
                            tune=(0, Ws, s))[0]
                    except (_lib.WendyError, ValueError):
                        pass   # window beyond the shape's limit
            if args.all:
                print('<!-- 2^%d %s: automatic %.3f ms; %s -->' % (
                    lg, 'lognormal' if lognormal else 'equal', 1e3 * auto,
                    ', '.join('%s %.3f' % (k, 1e3 * v) for k, v in alts.items())),
                    flush=True)
            best = min(alts, key=alts.get) if alts else None
            bt = alts[best] if best else float('nan')
            print('| 2^%d | %s | %s | %.3f | %s | %.3f | %.3f |'
                  % (lg, 'lognormal' if lognormal else 'equal', label, 1e3 * auto,
                     best, 1e3 * bt, auto / bt), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
