"""Benchmark of the approximate-integration path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--workload plain24|unequal27] [--no-ensemble] [--no-serve]

Workload (the single-GPU drop-in configuration): N = 2^24 equal-mass sheets
(m = 2^-24, twopiG = 1) drawn from an isothermal sech^2 disk (z_h = 1,
sigma = 1, RandomState(0)), external harmonic omega = 0.1, dt = 0.005,
nleap = 20.  One bench step is one output interval (one ``next()`` of the
generator): nleap = 20 leapfrog steps, each a sort of the positions, the
exact force, a kick and a drift.  The state is resident in HBM before the
timed region; the timed region is K output intervals of device work bracketed
by stream synchronisations.  ``value`` is N * nleap * K / seconds
(FP64 sheet-steps/s).  The end-to-end rate through the generator loop (each
output de-sorted and copied to page-locked host arrays while the next
interval runs) is reported beside it as ``e2e_value``, in steady state: timed
from the first overlapped output on.
``--workload unequal27`` measures the large unequal-mass disk instead (N =
2^27, lognormal(0, 0.5) masses, +0.3 bulk velocity on x > 0, dt = 0.005,
nleap = 10) -- not the headline line.

Beside the headline line the default run reports, in the same JSON object,
``ensemble`` (4096 disks of N log-uniform in [2^8, 2^16] advanced together
in one launch per output interval, against looping the plain path over a
sample of the same disks) and ``serving`` (a bounded window of the Poisson
request load through wendy_amd.serve, p50/p99 completion latency).

Roofline: the dominant kernel is timed with HIP events on the simulation's
stream (an event pair around every launch) over a second timed region of the
same work, min(K, 20) output intervals right after the first; ``achieved``
is its algorithmic bytes per launch divided by that per-launch average,
against the MI355X HBM peak of 8 TB/s.  ``traffic`` is the HBM bytes per
launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE counters (committed
summary under profiles/).

CPU baseline: the reference's own C (built from its sources into oracle/_ref
by oracle/Makefile, sort='parallel', OpenMP) timed on this host on a bounded
sample of the same workload (see ``cpu_baseline.sample``).
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'FP64 sheet-steps/s @N=2^24 and batched-vs-plain; p50/p99 sim latency'
HBM_PEAK = 8.0e12

# kernel classes (include/wendy_amd.h, wdy_kernel_time)
KC_SORT_HIST, KC_SORT_PASS, KC_LEAP, KC_SCAN, KC_OTHER, KC_WARM = range(6)


def workload(n, seed=0, lognormal=None, vshift=None):
    """sech^2 disk, z_h = sigma = 1, from RandomState(seed) drawing u, then
    v, then masses: equal 1/n, or lognormal(0, s) normalised to 1; vshift
    adds a bulk velocity to the sheets with x > 0."""
    rs = numpy.random.RandomState(seed)
    u = rs.uniform(size=n)
    bad = u == 0.
    while numpy.any(bad):
        u[bad] = rs.uniform(size=int(bad.sum()))
        bad = u == 0.
    x = 2. * numpy.arctanh(2. * u - 1.)
    v = rs.normal(size=n)
    if lognormal is None:
        m = numpy.full(n, 1. / n)
    else:
        m = rs.lognormal(0., lognormal, size=n)
        m /= m.sum()
    if vshift is not None:
        v[x > 0] += vshift
    return x, v, m


WORKLOADS = {
    'plain24': dict(n=1 << 24, nleap=20, dt=0.005, omega=0.1, lognormal=None,
                    vshift=None,
                    label=('plain single-simulation drop-in: N=2^24 equal-mass '
                           'sech^2 disk (z_h=1, sigma=1), omega=0.1, dt=0.005, '
                           'nleap=20; one step = one output interval')),
    'unequal24': dict(n=1 << 24, nleap=20, dt=0.005, omega=0.1, lognormal=0.5,
                      vshift=None,
                      label=('unequal-mass disk: N=2^24, lognormal(0,0.5) masses '
                             'normalised to 1, otherwise the plain workload '
                             '(omega=0.1, dt=0.005, nleap=20)')),
    'unequal27': dict(n=1 << 27, nleap=10, dt=0.005, omega=None, lognormal=0.5,
                      vshift=0.3,
                      label=('large unequal-mass disk: N=2^27, lognormal(0,0.5) '
                             'masses normalised to 1, v+=0.3 for x>0, no omega, '
                             'dt=0.005, nleap=10; one step = one output '
                             'interval')),
}


def disk(n, rs, zh=1., sigma=1., lognormal=None):
    """sech^2 disk of n sheets from RandomState rs: x = 2 zh atanh(2u-1),
    v = sigma N(0,1), masses equal 1/n or lognormal(0, s) normalised to 1."""
    u = rs.uniform(size=n)
    bad = u == 0.
    while numpy.any(bad):
        u[bad] = rs.uniform(size=int(bad.sum()))
        bad = u == 0.
    x = 2. * zh * numpy.arctanh(2. * u - 1.)
    v = sigma * rs.normal(size=n)
    if lognormal is None:
        m = numpy.full(n, 1. / n)
    else:
        m = rs.lognormal(0., lognormal, size=n)
        m /= m.sum()
    return x, v, m


def ensemble_workload(nsims, seed=1, lo=8, hi=16):
    """The mixed-size ensemble: N_k log-uniform in [2^lo, 2^hi], equal masses
    1/N_k, z_h ~ U(0.5, 2), sigma = sqrt(z_h), omega_k ~ U(0, 0.2)."""
    rs = numpy.random.RandomState(seed)
    ns = numpy.floor(2. ** rs.uniform(lo, hi, nsims)).astype(int)
    zhs = rs.uniform(0.5, 2., nsims)
    oms = rs.uniform(0., 0.2, nsims)
    xs, vs, ms = [], [], []
    for n, zh in zip(ns, zhs):
        x, v, m = disk(int(n), rs, zh, numpy.sqrt(zh))
        xs.append(x)
        vs.append(v)
        ms.append(m)
    return xs, vs, ms, [float(o) for o in oms]


def bench_ensemble(args, log):
    """Batched ensemble vs looping the plain GPU path over the same disks.

    Batched: all simulations advance together, one launch per output
    interval; the timed region is ``ens_steps`` intervals after one warm-up.
    Plain: a sample of the same simulations (the first ``ens_sample``; sizes
    are i.i.d.) each run alone through wendy_amd.Sim, one after the other,
    with the same warm-up and step count; its rate is the sample's sheet-steps
    over the sample's summed wall time."""
    import wendy_amd
    nleap, dt = 10, 0.01
    h = dt / nleap
    K = args.ens_steps
    xs, vs, ms, oms = ensemble_workload(args.ens_sims)
    ntot = int(sum(len(x) for x in xs))
    log('ensemble: %d sims, %d sheets' % (len(xs), ntot))
    b = wendy_amd.Batch(xs, vs, ms, oms)
    b.advance(h, nleap)
    b.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        b.advance(h, nleap)
    b.sync()
    tb = time.perf_counter() - t0
    b.close()
    rate_b = ntot * nleap * K / tb
    log('ensemble: batched %.3e sheet-steps/s (%.2f ms/output)'
        % (rate_b, 1e3 * tb / K))
    S = min(args.ens_sample, len(xs))
    work, tp = 0, 0.
    for j in range(S):
        m = ms[j]
        sim = wendy_amd.Sim(xs[j], vs[j], m, numpy.sum(m), oms[j] ** 2.)
        sim.advance(h, nleap)
        sim.read()
        t0 = time.perf_counter()
        for _ in range(K):
            sim.advance(h, nleap)
        sim.read()
        tp += time.perf_counter() - t0
        sim.close()
        work += len(m) * nleap * K
    rate_p = work / tp
    log('ensemble: plain loop %.3e sheet-steps/s over %d sims' % (rate_p, S))
    return {'workload': ('%d sech^2 disks, N log-uniform [2^8, 2^16], equal '
                         'masses 1/N, z_h~U(0.5,2), sigma=sqrt(z_h), '
                         'omega~U(0,0.2), dt=0.01, nleap=10' % len(xs)),
            'sims': len(xs), 'sheets': ntot, 'steps': K,
            'batched_value': rate_b, 'batched_ms_per_step': 1e3 * tb / K,
            'plain_loop_value': rate_p, 'plain_loop_sample_sims': S,
            'speedup': rate_b / rate_p, 'unit': 'sheet-steps/s'}


def serve_workload(window_s, rate=200., seed=0, pool_log2=21):
    """Requests of the serving load: Poisson arrivals at ``rate``/s over
    ``window_s`` seconds; N log-uniform in [2^10, 2^20]; z_h ~ U(0.5, 2),
    sigma = sqrt(z_h); lognormal(0, 0.3) masses normalised to 1; omega ~
    U(0, 0.2); duration log-uniform in [100, 10000] leapfrog steps of h = 0.01.
    Initial conditions are windows of one pre-drawn sample (sech^2 unit
    positions, normal velocities, lognormal masses; 2^21 sheets) at random
    offsets, scaled per request -- cheap enough to make at arrival time."""
    from wendy_amd import serve
    rs = numpy.random.RandomState(seed)
    t, arr = 0., []
    while True:
        t += rs.exponential(1. / rate)
        if t > window_s:
            break
        arr.append(t)
    k = len(arr)
    ns = numpy.floor(2. ** rs.uniform(10, 20, k)).astype(int)
    zhs = rs.uniform(0.5, 2., k)
    oms = rs.uniform(0., 0.2, k)
    steps = numpy.floor(10. ** rs.uniform(2, 4, k)).astype(int)
    offs = rs.randint(0, 1 << 30, size=(k, 3))
    base = numpy.random.RandomState(seed + 1)
    P = 1 << pool_log2
    u = base.uniform(size=P)
    u[u == 0.] = 0.5
    xb = 2. * numpy.arctanh(2. * u - 1.)
    gb = base.normal(size=P)
    lb = base.lognormal(0., 0.3, size=P)

    def maker(n, zh, o):
        def make():
            a, b, c = (int(q) % (P - n + 1) for q in o)
            m = lb[c:c + n].copy()
            return zh * xb[a:a + n], numpy.sqrt(zh) * gb[b:b + n], m / m.sum()
        return make
    return [serve.Request(i, int(ns[i]), maker(int(ns[i]), zhs[i], offs[i]),
                          int(steps[i]), 0.01, omega=float(oms[i]), arrival=arr[i])
            for i in range(k)]


# largest request the serving pool holds (one workgroup per simulation);
# larger ones advance in gangs on the whole-GPU path.  Measured over a 3 s
# window of the serving load: split 2^14 served 2.87e10 sheet-steps/s (p50
# 1.4 s), 2^13 2.82e10 (p50 2.3 s), 2^15 2.78e10; over a 1 s window 2^16
# (the pool then bounded by its largest, one-workgroup members) served
# about half as much.
SERVE_SPLIT = 1 << 14


def bench_serve(args, log):
    """Continuous batching of the serving load on one GPU (wendy_amd.serve):
    a 200 GB HBM budget, pool of 2048 slots for N <= SERVE_SPLIT, larger
    requests on the large path in two gangs (up to 128 at once), intervals of
    10 steps."""
    from wendy_amd import serve
    reqs = serve_workload(args.serve_window, args.serve_rate)
    offered = sum(r.n * r.steps for r in reqs) / args.serve_window
    log('serve: %d requests over %.1f s (offered %.3e sheet-steps/s)'
        % (len(reqs), args.serve_window, offered))
    s = serve.Server(budget_bytes=200e9, max_sims=2048, split=SERVE_SPLIT,
                     tick=10, large='gang', ngangs=2)
    try:
        res = s.run(reqs, progress=lambda d: log('serve: progress %s' % d))
    finally:
        s.close()
    summ = serve.latency_summary(res)
    log('serve: p50 %.3f s p99 %.3f s, %.3e sheet-steps/s, %s'
        % (summ['p50_s'], summ['p99_s'], summ['sheet_steps_per_s'], s.stats))
    where = [r.where for r in res.values()]
    return {'workload': ('Poisson arrivals %.0f/s for %.1f s, N log-uniform '
                         '[2^10, 2^20], lognormal(0,0.3) masses, z_h~U(0.5,2), '
                         'sigma=sqrt(z_h), omega~U(0,0.2), 100-10000 steps '
                         'log-uniform, h=0.01, 200 GB HBM budget; pool for '
                         'N <= 2^%d, larger N in two gangs'
                         % (args.serve_rate, args.serve_window,
                            SERVE_SPLIT.bit_length() - 1)),
            'requests': len(reqs), 'offered_sheet_steps_per_s': offered,
            'p50_latency_s': summ['p50_s'], 'p99_latency_s': summ['p99_s'],
            'mean_latency_s': summ['mean_s'],
            'sheet_steps_per_s': summ['sheet_steps_per_s'],
            'pool_requests': where.count('pool'), 'large_path_requests': where.count('sim'),
            'ticks': s.stats['ticks'], 'admission_stalls': s.stats['admission_stalls'],
            'wall_s': s.stats['wall_s']}


def host_cpus():
    """(CPU model, logical CPUs of the machine, CPUs this process may run on,
    threads to use: the latter capped by OMP_NUM_THREADS when it is set)."""
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    total = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = total
    use = avail
    env = os.environ.get('OMP_NUM_THREADS', '')
    if env.isdigit() and int(env) > 0:
        use = min(use, int(env))
    return model, total, avail, use


REF_TEAMS = (1, 2, 4, 8, 16, 32, 64, 128)


def cpu_baseline(x, v, m, dt, nleap, omega, threads=None, budget_s=25.):
    """Time the reference C (sort='parallel') on a bounded sample.

    The reference fixes its parallel-sort team size at compile time
    (wendy/parallel_sort.h:13-15); oracle/Makefile builds one library per
    team size, and the one for the threads used here is loaded, with
    OMP_NUM_THREADS set to the same count.  One warm-up output (it contains
    the cold sort of the random initial order), then outputs of ``nleap``
    leapfrog steps (the workload's own) until ``budget_s`` of CPU work;
    rate = N * steps / time_elapsed, using the reference's own in-C wall
    clock (wendy/wendy.c:396,416-417)."""
    from oracle import ref_harness as R
    model, total, avail, use = host_cpus()
    if threads is not None:
        use = threads
    team = max(k for k in REF_TEAMS if k <= max(1, use))
    variant = '_t%d' % team
    if not R.available(variant):
        return None
    os.environ['OMP_NUM_THREADS'] = str(team)
    n = len(x)
    h = dt / nleap
    ref = R.RefApprox(x, v, m, h * nleap, nleap, omega=omega, sort='parallel',
                      variant=variant)
    t0 = time.time()
    ref.onestep()
    warm = time.time() - t0
    steps, secs = 0, 0.
    while secs < budget_s and (time.time() - t0) < budget_s + warm + 30:
        _, _, te = ref.onestep()
        steps += nleap
        secs += te
    if steps == 0:
        return None
    return {'value': n * steps / secs, 'unit': 'sheet-steps/s',
            'cores': team, 'kind': 'reference',
            'cpu_model': model, 'nproc': total, 'cpus_available': avail,
            'sample': ('N=%d sheets of this workload, reference C '
                       "_wendy_nbody_approx_onestep sort='parallel' "
                       '(OpenMP, PARALLEL_SORT_NUM_THREADS=%d, '
                       'OMP_NUM_THREADS=%d) on %d of the %d CPUs this process '
                       'may use (%s, %d logical CPUs in the machine): %d '
                       'outputs of nleap=%d leapfrog steps after one discarded '
                       "cold-sort output, %.1f s timed by the reference's own "
                       'wall clock' % (n, team, team, team, avail, model, total,
                                       steps // nleap, nleap, secs))}


def measured_traffic(name):
    """HBM bytes per launch recorded in a committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size); None
    when the summary is absent."""
    path = os.path.join(ROOT, 'profiles', name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch:'):
                    return int(line.split(':')[1]), 'profiles/' + name
    except OSError:
        pass
    return None, None


def copy_probe(mib=1024, launches=5):
    """A plain device copy of known size (libwendy_amd's wdy_copy_probe: read
    mib MiB, write mib MiB per launch) whose HBM counters calibrate the
    FETCH_SIZE / WRITE_SIZE readings of the kernels above."""
    import ctypes
    from wendy_amd import _lib
    ms = ctypes.c_double(0.)
    nbytes = mib * (1 << 20)
    _lib.check(_lib.lib().wdy_copy_probe(nbytes, launches, ctypes.byref(ms)))
    print(json.dumps({'copy_probe_read_bytes': nbytes,
                      'copy_probe_write_bytes': nbytes, 'launches': launches,
                      'ms_per_launch': ms.value,
                      'GB_per_s': 2 * nbytes / (ms.value * 1e-3) / 1e9}))


def bench_plain(name, steps, warmup, log, e2e=True, warm=True):
    """One simulation of workload ``name`` on the GPU: ``warmup`` output
    intervals, then ``steps`` timed intervals of device work bracketed by
    synchronisations (the rate), a second timed region with an event pair
    around every launch (the dominant kernel's time per launch), and with
    ``e2e`` the generator's loop in steady state.  ``warm=False`` keeps the
    simulation on the radix path (a full sort every step)."""
    import wendy_amd
    from wendy_amd import _lib as wlib
    wl = WORKLOADS[name]
    n, nleap, dt, omega = wl['n'], wl['nleap'], wl['dt'], wl['omega']
    x, v, m = workload(n, lognormal=wl['lognormal'], vshift=wl['vshift'])
    equal = wl['lognormal'] is None
    h = dt / nleap
    log('bench %s: n=%d nleap=%d steps=%d warmup=%d' % (name, n, nleap, steps,
                                                        warmup))
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1. if omega is None else omega ** 2.,
                        warm=warm)
    del x, v, m
    for _ in range(warmup):
        sim.advance(h, nleap)
    sim.read()
    log('bench %s: warmup done' % name)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.advance(h, nleap)
    sim.read()  # synchronise
    elapsed = time.perf_counter() - t0
    rate = n * nleap * steps / elapsed
    log('bench %s: device-only %.3e sheet-steps/s (%.2f ms/output)'
        % (name, rate, 1e3 * elapsed / steps))
    # per-kernel device time: a second timed region of the same work with a
    # HIP event pair around every launch on the simulation's stream (the
    # events themselves add a little time between launches, so the rate
    # above is taken without them)
    kb = max(1, min(steps, 20))
    sim.set_timing(True)
    for _ in range(kb):
        sim.advance(h, nleap)
    sim.read()
    sim.set_timing(False)
    kt = {c: sim.kernel_time(c) for c in range(6)}
    ps = sim.path_stats()
    paths = {'warm_steps': ps[0], 'cold_steps': ps[1], 'redos': ps[2],
             'failures_capacity_overflow_stall': list(sim.warm_failures())}
    shape, target, window = sim.warm_config()
    log('bench %s: paths %s, warm shape %d target %d window %d'
        % (name, paths, shape, target, window))
    e2e_rate = None
    if e2e:
        # end-to-end as the nbody generator runs it: each output is de-sorted
        # and copied to page-locked host arrays while the next interval runs
        xo = wlib.pinned_empty(n)
        vo = wlib.pinned_empty(n)
        ne = max(1, min(10, steps))
        sim.read()
        sim.advance(h, nleap)
        sim.read_and_advance(xo, vo, h, nleap)  # the pipeline is full from here
        t0 = time.perf_counter()
        for _ in range(ne):
            sim.read_and_advance(xo, vo, h, nleap)
        e2e_rate = n * nleap * ne / (time.perf_counter() - t0)
        log('bench %s: end-to-end %.3e sheet-steps/s' % (name, e2e_rate))
    sim.close()

    # dominant kernel and its algorithmic bytes per launch
    busy = {c: kt[c][0] for c in kt}
    dom = max(busy, key=busy.get)
    ms, cnt = kt[dom]
    if dom == KC_SORT_PASS:
        kname, bpl = 'radix scatter pass', 24 * n
        note = 'read key 8 B + payload 4 B, write key 8 B + payload 4 B per sheet'
    elif dom == KC_WARM:
        per = 40 if equal else 56
        kname, bpl = 'warm-start bucket sort + leap', per * n
        note = ('read key, v, idx%s; write the same (%d B per sheet)'
                % ('' if equal else ', mass', per))
    elif dom == KC_LEAP:
        kname, bpl = 'fused force+kick+drift', 44 * n
        note = 'read key, perm, v, idx; write key, v, idx'
    else:
        kname, bpl = 'class %d' % dom, 16 * n
        note = 'approximate'
    avg_s = ms * 1e-3 / max(cnt, 1)
    achieved = bpl / avg_s if avg_s > 0 else 0.
    traffic, traffic_src = None, None
    if dom == KC_WARM:
        traffic, traffic_src = measured_traffic('warm_pipe_n%d_hbm_traffic.md' % n)
    roofline = {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9,
                'unit': 'GB/s', 'frac': achieved / HBM_PEAK, 'traffic': traffic,
                'traffic_source': traffic_src,
                'kernel': kname, 'bytes_per_launch': bpl,
                'avg_launch_us': avg_s * 1e6, 'launches': cnt,
                'bytes_note': note,
                # whole-step rate against the full-radix-resort model
                # (236 / 272 B per sheet-step) and the compulsory state
                # traffic (40 / 48 B), equal / unequal masses
                'step_frac_Balg': rate * (236 if equal else 272) / HBM_PEAK,
                'step_frac_Bmin': rate * (40 if equal else 48) / HBM_PEAK}
    return {'workload': wl['label'], 'n': n, 'value': rate,
            'unit': 'sheet-steps/s', 'ms_per_step': 1e3 * elapsed / steps,
            'steps': steps, 'warmup': warmup, 'elapsed': elapsed,
            'e2e_value': e2e_rate,
            'kernel_ms': {str(c): kt[c][0] for c in kt}, 'paths': paths,
            'warm_config': {'shape': shape, 'bucket_target': target,
                            'window': window},
            'roofline': roofline}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', choices=sorted(WORKLOADS), default='plain24')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-threads', type=int, default=None,
                    help='threads for the CPU baseline (default: the CPUs '
                         'this process may use, capped by OMP_NUM_THREADS)')
    ap.add_argument('--no-ensemble', action='store_true')
    ap.add_argument('--ens-sims', type=int, default=4096)
    ap.add_argument('--ens-steps', type=int, default=10)
    ap.add_argument('--ens-sample', type=int, default=256)
    ap.add_argument('--no-serve', action='store_true')
    ap.add_argument('--copy-probe', action='store_true',
                    help='only run the calibration copy kernel')
    ap.add_argument('--serve-window', type=float, default=3.)
    ap.add_argument('--serve-rate', type=float, default=200.)
    ap.add_argument('--no-unequal27', action='store_true')
    ap.add_argument('--no-warm', action='store_true',
                    help='radix path only (a full sort every step), for '
                         'comparison and profiling; not the headline')
    ap.add_argument('--unequal27-steps', type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; --gpus %d is not '
              'supported (multi-GPU sharding is design-only)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe()
        return
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    res = bench_plain(args.workload, args.steps, args.warmup, log, e2e=True,
                      warm=not args.no_warm)
    wl = WORKLOADS[args.workload]
    n, nleap, dt, omega = wl['n'], wl['nleap'], wl['dt'], wl['omega']
    rate, elapsed = res['value'], res['elapsed']
    main_line = args.workload == 'plain24' and not args.no_warm
    ens = None if args.no_ensemble or not main_line else bench_ensemble(args, log)
    srv = None if args.no_serve or not main_line else bench_serve(args, log)

    # the large unequal-mass disk (N = 2^27), time-bounded, beside the line
    big = None
    if main_line and not args.no_unequal27:
        big = bench_plain('unequal27', args.unequal27_steps, 2, log, e2e=False)
        big.pop('elapsed', None)

    cpu = None
    if not args.no_cpu_baseline and main_line:
        log('bench: cpu baseline ...')
        x, v, m = workload(n, lognormal=wl['lognormal'], vshift=wl['vshift'])
        cpu = cpu_baseline(x, v, m, dt, nleap, omega, args.cpu_threads)
        if cpu:
            log('bench: cpu baseline %.3e sheet-steps/s' % cpu['value'])

    out = {
        'metric': METRIC, 'value': rate, 'unit': 'sheet-steps/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': wl['label'] + (' [radix path only]' if args.no_warm else ''),
                   'n': n, 'nleap': nleap, 'dt': dt, 'omega': omega},
        'e2e_value': res['e2e_value'],
        'kernel_ms': res['kernel_ms'],
        'paths': res['paths'],
        'warm_config': res['warm_config'],
        'roofline': res['roofline'],
        'cpu_baseline': cpu,
        'unequal27': big,
        'ensemble': ens,
        'serving': srv,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
