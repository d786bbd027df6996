"""numpy restatement of wendy's approximate-integration step (test oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Every function follows the reference's operation order so that, on inputs
without exactly tied positions, results equal the compiled reference bit for
bit:

* multiplications and additions are separate roundings (the reference is
  built for x86-64 without ``-march``, so gcc emits no FMA);
* the force is ``((totmass - 2*C_i) - m_i) - omega2*x_i`` added to a zeroed
  acceleration, i.e. ``0.0 + F`` (wendy/wendy.c:371-383);
* the exclusive mass prefix ``C`` is a strictly sequential sum in sorted order
  (wendy/wendy.c:359-360); ``numpy.cumsum`` on float64 is a sequential loop;
* ``totmass = numpy.sum(m)`` is numpy's pairwise sum (wendy/wendy.py:383).

Exactly tied positions are ordered by original sheet index (``lexsort``);
the reference leaves their order to its comparison sort.
"""
import copy
import math

import numpy

# wendy/wendy.py:102
SORT_TYPES = {'quick': 0, 'merge': 1, 'tim': 2, 'qsort': 3, 'parallel': 4}


def canonical_order(val, idx):
    """Permutation sorting positions ascending, ties by original index.

    The reference comparator (wendy/wendy.c:19-31) compares ``val`` only, so
    -0.0 and +0.0 compare equal; ``lexsort`` compares with ``<`` too.
    """
    return numpy.lexsort((idx, val))


def canonical_order_presorted(val, idx):
    """``canonical_order`` for input close to sorted (the previous step's
    order, as in every step after the first): a stable argsort on ``val``,
    which numpy runs as timsort, near-linear on long sorted runs.  Without
    exactly equal positions the order by ``val`` alone is the (val, idx)
    order; if any two positions compare equal (-0.0 and +0.0 included) it
    defers to ``canonical_order``."""
    order = numpy.argsort(val, kind='stable')
    sv = val[order]
    if len(sv) > 1 and numpy.any(sv[1:] == sv[:-1]):
        return canonical_order(val, idx)
    return order


def exclusive_mass_prefix(m_sorted):
    """cumulmass[0] = 0, cumulmass[i+1] = cumulmass[i] + m_sorted[i]
    (wendy/wendy.c:359-360; cumulmass[0] is never written, wendy/wendy.py:372)."""
    out = numpy.zeros_like(m_sorted)
    if len(m_sorted) > 1:
        out[1:] = numpy.cumsum(m_sorted[:-1])
    return out


def force_sorted(val, m_sorted, totmass, omega2, ext=None):
    """Acceleration of each sheet in sorted order (wendy/wendy.c:371-383).

    ``m_sorted`` is already multiplied by twopiG; ``omega2 < 0`` means no
    harmonic term (the -1 sentinel of wendy/wendy.py:363-366).  ``ext``: the
    external force in sorted order, which the reference stores in a[] before
    adding the self-gravity term (wendy/wendy.c:362-370); else a[] is zeroed.
    """
    cumul = exclusive_mass_prefix(m_sorted)
    f = totmass - 2. * cumul - m_sorted
    if omega2 >= 0:
        f = f - omega2 * val
    return (0.0 if ext is None else ext) + f


EXTERNAL_SWITCH = 10  # wendy/wendy.h:9-11


def ext_force_sorted(ext_force, val, idx, x, t):
    """The external force at the sorted positions, evaluated the way
    wendy/wendy.c:362-370 does with the ctypes wrapper of wendy/wendy.py:
    406-418: for N > EXTERNAL_SWITCH the positions are de-sorted into x and
    the callable is evaluated once on the array; otherwise once per sheet on
    the scalar position, in sorted order."""
    n = len(val)
    if n > EXTERNAL_SWITCH:
        x[idx] = val
        a = numpy.empty(n)
        a[:] = ext_force(x, t)
        return a[idx]
    return numpy.array([ext_force(float(val[i]), t) for i in range(n)])


class ApproxState:
    """State of one approx=True simulation, laid out like the reference.

    ``val``/``idx`` mirror ``struct array_w_index xi[N]`` (wendy/wendy.h:12-16),
    kept in sorted order between calls; ``v`` and ``m`` stay in original sheet
    order and are reached through ``idx``, as in wendy/wendy.c:318-333.
    Setup restates wendy/wendy.py:363-387 and :422.
    """

    def __init__(self, x, v, m, dt, nleap, omega=None, twopiG=1.,
                 ext_force=None, t0=0., presorted_order=False):
        self.omega2 = -1. if omega is None else omega ** 2.
        # canonical_order_presorted instead of canonical_order (same result;
        # fast on the nearly sorted input of large runs)
        self.order_fn = canonical_order_presorted if presorted_order \
            else canonical_order
        self.ext_force = ext_force
        self.t = float(t0)  # the *t0 clock (wendy/wendy.c:401-402,407-408)
        n = len(x)
        self.x = numpy.require(copy.copy(x), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.v = numpy.require(copy.copy(v), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.m = numpy.require(twopiG * copy.copy(m), dtype=numpy.float64,
                               requirements=['C', 'W'])
        self.totmass = numpy.sum(self.m)
        self.val = self.x.copy()
        self.idx = numpy.arange(n, dtype=numpy.int32)
        self.nleap = int(nleap)
        self.h = dt / nleap
        self.n = n

    def _force_and_leap(self, dt_kick, dt_drift):
        # _nbody_force: sort then force (wendy/wendy.c:334-384)
        order = self.order_fn(self.val, self.idx)
        self.val = self.val[order]
        self.idx = self.idx[order]
        ext = None
        if self.ext_force is not None:
            ext = ext_force_sorted(self.ext_force, self.val, self.idx, self.x,
                                   self.t)
            self.t = self.t + dt_kick
        a_sorted = force_sorted(self.val, self.m[self.idx], self.totmass,
                                self.omega2, ext)
        # leapfrog_leappq (wendy/wendy.c:324-333)
        vs = self.v[self.idx] + dt_kick * a_sorted
        self.v[self.idx] = vs
        self.val = self.val + dt_drift * vs

    def onestep(self):
        """One output interval: _wendy_nbody_approx_onestep (wendy/wendy.c:385-418)."""
        h = self.h
        # leapfrog_leapq with dt/2. (wendy/wendy.c:318-323,398)
        self.val = self.val + (h / 2.) * self.v[self.idx]
        for _ in range(self.nleap - 1):
            self._force_and_leap(h, h)
        self._force_and_leap(h, h / 2.)
        # de-sort (wendy/wendy.c:413-415)
        self.x[self.idx] = self.val
        return self.x, self.v


def nbody_approx(x, v, m, dt, nleap, omega=None, twopiG=1.):
    """Generator mirroring wendy.nbody(..., approx=True) (wendy/wendy.py:112-157,336-437)."""
    st = ApproxState(x, v, m, dt, nleap, omega=omega, twopiG=twopiG)
    while True:
        xo, vo = st.onestep()
        yield xo.copy(), vo.copy()


def energy(x, v, m, twopiG=1., individual=False, omega=None):
    """Restatement of wendy.energy (wendy/wendy.py:439-475)."""
    if omega is not None:
        out = m * omega ** 2. * x ** 2. / 2.
    else:
        out = 0.
    if individual:
        return out + twopiG * m * numpy.sum(
            m * numpy.fabs(x - numpy.atleast_2d(x).T), axis=1) + m * v ** 2. / 2.
    sindx = numpy.argsort(x)
    mass_below = numpy.roll(numpy.cumsum(m[sindx]), 1)
    mass_below[0] = 0.
    xmass_below = numpy.roll(numpy.cumsum((m * x)[sindx]), 1)
    xmass_below[0] = 0.
    return numpy.sum(out) \
        + twopiG * numpy.sum(m[sindx] * (mass_below * x[sindx] - xmass_below)) \
        + numpy.sum(m * v ** 2. / 2.)


def momentum(v, m):
    """wendy.momentum (wendy/wendy.py:477-491)."""
    return numpy.sum(m * v)


def potential(y, x, v, m, twopiG=1., omega=None):
    """wendy.potential (wendy/wendy.py:493-517)."""
    if omega is not None:
        out = omega ** 2. * y ** 2. / 2.
    else:
        out = 0.
    return out + twopiG * numpy.sum(m * numpy.fabs(x - numpy.atleast_2d(y).T),
                                    axis=1)


# ----------------------------------------------------------------------------
# Closed forms and brute-force checks (independent of the restatement above)
# ----------------------------------------------------------------------------

def brute_force_accel(x, m, twopiG=1., omega=None):
    """a_i = twopiG * sum_j m_j sign(x_j - x_i) - omega^2 x_i, fsum per sheet.

    Independent of sorting: the 1-D force is the mass above minus the mass
    below.  Sheets exactly tied with i are split by index (lower index below),
    matching the canonical order.
    """
    n = len(x)
    out = numpy.empty(n)
    for i in range(n):
        terms = []
        for j in range(n):
            if j == i:
                continue
            above = x[j] > x[i] or (x[j] == x[i] and j > i)
            terms.append(m[j] if above else -m[j])
        a = twopiG * math.fsum(terms)
        if omega is not None:
            a -= omega ** 2 * x[i]
        out[i] = a
    return out


def brute_force_energy(x, v, m, twopiG=1., omega=None):
    """E = sum m v^2/2 + twopiG sum_{i<j} m_i m_j |x_i-x_j| (+ sum m w^2 x^2/2)."""
    n = len(x)
    terms = [0.5 * m[i] * v[i] ** 2 for i in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            terms.append(twopiG * m[i] * m[j] * abs(x[i] - x[j]))
    if omega is not None:
        terms += [0.5 * m[i] * omega ** 2 * x[i] ** 2 for i in range(n)]
    return math.fsum(terms)


def sech2_disk(n, seed=0, zh=1., sigma=1., mass='equal', vshift=None):
    """Isothermal sech^2 disk: x = 2 zh atanh(2u-1), v = sigma N(0,1).

    Draws u, then v, then masses from ``numpy.random.RandomState(seed)``;
    exact u == 0.0 draws are resampled.  ``mass`` is 'equal' (1/n) or
    ('lognormal', s) for lognormal(0, s) masses normalised to total 1.
    ``vshift`` adds a bulk velocity to sheets with x > 0.
    """
    rs = numpy.random.RandomState(seed)
    u = rs.uniform(size=n)
    bad = u == 0.
    while numpy.any(bad):
        u[bad] = rs.uniform(size=int(bad.sum()))
        bad = u == 0.
    x = 2. * zh * numpy.arctanh(2. * u - 1.)
    v = sigma * rs.normal(size=n)
    if mass == 'equal':
        m = numpy.full(n, 1. / n)
    else:
        kind, s = mass
        assert kind == 'lognormal'
        m = rs.lognormal(0., s, size=n)
        m /= m.sum()
    if vshift is not None:
        v[x > 0] += vshift
    return x, v, m
