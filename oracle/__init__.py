"""CPU oracle for the approximate (leapfrog) integration path of wendy.

TEST INFRASTRUCTURE ONLY.  Nothing in this package is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the CPU baseline,
never as the thing measured or shipped.  The product (``wendy_amd``) never
imports it and has no CPU fallback.

Contents
--------
``nbody_oracle``  numpy restatement of the reference's approx step
                  (wendy/wendy.py:336-437, wendy/wendy.c:318-418), its energy
                  and momentum diagnostics (wendy/wendy.py:439-517), plus
                  closed forms used to pin it.
``ref_harness``   ctypes drivers for the reference C library built from the
                  reference's own sources by ``oracle/Makefile`` into
                  ``oracle/_ref/`` (present only where it was built): the
                  approx step (``RefApprox``, with an optional external force
                  passed through a ctypes callback as wendy/wendy.py:406-418
                  does) and the exact event-driven solver (``RefExact``).

Pinning status of every oracle function (what checks it, so that a dropped
term, a wrong sign or index, or a transposed operand fails a test):

* ``approx_onestep`` / ``ApproxState`` -- bitwise against the compiled
  reference (``oracle/_ref``) on tie-free inputs, against the committed golden
  fixtures in ``tests/golden/`` (written from the compiled reference), and
  against closed forms: the two-sheet free-fall parabola (constant force,
  where drift-kick-drift is exact) and the harmonic rotation map of tracer
  sheets.
* ``ext_force_sorted`` and the external-force clock of ``ApproxState`` --
  bitwise against the compiled reference driven with the same callable, on
  both sides of EXTERNAL_SWITCH, clock included; against the closed form of
  massless sheets in an external harmonic force (equal to the built-in omega
  term); and against the momentum change a uniform time-dependent force must
  produce, M h sum_k g(t0 + k h), which a half-step clock error fails.
* ``canonical_order`` -- against brute-force ranking of tiny inputs.
* ``canonical_order_presorted`` -- equal to ``canonical_order`` on tie-free,
  tied (-0.0 vs +0.0 included), random and nearly sorted inputs, and a whole
  ``ApproxState`` trajectory is bitwise the same with either.
* ``force_sorted`` -- against the brute-force pair sum
  ``sum_j m_j sign(x_j - x_i)`` with ``math.fsum``.
* ``energy`` / ``momentum`` / ``potential`` / ``energy_individual`` --
  against brute-force pair sums with ``math.fsum`` and the reference's
  hand-expanded closed forms (tests/test_wendy.py:81-104).

No function is left "parity unpinned" except the order of *exactly tied*
positions: the reference's tie order is whatever its chosen comparison sort
(quick, merge, tim, qsort, parallel) happens to produce, which no reference
test pins; the oracle and the product both break ties by original sheet index.
"""
