"""The radix sort's look-back waits are bounded: a wait that gives up sets a
device error word, still publishes (so every kernel finishes) and the
library reports WDY_ERUNTIME instead of a result.  The condition is forced
from the host with a look-back bound of 0 (WDY_RADIX_SPIN_LIMIT=0: every
wait gives up at once); a tile that gives up only ever adds genuine counts
of its predecessors, so its writes stay inside the arrays."""
import numpy
import pytest

import wendy_amd
from conftest import bits_equal
from oracle import nbody_oracle as O
from wendy_amd import _lib

pytestmark = pytest.mark.gpu


def test_bounded_lookback_reports_instead_of_hanging(gpu_lib, monkeypatch):
    n = 1 << 16            # 16 tiles of 4096 keys per pass
    x, v, m = O.sech2_disk(n, seed=1)
    monkeypatch.setenv('WDY_RADIX_SPIN_LIMIT', '0')
    with pytest.raises(_lib.WendyError) as e:
        wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large', warm=False)
    assert 'gave up' in str(e.value)
    with pytest.raises(_lib.WendyError):
        wendy_amd.energy(x, v, m)
    monkeypatch.delenv('WDY_RADIX_SPIN_LIMIT')

    # a simulation whose later interval fails: the next call reports it, and
    # every call after that too (the state is lost)
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large', warm=False)
    try:
        sim.advance(0.001, 2)
        sim.read()
        monkeypatch.setenv('WDY_RADIX_SPIN_LIMIT', '0')
        sim.advance(0.001, 2)
        monkeypatch.delenv('WDY_RADIX_SPIN_LIMIT')
        with pytest.raises(_lib.WendyError):
            sim.read()
        with pytest.raises(_lib.WendyError):
            sim.advance(0.001, 2)
    finally:
        sim.close()

    # with the default bound the same run is the oracle's, bit for bit
    sim = wendy_amd.Sim(x, v, m, numpy.sum(m), -1., path='large', warm=False)
    st = O.ApproxState(x, v, m, 0.002, 2)
    xo = numpy.empty(n)
    vo = numpy.empty(n)
    try:
        for _ in range(2):
            sim.advance(0.001, 2)
            sim.read(xo, vo)
            wx, wv = st.onestep()
            assert bits_equal(xo, wx) and bits_equal(vo, wv)
    finally:
        sim.close()
