"""External forces evaluated on the GPU.

The reference takes ``ext_force`` as any Python callable ``f(x, t)`` and calls
it from C every leapfrog step (wendy/wendy.py:389-420, wendy/wendy.c:362-370).
Calling back into the host each step would take the state off the device, so
``wendy_amd.nbody(..., approx=True, ext_force=...)`` accepts the forms below.
Each is also an ordinary callable ``f(x, t)`` computing the same expression
with numpy, so the same object can be handed to the reference's ``nbody`` and
the two results compared.
"""
import numpy

KIND_NONE, KIND_HARMONIC, KIND_ISOTHERMAL, KIND_PULSE = 0, 1, 2, 3


class ExtForce:
    """Base class: ``kind`` and ``params`` are what the device evaluates."""
    kind = KIND_NONE

    def params(self):
        return (0., 0., 0.)


class Harmonic(ExtForce):
    """-omega**2 x: an external harmonic potential (the form used throughout
    the reference's tests/test_approx_external.py)."""
    kind = KIND_HARMONIC

    def __init__(self, omega):
        self.omega2 = float(omega) ** 2.

    def params(self):
        return (self.omega2, 0., 0.)

    def __call__(self, x, t):
        return -self.omega2 * x


class Isothermal(ExtForce):
    """amp * tanh(0.5 x / zh) / zh: the force of an isothermal sech^2 slab of
    scale height zh (amp = -(1 - alpha) sigma^2 gives the dark part of a disk
    whose live fraction is alpha, as in the reference's phase-spiral
    example)."""
    kind = KIND_ISOTHERMAL

    def __init__(self, amp, zh):
        self.amp = float(amp)
        self.zh = float(zh)
        if self.zh == 0.:
            raise ValueError('zh must be nonzero')

    def params(self):
        return (self.amp, self.zh, 0.)

    def __call__(self, x, t):
        return self.amp * numpy.tanh(0.5 * x / self.zh) / self.zh


class Pulse(ExtForce):
    """amp * exp(-0.5 ((t - tc) / tau)**2): a uniform vertical acceleration
    that rises and falls over a time tau around tc (the distant-tide
    approximation of a passing perturber)."""
    kind = KIND_PULSE

    def __init__(self, amp, tc, tau):
        self.amp = float(amp)
        self.tc = float(tc)
        self.tau = float(tau)
        if self.tau == 0.:
            raise ValueError('tau must be nonzero')

    def params(self):
        return (self.amp, self.tc, self.tau)

    def __call__(self, x, t):
        u = (t - self.tc) / self.tau
        return self.amp * numpy.exp(-0.5 * (u * u)) + 0. * x
