"""wendy_amd: MI355X-native approximate integration for 1-D self-gravitating sheets.

Same API as the reference package's top level (wendy/__init__.py:1, which
re-exports wendy/wendy.py): ``nbody``, ``energy``, ``momentum``, ``potential``.
The approx=True leapfrog runs on the GPU through libwendy_amd.so (C ABI,
include/wendy_amd.h); there is no CPU fallback.
"""
from . import forces  # noqa: F401
from .batch import Batch  # noqa: F401
from .nbody import Gang, Sim, energy, momentum, nbody, potential  # noqa: F401

__version__ = '0.1.0'
