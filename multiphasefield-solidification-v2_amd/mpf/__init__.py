"""MI355X multi-phase-field solidification solver.

The compute path is libmpf.so (hand-written HIP for gfx950) reached through
the C ABI in include/mpf.h; this package is its host-side front end.
"""
from ._lib import MPFError, CapacityError, lib  # noqa: F401
from .solver import Solver, KMaxOverflowError  # noqa: F401
from .batch import Batch  # noqa: F401

KMAX = 50
LIQ = 0
