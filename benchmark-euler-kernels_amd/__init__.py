"""MI355X-native Godunov step (HLLC, perfect gas) for the 3-D Euler equations.

The compute path is the HIP library libeuler_hip.so (C ABI in include/euler_capi.h);
this package is the host-side mirror of the reference's euler_operators interface.
"""
from . import _lib  # noqa: F401  (loads the HIP library; raises if it is missing)
from .operators import *  # noqa: F401,F403
from .operators import __all__ as _ops_all
from . import simulation  # noqa: F401
from . import adjoint  # noqa: F401
from . import workloads  # noqa: F401
from . import batch  # noqa: F401

__all__ = list(_ops_all)
