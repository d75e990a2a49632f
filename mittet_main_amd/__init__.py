"""MI355X-native fictitious-wave-domain FDTD (Mittet) forward/adjoint solver.

Python host orchestration over ``libmittet_fdtd.so`` (hand-written HIP for
gfx950, C ABI in ``include/mittet_fdtd.h``), reached through ctypes.  The
library must be built (``python -m mittet_main_amd.build``); nothing falls
back to a CPU path.
"""
from .fwi3d import (  # noqa: F401
    Batch, Grid, Model, Shot, f_to_d_h, firstderiv_gauss, from_1based, lattice_time_3d,
    shot_bytes, source_spectrum,
)
from ._lib import Fwi3dError  # noqa: F401
