"""MI355X-native isothermal-MHD cell-update path (VL2 + CT with HLLD/HLLE/Roe, K98
Roe-TVD/HLLE + CTU + CT, passive scalars, CFL reduction, ghost fill) behind the
reference's procedure interface.

The compute path is libimhd.so (hand-written HIP for gfx950, C ABI in include/imhd.h);
this package only binds it.
"""
from ._lib import ImhdError  # noqa: F401
from .solver import TvdSolver, VanLeerSolver, driver_step, padded_shapes  # noqa: F401

__all__ = ["VanLeerSolver", "TvdSolver", "driver_step", "padded_shapes", "ImhdError"]
