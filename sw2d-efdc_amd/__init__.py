"""sw2d-efdc_amd: MI355X-native SW2D overland-flow time step (HIP, gfx950).

The compute lives in libsw2d.so (csrc/, C ABI in include/sw2d.h); this package
holds its ctypes binding (sw2d.py), the build script (build.py), synthetic
case generation (cases.py) and the reference file formats (io.py).

The directory name has a hyphen, so load it with load_package() below (or
importlib) under the module name ``sw2d_efdc_amd``.
"""
from . import sw2d  # noqa: F401
from .sw2d import Case, Sw2dError, couple_bc, device_math, lake_catchment_smooth  # noqa: F401

__all__ = ["Case", "Sw2dError", "couple_bc", "device_math", "lake_catchment_smooth", "sw2d"]
