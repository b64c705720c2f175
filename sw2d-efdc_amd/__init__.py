"""sw2d-efdc_amd: MI355X-native SW2D overland-flow time step (HIP, gfx950).

The compute lives in libsw2d.so (csrc/, C ABI in include/sw2d.h); this package
holds its ctypes binding (sw2d.py), the build script (build.py) and synthetic
case generation (cases.py).  The sw2d-efdc driver executable (csrc/
sw2d_driver.cpp, built next to the library) reads and writes the reference's
db/ case files.

The directory name has a hyphen, so load it with importlib under the module
name ``sw2d_efdc_amd``.
"""
from . import sw2d  # noqa: F401
from .sw2d import (Batch, Case, Server, Sw2dError, couple_bc, device_math,  # noqa: F401
                   footprint, lake_catchment_smooth)

__all__ = ["Batch", "Case", "Server", "Sw2dError", "couple_bc", "device_math", "footprint",
           "lake_catchment_smooth", "sw2d"]
