"""niftyreg_amd: NiftyReg's F3D objective/gradient step and Aladin block matching on MI355X.

The compute lives in libnrhip.so (hand-written HIP for gfx950, C ABI in
include/niftyreg_hip.h); this package holds its Python binding, geometry helpers and the
synthetic workload generators.
"""
from ._lib import NiftyRegError, load, declared_symbols  # noqa: F401
from .geometry import Geom, create_control_point_grid, mat44_inverse  # noqa: F401

__all__ = ["NiftyRegError", "Geom", "create_control_point_grid", "mat44_inverse", "load", "declared_symbols"]
