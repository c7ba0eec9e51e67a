"""Dirichlet splatting on MI355X: HIP kernels for gfx950 behind the reference's API.

fmcw        FMCW sensor, scans and the surfel measurement model
surfel      oriented surfels: scattering, rendering, the sqrt(area) chain
kernels     HIP kernels (libdsplat_hip.so, C ABI in include/dsplat_hip.h) and bindings
"""

from .fmcw import FMCWConfig, Scan

__all__ = ["FMCWConfig", "Scan"]
__version__ = "0.1.0"
