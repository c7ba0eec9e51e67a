"""Dirichlet splatting on MI355X: HIP kernels for gfx950 behind the reference's API.

parameters  per-atom parameter constraints
spectral    Dirichlet atoms on a spectral grid
surfel      oriented surfels: scattering, rendering, the sqrt(area) chain
fmcw        FMCW sensor, scans and the surfel measurement model
groundtruth planar arrays and FMCW scans
dsfw        the solver loop over varpro, certificate, replacement, refinement, spatial
kernels     HIP kernels (libdsplat_hip.so, C ABI in include/dsplat_hip.h) and bindings
"""

from .certificate import CertificateSearchConfig
from .dsfw import DSFWConfig, fit, initialize, step
from .fmcw import FMCWConfig, FMCWModel, Scan
from .parameters import ParameterConfig
from .spatial import SpatialSearchConfig
from .spectral import SpectralGrid, SpectralModel

__all__ = [
    "CertificateSearchConfig",
    "DSFWConfig",
    "FMCWConfig",
    "FMCWModel",
    "ParameterConfig",
    "Scan",
    "SpatialSearchConfig",
    "SpectralGrid",
    "SpectralModel",
    "fit",
    "initialize",
    "step",
]
__version__ = "0.1.0"
