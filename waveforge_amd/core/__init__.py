"""Drop-in for the reference's ``src/core`` 2D API (``src/core/__init__.py:1-51``).

``from waveforge_amd.core import YeeGrid, FieldSet, MurABC, FDTD2D, ...``
gives the same names as the reference's ``core`` package; stepping and
imaging run on hand-written HIP kernels for gfx950.
"""

from .grid import C0, GridMetadata, YeeGrid, compute_stable_dt
from .fields import FIELD_COMPONENTS, YEE_OFFSETS, FieldSet
from .materials import EPS0, MU0, TISSUE_LIBRARY, Material, MaterialMap, coefficients_from_maps
from .sources import (
    VALID_COMPONENTS,
    Chirp,
    GaussianDerivative,
    GaussianPulse,
    LineSource,
    ModulatedGaussian,
    PlaneSource,
    PointSource,
    RickerWavelet,
    SinusoidalSource,
    SourceCollection,
    UWBPulse,
    Waveform,
)
from .boundaries import MurABC
from .detectors import DetectorSet
from .imaging import das_backprojection, get_antenna_positions
from .fdtd2d import FDTD2D, SimulationDivergedError

__all__ = [
    "C0", "GridMetadata", "YeeGrid", "compute_stable_dt",
    "FIELD_COMPONENTS", "YEE_OFFSETS", "FieldSet",
    "EPS0", "MU0", "TISSUE_LIBRARY", "Material", "MaterialMap", "coefficients_from_maps",
    "VALID_COMPONENTS", "Waveform", "GaussianPulse", "SinusoidalSource", "RickerWavelet",
    "ModulatedGaussian", "Chirp", "UWBPulse", "GaussianDerivative",
    "PlaneSource", "PointSource", "LineSource", "SourceCollection",
    "MurABC", "DetectorSet", "FDTD2D", "SimulationDivergedError",
    "das_backprojection", "get_antenna_positions",
]
