"""Drop-in for the reference's ``src/core`` API (``src/core/__init__.py:1-51``): 2D TM and 3D.

``from waveforge_amd.core import YeeGrid, FieldSet, MurABC, FDTD2D, ...``
gives the same names as the reference's ``core`` package; stepping and
imaging run on hand-written HIP kernels for gfx950.
"""

from .grid import C0, GridMetadata, YeeGrid, compute_stable_dt
from .fields import FIELD_COMPONENTS, YEE_OFFSETS, FieldSet
from .materials import EPS0, MU0, TISSUE_LIBRARY, Material, MaterialMap, MaterialMap3D, coefficients_from_maps
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
from .boundaries import MurABC, MurABC3D
from .detectors import DetectorSet
from .imaging import das_backprojection, get_antenna_positions
from .fdtd2d import FDTD2D, SimulationDivergedError
from .fdtd3d import FDTD3D

__all__ = [
    "C0", "GridMetadata", "YeeGrid", "compute_stable_dt",
    "FIELD_COMPONENTS", "YEE_OFFSETS", "FieldSet",
    "EPS0", "MU0", "TISSUE_LIBRARY", "Material", "MaterialMap", "MaterialMap3D", "coefficients_from_maps",
    "VALID_COMPONENTS", "Waveform", "GaussianPulse", "SinusoidalSource", "RickerWavelet",
    "ModulatedGaussian", "Chirp", "UWBPulse", "GaussianDerivative",
    "PlaneSource", "PointSource", "LineSource", "SourceCollection",
    "MurABC", "MurABC3D", "DetectorSet", "FDTD2D", "FDTD3D", "SimulationDivergedError",
    "das_backprojection", "get_antenna_positions",
]
