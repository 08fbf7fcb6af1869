"""Yee grid geometry (drop-in for the reference ``core.grid``).

Mirrors ``src/core/grid.py``: the ``C0`` constant (:23), ``GridMetadata``
(:30-113), ``compute_stable_dt`` (:120-169) and ``YeeGrid`` (:176-665) with
the same validation rules, messages and device fallback.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import torch

C0: float = 299_792_458.0  # m/s


@dataclass
class GridMetadata:
    """Grid dimensions, spacing, time step and device (``src/core/grid.py:30-113``)."""

    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    courant: float
    device: torch.device
    dtype: torch.dtype = torch.float32
    is_3d: bool = False


def compute_stable_dt(dx: float, dy: float, dz: float | None = None, courant: float = 0.98) -> float:
    """CFL time step ``courant / (C0 * sqrt(1/dx^2 + 1/dy^2 [+ 1/dz^2]))``
    (``src/core/grid.py:120-169``)."""
    if not (0.0 < courant < 1.0):
        raise ValueError(f"courant must satisfy 0 < courant < 1, got {courant!r}")
    if dx <= 0.0:
        raise ValueError(f"dx must be > 0, got {dx!r}")
    if dy <= 0.0:
        raise ValueError(f"dy must be > 0, got {dy!r}")
    if dz is not None and dz <= 0.0:
        raise ValueError(f"dz must be > 0 when provided, got {dz!r}")
    s = (1.0 / dx) ** 2 + (1.0 / dy) ** 2
    if dz is not None:
        s += (1.0 / dz) ** 2
    return courant / (C0 * math.sqrt(s))


def _resolve_device(device) -> torch.device:
    dev = torch.device(device) if isinstance(device, str) else device
    if dev.type == "cuda" and not torch.cuda.is_available():
        warnings.warn("CUDA requested but not available. Falling back to CPU.", UserWarning, stacklevel=3)
        return torch.device("cpu")
    return dev


class YeeGrid:
    """Yee-lattice geometry with CFL time step and stagger coordinates.

    Same constructor, validation and attributes as ``src/core/grid.py:176-665``.
    The HIP engine runs on the ROCm device (``"cuda"`` in PyTorch-ROCm); a
    CPU grid can be built (the reference falls back to it when no GPU is
    visible) but cannot be stepped by this package.
    """

    def __init__(
        self,
        Nx: int,
        Ny: int,
        dx: float,
        dy: float,
        *,
        Nz: int = 1,
        dz: float | None = None,
        courant: float = 0.98,
        dt: float | None = None,
        device: str | torch.device = "cuda",
        dtype: torch.dtype = torch.float32,
    ) -> None:
        for name, val in (("Nx", Nx), ("Ny", Ny)):
            if not isinstance(val, int) or val < 2:
                raise ValueError(f"{name} must be an integer >= 2, got {val!r}")
        if not isinstance(Nz, int) or Nz < 1:
            raise ValueError(f"Nz must be an integer >= 1 (use Nz=1 for 2-D mode), got {Nz!r}")
        if dx <= 0.0:
            raise ValueError(f"dx must be > 0, got {dx!r}")
        if dy <= 0.0:
            raise ValueError(f"dy must be > 0, got {dy!r}")
        is_3d = Nz > 1
        if is_3d:
            if dz is None:
                raise ValueError("dz is required when Nz > 1 (3-D mode)")
            if dz <= 0.0:
                raise ValueError(f"dz must be > 0, got {dz!r}")
        elif dz is None:
            dz = dx
        if not (0.0 < courant < 1.0):
            raise ValueError(f"courant must satisfy 0 < courant < 1, got {courant!r}")
        dev = _resolve_device(device)
        stable = compute_stable_dt(dx, dy, dz if is_3d else None, courant)
        if dt is None:
            dt = stable
        else:
            if dt > stable:
                raise ValueError(
                    f"Provided dt={dt:.6e} exceeds CFL-stable limit {stable:.6e}. Simulation will be unstable."
                )
            if dt < 0.10 * stable:
                warnings.warn(
                    f"Provided dt={dt:.6e} is less than 10 % of the CFL-stable dt={stable:.6e}. This is "
                    "suspiciously small and may indicate a unit error.",
                    UserWarning,
                    stacklevel=2,
                )
        self.meta = GridMetadata(Nx, Ny, Nz, dx, dy, dz, dt, courant, dev, dtype, is_3d)
        # E components sit at half-integer offsets, H at integer offsets.
        self.xs = self._coords(Nx, dx, 0.5)
        self.ys = self._coords(Ny, dy, 0.5)
        self.zs = self._coords(Nz, dz, 0.5)
        self.xs_h = self._coords(Nx, dx, 0.0)
        self.ys_h = self._coords(Ny, dy, 0.0)
        self.zs_h = self._coords(Nz, dz, 0.0)

    def _coords(self, n: int, spacing: float, offset: float) -> torch.Tensor:
        idx = torch.arange(n, device=self.meta.device, dtype=self.meta.dtype)
        if offset:
            idx = idx + offset
        return idx * spacing

    # --- read-only properties ------------------------------------------
    Nx = property(lambda self: self.meta.Nx, doc="Cells along x.")
    Ny = property(lambda self: self.meta.Ny, doc="Cells along y.")
    Nz = property(lambda self: self.meta.Nz, doc="Cells along z.")
    dx = property(lambda self: self.meta.dx, doc="Cell spacing along x (m).")
    dy = property(lambda self: self.meta.dy, doc="Cell spacing along y (m).")
    dz = property(lambda self: self.meta.dz, doc="Cell spacing along z (m).")
    dt = property(lambda self: self.meta.dt, doc="Time step (s).")
    courant = property(lambda self: self.meta.courant, doc="Courant number.")
    device = property(lambda self: self.meta.device, doc="Torch device of field tensors.")
    dtype = property(lambda self: self.meta.dtype, doc="Field dtype.")

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.meta.Nx, self.meta.Ny, self.meta.Nz)

    @property
    def cell_volume(self) -> float:
        return self.meta.dx * self.meta.dy * self.meta.dz

    # --- helpers ----------------------------------------------------------
    def allocate_field(self, fill: float = 0.0) -> torch.Tensor:
        return torch.full(self.shape, fill, device=self.meta.device, dtype=self.meta.dtype)

    def allocate_material_coeff(self, fill: float = 1.0) -> torch.Tensor:
        return torch.full(self.shape, fill, device=self.meta.device, dtype=self.meta.dtype)

    def cell_index(self, i: int, j: int, k: int = 0) -> int:
        return i * self.meta.Ny * self.meta.Nz + j * self.meta.Nz + k

    def physical_size(self) -> tuple[float, float, float]:
        m = self.meta
        return (m.Nx * m.dx, m.Ny * m.dy, m.Nz * m.dz)

    def wavelength_resolution(self, freq_hz: float) -> float:
        return (C0 / freq_hz) / min(self.meta.dx, self.meta.dy, self.meta.dz)

    def validate_resolution(self, freq_hz: float, min_cells_per_wavelength: float = 10) -> None:
        cpw = self.wavelength_resolution(freq_hz)
        if cpw < min_cells_per_wavelength:
            raise ValueError(
                f"Grid resolution at {freq_hz:.3e} Hz is {cpw:.2f} cells per wavelength, below the required "
                f"minimum of {min_cells_per_wavelength}. Reduce cell spacing or lower the target frequency."
            )

    def to(self, device) -> "YeeGrid":
        new = YeeGrid.__new__(YeeGrid)
        dev = _resolve_device(torch.device(device) if isinstance(device, str) else device)
        m = self.meta
        new.meta = GridMetadata(m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.dz, m.dt, m.courant, dev, m.dtype, m.is_3d)
        for name in ("xs", "ys", "zs", "xs_h", "ys_h", "zs_h"):
            setattr(new, name, getattr(self, name).to(dev))
        return new

    def __repr__(self) -> str:
        m = self.meta
        Lx, Ly, Lz = self.physical_size()
        return (
            f"YeeGrid(mode={'3D' if m.is_3d else '2D'}, shape={self.shape}, "
            f"physical=({Lx:.3g} m, {Ly:.3g} m, {Lz:.3g} m), "
            f"spacing=(dx={m.dx:.3g}, dy={m.dy:.3g}, dz={m.dz:.3g}) m, dt={m.dt:.6e} s, "
            f"courant={m.courant}, device={m.device}, dtype={m.dtype})"
        )
