"""Per-cell material coefficients (drop-in for the reference ``core.materials``, 2D part).

Mirrors ``src/core/materials.py``: ``Material`` with the semi-implicit lossy
coefficients ``Ca``/``Cb`` (:61-82), ``TISSUE_LIBRARY`` (:101-108) and
``MaterialMap`` (:115-274: painter's-order regions, masks evaluated on CPU
fp32 index grids, coefficients computed in Python fp64 and stored fp32, the
``Ca <= 0`` error and ``Ca < 0.5`` warning). ``coefficients_from_maps`` is
the same fp64 formula applied per cell for smoothly varying media (random
blob media and head phantoms need per-cell eps_r/sigma).
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Callable

import torch

EPS0: float = 8.8541878128e-12
MU0: float = 1.2566370614e-6


@dataclass
class Material:
    """One homogeneous medium (``src/core/materials.py:33-94``)."""

    name: str
    eps_r: float
    sigma: float
    mu_r: float = 1.0

    def __post_init__(self) -> None:
        if self.eps_r < 1.0:
            raise ValueError(f"eps_r must be >= 1, got {self.eps_r} for '{self.name}'")
        if self.sigma < 0.0:
            raise ValueError(f"sigma must be >= 0, got {self.sigma} for '{self.name}'")
        if self.mu_r <= 0.0:
            raise ValueError(f"mu_r must be > 0, got {self.mu_r} for '{self.name}'")

    def _alpha(self, dt: float) -> float:
        return self.sigma * dt / (2.0 * EPS0 * self.eps_r)

    def ca(self, dt: float) -> float:
        """``(1 - a)/(1 + a)``, ``a = sigma dt / (2 eps0 eps_r)``."""
        a = self._alpha(dt)
        return (1.0 - a) / (1.0 + a)

    def cb(self, dt: float) -> float:
        """``(dt / (eps0 eps_r)) / (1 + a)``."""
        a = self._alpha(dt)
        return (dt / (EPS0 * self.eps_r)) / (1.0 + a)

    def skin_depth(self, freq: float) -> float:
        if self.sigma == 0.0:
            return float("inf")
        return math.sqrt(2.0 / (2.0 * math.pi * freq * MU0 * self.sigma))

    def wavelength(self, freq: float) -> float:
        c0 = 1.0 / math.sqrt(EPS0 * MU0)
        return c0 / (freq * math.sqrt(self.eps_r * self.mu_r))


TISSUE_LIBRARY: dict[str, Material] = {
    "free_space": Material("free_space", eps_r=1.0, sigma=0.000, mu_r=1.0),
    "skull": Material("skull", eps_r=8.0, sigma=0.100, mu_r=1.0),
    "brain": Material("brain", eps_r=40.0, sigma=1.500, mu_r=1.0),
    "tumor": Material("tumor", eps_r=55.0, sigma=2.100, mu_r=1.0),
    "blood": Material("blood", eps_r=55.0, sigma=2.100, mu_r=1.0),
    "saline": Material("saline", eps_r=70.0, sigma=2.000, mu_r=1.0),
}


def _validate_ca(ca_min: float) -> None:
    if ca_min <= 0.0:
        raise ValueError(
            f"Ca <= 0 detected (min={ca_min:.6f}). Check material conductivities or reduce dt. "
            "Physical passive media require Ca in (0, 1]."
        )
    if ca_min < 0.5:
        warnings.warn(
            f"Some Ca values are below 0.5 (min={ca_min:.4f}). Very high conductivity — verify material "
            "parameters.",
            UserWarning,
            stacklevel=3,
        )


class MaterialMap:
    """Region painter producing ``(Ca, Cb)`` of shape ``(Nx, Ny, 1)``."""

    def __init__(self, grid, default: Material | None = None) -> None:
        self._grid = grid
        self._default = default or TISSUE_LIBRARY["free_space"]
        self._regions: list[Callable[[torch.Tensor, torch.Tensor], torch.Tensor]] = []
        self._materials: list[Material] = []

    def add_circle(self, center, radius: float, material: Material) -> "MaterialMap":
        ci, cj = float(center[0]), float(center[1])
        r2 = radius * radius
        self._regions.append(lambda I, J: ((I - ci) ** 2 + (J - cj) ** 2) <= r2)
        self._materials.append(material)
        return self

    def add_rectangle(self, i_range, j_range, material: Material) -> "MaterialMap":
        i0, i1 = i_range
        j0, j1 = j_range
        self._regions.append(lambda I, J: (I >= i0) & (I <= i1) & (J >= j0) & (J <= j1))
        self._materials.append(material)
        return self

    def add_custom(self, mask_fn, material: Material) -> "MaterialMap":
        self._regions.append(mask_fn)
        self._materials.append(material)
        return self

    def build(self) -> tuple[torch.Tensor, torch.Tensor]:
        Nx, Ny = self._grid.Nx, self._grid.Ny
        dt = self._grid.dt
        ca = torch.full((Nx, Ny), self._default.ca(dt), dtype=torch.float32)
        cb = torch.full((Nx, Ny), self._default.cb(dt), dtype=torch.float32)
        I = torch.arange(Nx, dtype=torch.float32).unsqueeze(1).expand(Nx, Ny)
        J = torch.arange(Ny, dtype=torch.float32).unsqueeze(0).expand(Nx, Ny)
        for mask_fn, mat in zip(self._regions, self._materials):
            m = mask_fn(I, J)
            ca[m] = mat.ca(dt)
            cb[m] = mat.cb(dt)
        _validate_ca(float(ca.min().item()))
        dev, dtype = self._grid.device, self._grid.dtype
        return ca.unsqueeze(2).to(device=dev, dtype=dtype), cb.unsqueeze(2).to(device=dev, dtype=dtype)

    def summary(self) -> str:
        dt = self._grid.dt
        lines = [
            f"MaterialMap — grid ({self._grid.Nx}×{self._grid.Ny}), dt={dt:.3e} s",
            f"  Default: {self._default.name} (eps_r={self._default.eps_r}, σ={self._default.sigma})",
        ]
        for k, mat in enumerate(self._materials):
            lines.append(
                f"  Region {k}: {mat.name} — eps_r={mat.eps_r}, σ={mat.sigma} S/m, "
                f"Ca={mat.ca(dt):.6f}, Cb={mat.cb(dt):.3e}"
            )
        return "\n".join(lines)

    def __repr__(self) -> str:
        return (
            f"MaterialMap(grid=({self._grid.Nx},{self._grid.Ny}), n_regions={len(self._regions)}, "
            f"default={self._default.name})"
        )


class MaterialMap3D(MaterialMap):
    """``MaterialMap`` with 3D primitives and ``build3d`` (``src/core/materials.py:281-477``):
    spheres, cylinders, boxes, ellipsoids and custom masks over fp32 index
    grids ``(Nx, Ny, Nz)``, painted in order over the default material."""

    def __init__(self, grid, default: Material | None = None) -> None:
        super().__init__(grid, default)
        self._regions3d: list = []
        self._materials3d: list[Material] = []

    def add_sphere(self, center, radius: float, material: Material) -> "MaterialMap3D":
        ci, cj, ck = float(center[0]), float(center[1]), float(center[2])
        r2 = radius * radius
        self._regions3d.append(lambda I, J, K: (I - ci) ** 2 + (J - cj) ** 2 + (K - ck) ** 2 <= r2)
        self._materials3d.append(material)
        return self

    def add_cylinder(self, center, axis: str, radius: float, height: float, material: Material) -> "MaterialMap3D":
        ci, cj, ck = float(center[0]), float(center[1]), float(center[2])
        r2 = radius * radius
        h2 = height / 2.0
        if axis == "x":
            def mask(I, J, K):
                return ((J - cj) ** 2 + (K - ck) ** 2 <= r2) & ((I - ci).abs() <= h2)
        elif axis == "y":
            def mask(I, J, K):
                return ((I - ci) ** 2 + (K - ck) ** 2 <= r2) & ((J - cj).abs() <= h2)
        else:
            def mask(I, J, K):
                return ((I - ci) ** 2 + (J - cj) ** 2 <= r2) & ((K - ck).abs() <= h2)
        self._regions3d.append(mask)
        self._materials3d.append(material)
        return self

    def add_box(self, corner_min, corner_max, material: Material) -> "MaterialMap3D":
        i0, j0, k0 = (float(v) for v in corner_min)
        i1, j1, k1 = (float(v) for v in corner_max)
        self._regions3d.append(lambda I, J, K: (I >= i0) & (I <= i1) & (J >= j0) & (J <= j1) & (K >= k0) & (K <= k1))
        self._materials3d.append(material)
        return self

    def add_ellipsoid(self, center, semi_axes, material: Material) -> "MaterialMap3D":
        ci, cj, ck = float(center[0]), float(center[1]), float(center[2])
        a2, b2, c2 = (float(v) ** 2 for v in semi_axes)
        self._regions3d.append(lambda I, J, K: (I - ci) ** 2 / a2 + (J - cj) ** 2 / b2 + (K - ck) ** 2 / c2 <= 1.0)
        self._materials3d.append(material)
        return self

    def add_custom3d(self, mask_fn, material: Material) -> "MaterialMap3D":
        self._regions3d.append(mask_fn)
        self._materials3d.append(material)
        return self

    def build3d(self) -> tuple[torch.Tensor, torch.Tensor]:
        Nx, Ny, Nz = self._grid.Nx, self._grid.Ny, self._grid.Nz
        dt = self._grid.dt
        ca = torch.full((Nx, Ny, Nz), self._default.ca(dt), dtype=torch.float32)
        cb = torch.full((Nx, Ny, Nz), self._default.cb(dt), dtype=torch.float32)
        I = torch.arange(Nx, dtype=torch.float32).view(-1, 1, 1).expand(Nx, Ny, Nz)
        J = torch.arange(Ny, dtype=torch.float32).view(1, -1, 1).expand(Nx, Ny, Nz)
        K = torch.arange(Nz, dtype=torch.float32).view(1, 1, -1).expand(Nx, Ny, Nz)
        for mask_fn, mat in zip(self._regions3d, self._materials3d):
            m = mask_fn(I, J, K)
            ca[m] = mat.ca(dt)
            cb[m] = mat.cb(dt)
        _validate_ca(float(ca.min().item()))
        dev, dtype = self._grid.device, self._grid.dtype
        return ca.to(device=dev, dtype=dtype), cb.to(device=dev, dtype=dtype)


def coefficients_from_maps(eps_r, sigma, dt: float, device=None, validate: bool = True):
    """Per-cell ``(Ca, Cb)`` from eps_r / sigma maps ``(Nx, Ny)``: the
    ``Material.ca/cb`` fp64 formula elementwise, stored fp32 as ``(Nx, Ny, 1)``."""
    eps = torch.as_tensor(eps_r, dtype=torch.float64).cpu()
    sig = torch.as_tensor(sigma, dtype=torch.float64).cpu()
    if eps.shape != sig.shape or eps.dim() != 2:
        raise ValueError("eps_r and sigma must be 2-D maps of equal shape")
    if float(eps.min()) < 1.0:
        raise ValueError("eps_r must be >= 1 everywhere")
    if float(sig.min()) < 0.0:
        raise ValueError("sigma must be >= 0 everywhere")
    alpha = sig * dt / (2.0 * EPS0 * eps)
    ca = ((1.0 - alpha) / (1.0 + alpha)).to(torch.float32)
    cb = ((dt / (EPS0 * eps)) / (1.0 + alpha)).to(torch.float32)
    if validate:
        _validate_ca(float(ca.min()))
    ca, cb = ca.unsqueeze(2), cb.unsqueeze(2)
    if device is not None:
        ca, cb = ca.to(device), cb.to(device)
    return ca, cb
