"""First-order Mur ABC (drop-in for the reference ``core.boundaries.MurABC``).

Mirrors ``src/core/boundaries.py:119-384``: coefficients
``C = (C0 dt - d)/(C0 dt + d)`` (:177-178) validated to lie in ``(-1, 0)``
(:181-190), the ``Hz_prev`` snapshot buffer, ``snapshot()`` and ``apply()``
(left, right, bottom, top edges, :259-274) and the read-only properties.

``FDTD2D`` does not call ``snapshot``/``apply``: the HIP kernels fold the Mur
update into the Faraday pass, reading the coefficients ``C_mur_x/y`` from
this object. The two methods remain for callers that use them directly; they
run as device tensor ops on the grid's device.

The reference's CPML coefficient scaffold (``pml_depth > 0``, never used by
its 2D stepper) is outside this package's scope.
"""

from __future__ import annotations

import torch

C0: float = 299_792_458.0
EPS0: float = 8.8541878128e-12
ETA0: float = 376.730313


class MurABC:
    """First-order Mur ABC on the four Hz edges."""

    def __init__(self, grid, Hz_field: torch.Tensor, *, pml_depth: int = 0) -> None:
        if pml_depth:
            raise NotImplementedError("waveforge_amd implements the Mur ABC only (pml_depth must be 0)")
        dt, dx, dy = grid.dt, grid.dx, grid.dy
        self.C_mur_x: float = (C0 * dt - dx) / (C0 * dt + dx)
        self.C_mur_y: float = (C0 * dt - dy) / (C0 * dt + dy)
        if not (-1.0 < self.C_mur_x < 0.0):
            raise ValueError(
                f"C_mur_x={self.C_mur_x!r} is not in the open interval (-1, 0). "
                "This indicates a CFL violation — reduce dt or increase dx."
            )
        if not (-1.0 < self.C_mur_y < 0.0):
            raise ValueError(
                f"C_mur_y={self.C_mur_y!r} is not in the open interval (-1, 0). "
                "This indicates a CFL violation — reduce dt or increase dy."
            )
        self.Hz_prev: torch.Tensor = torch.zeros_like(Hz_field)
        self.Hz_curr: torch.Tensor = Hz_field
        self._Nx: int = grid.Nx
        self._Ny: int = grid.Ny
        self._pml_depth = 0
        self._has_pml = False

    def snapshot(self) -> None:
        """Copy the current Hz into ``Hz_prev``."""
        self.Hz_prev.copy_(self.Hz_curr)

    def apply(self, Hz: torch.Tensor) -> None:
        """Mur update of the four edges of ``Hz`` from ``Hz_prev`` (x-edges first)."""
        Nx, Ny, cx, cy, P = self._Nx, self._Ny, self.C_mur_x, self.C_mur_y, self.Hz_prev
        Hz[..., 0, :, 0] = P[..., 1, :, 0] + cx * (Hz[..., 1, :, 0] - P[..., 0, :, 0])
        Hz[..., Nx - 1, :, 0] = P[..., Nx - 2, :, 0] + cx * (Hz[..., Nx - 2, :, 0] - P[..., Nx - 1, :, 0])
        Hz[..., :, 0, 0] = P[..., :, 1, 0] + cy * (Hz[..., :, 1, 0] - P[..., :, 0, 0])
        Hz[..., :, Ny - 1, 0] = P[..., :, Ny - 2, 0] + cy * (Hz[..., :, Ny - 2, 0] - P[..., :, Ny - 1, 0])

    n_boundary_cells = property(lambda self: 2 * (self._Nx + self._Ny))
    mur_coefficient_x = property(lambda self: self.C_mur_x)
    mur_coefficient_y = property(lambda self: self.C_mur_y)
    has_pml = property(lambda self: self._has_pml)
    pml_depth = property(lambda self: self._pml_depth)

    def __repr__(self) -> str:
        return (
            f"MurABC(shape=({self._Nx}, {self._Ny}, 1), C_mur_x={self.C_mur_x:.6f}, "
            f"C_mur_y={self.C_mur_y:.6f}, has_pml={self._has_pml}, pml_depth={self._pml_depth}, "
            f"n_boundary_cells={self.n_boundary_cells})"
        )


class MurABC3D:
    """First-order Mur ABC on the six faces for Hx, Hy, Hz
    (``src/core/boundaries.py:391-654``): coefficients per axis (:468-470)
    validated to lie in ``(-1, 0)`` (:473-485), snapshot buffers, and faces
    x_min/x_max (Hy, Hz), y_min/y_max (Hx, Hz), z_min/z_max (Hx, Hy, only
    when Nz > 1) in that order (:506-575).

    ``FDTD3D`` does not call ``snapshot``/``apply``: its kernels snapshot the
    planes Mur reads and apply the faces in the same order. The methods
    remain for direct callers, as device tensor ops."""

    def __init__(self, grid, Hx_field: torch.Tensor, Hy_field: torch.Tensor, Hz_field: torch.Tensor) -> None:
        dt = grid.dt
        self.C_mur_x: float = (C0 * dt - grid.dx) / (C0 * dt + grid.dx)
        self.C_mur_y: float = (C0 * dt - grid.dy) / (C0 * dt + grid.dy)
        self.C_mur_z: float = (C0 * dt - grid.dz) / (C0 * dt + grid.dz)
        for name, coeff in (("C_mur_x", self.C_mur_x), ("C_mur_y", self.C_mur_y), ("C_mur_z", self.C_mur_z)):
            if not (-1.0 < coeff < 0.0):
                raise ValueError(
                    f"{name}={coeff!r} is not in the open interval (-1, 0). "
                    "This indicates a CFL violation — reduce dt or increase the corresponding cell spacing."
                )
        self.Hx_prev = torch.zeros_like(Hx_field)
        self.Hy_prev = torch.zeros_like(Hy_field)
        self.Hz_prev = torch.zeros_like(Hz_field)
        self._Hx, self._Hy, self._Hz = Hx_field, Hy_field, Hz_field
        self._Nx, self._Ny, self._Nz = grid.Nx, grid.Ny, grid.Nz

    def snapshot(self) -> None:
        self.Hx_prev.copy_(self._Hx)
        self.Hy_prev.copy_(self._Hy)
        self.Hz_prev.copy_(self._Hz)

    def apply(self, Hx: torch.Tensor, Hy: torch.Tensor, Hz: torch.Tensor) -> None:
        Nx, Ny, Nz = self._Nx, self._Ny, self._Nz
        cx, cy, cz = self.C_mur_x, self.C_mur_y, self.C_mur_z
        Px, Py, Pz = self.Hx_prev, self.Hy_prev, self.Hz_prev
        for f, P in ((Hy, Py), (Hz, Pz)):
            f[..., 0, :, :] = P[..., 1, :, :] + cx * (f[..., 1, :, :] - P[..., 0, :, :])
        for f, P in ((Hy, Py), (Hz, Pz)):
            f[..., Nx - 1, :, :] = P[..., Nx - 2, :, :] + cx * (f[..., Nx - 2, :, :] - P[..., Nx - 1, :, :])
        for f, P in ((Hx, Px), (Hz, Pz)):
            f[..., :, 0, :] = P[..., :, 1, :] + cy * (f[..., :, 1, :] - P[..., :, 0, :])
        for f, P in ((Hx, Px), (Hz, Pz)):
            f[..., :, Ny - 1, :] = P[..., :, Ny - 2, :] + cy * (f[..., :, Ny - 2, :] - P[..., :, Ny - 1, :])
        if Nz > 1:
            for f, P in ((Hx, Px), (Hy, Py)):
                f[..., :, :, 0] = P[..., :, :, 1] + cz * (f[..., :, :, 1] - P[..., :, :, 0])
            for f, P in ((Hx, Px), (Hy, Py)):
                f[..., :, :, Nz - 1] = P[..., :, :, Nz - 2] + cz * (f[..., :, :, Nz - 2] - P[..., :, :, Nz - 1])

    n_boundary_cells = property(lambda self: 2 * (self._Nx * self._Ny + self._Ny * self._Nz + self._Nx * self._Nz))
    mur_coefficient_x = property(lambda self: self.C_mur_x)
    mur_coefficient_y = property(lambda self: self.C_mur_y)
    mur_coefficient_z = property(lambda self: self.C_mur_z)

    def __repr__(self) -> str:
        return (
            f"MurABC3D(shape=({self._Nx}, {self._Ny}, {self._Nz}), C_mur_x={self.C_mur_x:.6f}, "
            f"C_mur_y={self.C_mur_y:.6f}, C_mur_z={self.C_mur_z:.6f}, n_boundary_cells={self.n_boundary_cells})"
        )
