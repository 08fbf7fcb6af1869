"""Torch-CPU restatement of the reference 3D Yee FDTD step (the oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``FDTD3D.step`` (``src/core/fdtd3d.py:262-373``) and ``MurABC3D``
(``src/core/boundaries.py:461-575``) operation by operation, with in-place
slice updates on ``(Nx, Ny, Nz)`` or ``(B, Nx, Ny, Nz)`` tensors, so fp32
results are bit-identical to the reference on the CPU:

1. snapshot Hx, Hy, Hz (``boundaries.py:495-504``);
2. soft sources in collection order (``OracleSource.inject``, with z indices);
3. Faraday for Hx on ``[:, :-1, :-1]``, Hy on ``[:-1, :, :-1]``, Hz on
   ``[:-1, :-1, :]`` (``fdtd3d.py:309-330``);
4. Mur on the faces in the order x_min (Hy, Hz), x_max (Hy, Hz), y_min (Hx,
   Hz), y_max (Hx, Hz), and, when Nz > 1, z_min (Hx, Hy), z_max (Hx, Hy), each
   ``f[face] = prev[in] + C (f[in] - prev[face])`` (``boundaries.py:506-575``);
5. Ampere for Ex on ``[:, 1:, 1:]``, Ey on ``[1:, :, 1:]``, Ez on ``[1:, 1:, :]``,
   free space ``E += De*curl`` or ``Ca*E + Cb*curl`` (``fdtd3d.py:336-359``);
6. stability check of Ex, Ey, Ez, Hx, Hy, Hz every ``n_check`` steps
   (``fdtd3d.py:390-425``).
"""

from __future__ import annotations

from typing import Optional, Sequence

import torch

from .fdtd2d_ref import C0, COMPONENTS, EPS0, MU0, OracleDivergedError, OracleSource


def mur_coefficients_3d(dt: float, dx: float, dy: float, dz: float):
    """``C = (C0 dt - d)/(C0 dt + d)`` per axis (``boundaries.py:468-470``)."""
    return tuple((C0 * dt - d) / (C0 * dt + d) for d in (dx, dy, dz))


class OracleFDTD3D:
    def __init__(self, Nx: int, Ny: int, Nz: int, dx: float, dy: float, dz: float, dt: float, *,
                 Ca: Optional[torch.Tensor] = None, Cb: Optional[torch.Tensor] = None,
                 sources: Sequence[OracleSource] = (), detectors: Sequence[tuple] = (), batch: int = 1,
                 dtype: torch.dtype = torch.float32, n_check: int = 100, stability_threshold: float = 1e10) -> None:
        """``detectors`` are ``(component, i, j, k)``, recorded after each step."""
        if (Ca is None) != (Cb is None):
            raise ValueError("Ca and Cb must be provided together or both omitted.")
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        shape = (Nx, Ny, Nz) if batch == 1 else (batch, Nx, Ny, Nz)
        self.fields = {c: torch.zeros(shape, dtype=dtype) for c in COMPONENTS}
        self.prev = {c: torch.zeros(shape, dtype=dtype) for c in ("Hx", "Hy", "Hz")}
        self.Dh = dt / MU0
        self.De = dt / EPS0
        self.cx, self.cy, self.cz = mur_coefficients_3d(dt, dx, dy, dz)
        self.Ca = None if Ca is None else Ca.to(dtype)
        self.Cb = None if Cb is None else Cb.to(dtype)
        self.sources = list(sources)
        self.detectors = list(detectors)
        self.traces: list[torch.Tensor] = []
        self.n_check, self.threshold = n_check, stability_threshold
        self.steps_completed = 0
        self.last_field_max = -1.0
        self.n_stability_checks = 0

    def apply_mur(self) -> None:
        F, P = self.fields, self.prev
        Hx, Hy, Hz = F["Hx"], F["Hy"], F["Hz"]
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        cx, cy, cz = self.cx, self.cy, self.cz
        Hy[..., 0, :, :] = P["Hy"][..., 1, :, :] + cx * (Hy[..., 1, :, :] - P["Hy"][..., 0, :, :])
        Hz[..., 0, :, :] = P["Hz"][..., 1, :, :] + cx * (Hz[..., 1, :, :] - P["Hz"][..., 0, :, :])
        Hy[..., Nx - 1, :, :] = P["Hy"][..., Nx - 2, :, :] + cx * (Hy[..., Nx - 2, :, :] - P["Hy"][..., Nx - 1, :, :])
        Hz[..., Nx - 1, :, :] = P["Hz"][..., Nx - 2, :, :] + cx * (Hz[..., Nx - 2, :, :] - P["Hz"][..., Nx - 1, :, :])
        Hx[..., :, 0, :] = P["Hx"][..., :, 1, :] + cy * (Hx[..., :, 1, :] - P["Hx"][..., :, 0, :])
        Hz[..., :, 0, :] = P["Hz"][..., :, 1, :] + cy * (Hz[..., :, 1, :] - P["Hz"][..., :, 0, :])
        Hx[..., :, Ny - 1, :] = P["Hx"][..., :, Ny - 2, :] + cy * (Hx[..., :, Ny - 2, :] - P["Hx"][..., :, Ny - 1, :])
        Hz[..., :, Ny - 1, :] = P["Hz"][..., :, Ny - 2, :] + cy * (Hz[..., :, Ny - 2, :] - P["Hz"][..., :, Ny - 1, :])
        if Nz > 1:
            Hx[..., :, :, 0] = P["Hx"][..., :, :, 1] + cz * (Hx[..., :, :, 1] - P["Hx"][..., :, :, 0])
            Hy[..., :, :, 0] = P["Hy"][..., :, :, 1] + cz * (Hy[..., :, :, 1] - P["Hy"][..., :, :, 0])
            Hx[..., :, :, Nz - 1] = P["Hx"][..., :, :, Nz - 2] + cz * (Hx[..., :, :, Nz - 2] - P["Hx"][..., :, :, Nz - 1])
            Hy[..., :, :, Nz - 1] = P["Hy"][..., :, :, Nz - 2] + cz * (Hy[..., :, :, Nz - 2] - P["Hy"][..., :, :, Nz - 1])

    def step(self) -> None:
        n = self.steps_completed
        F = self.fields
        Ex, Ey, Ez, Hx, Hy, Hz = (F[c] for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
        Dh, De, dx, dy, dz = self.Dh, self.De, self.dx, self.dy, self.dz
        for c in ("Hx", "Hy", "Hz"):
            self.prev[c].copy_(F[c])
        for src in self.sources:
            src.inject(F, n)
        Hx[..., :, :-1, :-1] += Dh * ((Ey[..., :, :-1, 1:] - Ey[..., :, :-1, :-1]) / dz
                                      - (Ez[..., :, 1:, :-1] - Ez[..., :, :-1, :-1]) / dy)
        Hy[..., :-1, :, :-1] += Dh * ((Ez[..., 1:, :, :-1] - Ez[..., :-1, :, :-1]) / dx
                                      - (Ex[..., :-1, :, 1:] - Ex[..., :-1, :, :-1]) / dz)
        Hz[..., :-1, :-1, :] += Dh * ((Ex[..., :-1, 1:, :] - Ex[..., :-1, :-1, :]) / dy
                                      - (Ey[..., 1:, :-1, :] - Ey[..., :-1, :-1, :]) / dx)
        self.apply_mur()
        cex = (Hz[..., :, 1:, 1:] - Hz[..., :, :-1, 1:]) / dy - (Hy[..., :, 1:, 1:] - Hy[..., :, 1:, :-1]) / dz
        cey = (Hx[..., 1:, :, 1:] - Hx[..., 1:, :, :-1]) / dz - (Hz[..., 1:, :, 1:] - Hz[..., :-1, :, 1:]) / dx
        cez = (Hy[..., 1:, 1:, :] - Hy[..., :-1, 1:, :]) / dx - (Hx[..., 1:, 1:, :] - Hx[..., 1:, :-1, :]) / dy
        if self.Ca is None:
            Ex[..., :, 1:, 1:] += De * cex
            Ey[..., 1:, :, 1:] += De * cey
            Ez[..., 1:, 1:, :] += De * cez
        else:
            Ca, Cb = self.Ca, self.Cb
            Ex[..., :, 1:, 1:] = Ca[..., :, 1:, 1:] * Ex[..., :, 1:, 1:] + Cb[..., :, 1:, 1:] * cex
            Ey[..., 1:, :, 1:] = Ca[..., 1:, :, 1:] * Ey[..., 1:, :, 1:] + Cb[..., 1:, :, 1:] * cey
            Ez[..., 1:, 1:, :] = Ca[..., 1:, 1:, :] * Ez[..., 1:, 1:, :] + Cb[..., 1:, 1:, :] * cez
        self.steps_completed += 1
        if self.detectors:
            self.traces.append(torch.stack([F[c][..., i, j, k] for (c, i, j, k) in self.detectors], dim=-1))
        if self.steps_completed % self.n_check == 0:
            self.check_stability()

    def check_stability(self) -> None:
        self.n_stability_checks += 1
        for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            val = float(self.fields[name].abs().max().item())
            if val > self.last_field_max:
                self.last_field_max = val
            if val > self.threshold:
                raise OracleDivergedError(f"Simulation diverged at step {self.steps_completed}: {name} max = "
                                          f"{val:.3e} exceeds threshold {self.threshold:.3e}.")

    def run(self, n_steps: int) -> None:
        for _ in range(n_steps):
            self.step()

    def trace_array(self) -> torch.Tensor:
        """``(n_det, n_steps)`` (or ``(B, n_det, n_steps)``)."""
        return torch.stack(self.traces, dim=-1)
