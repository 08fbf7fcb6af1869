"""Torch-CPU restatement of the reference 2D TM FDTD step (the oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by tests,
``smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, never by the
product.

Every operation below follows the reference's eager tensor-op sequence so
that, on the CPU, fp32 results are bit-identical to the reference itself:

* step order: ``src/core/fdtd2d.py:187-266`` (snapshot, inject, Faraday,
  Mur, Ampere, counter, stability check every ``n_check`` steps);
* Faraday: ``src/core/fdtd2d.py:232-235``
  ``Hz[:-1,:-1] += Dh*((Ex[:-1,1:]-Ex[:-1,:-1])/dy - (Ey[1:,:-1]-Ey[:-1,:-1])/dx)``;
* Mur: ``src/core/boundaries.py:177-178`` (coefficients) and ``:259-274``
  (left, right, bottom, top, each ``P[e±1] + C*(Hz[e±1] - P[e])``);
* Ampere: ``src/core/fdtd2d.py:244-257`` (free-space ``+= De*g`` /
  material ``Ca*E + Cb*g``);
* stability: ``src/core/fdtd2d.py:268-295``;
* soft sources: ``src/core/sources.py:882-887`` (point),
  ``:1031-1035`` (line), ``:715-736`` (plane), ``:1060-1071`` (order).

The class works in fp32 (bit-level oracle) or fp64 (error reference);
with fp64, pass Ca/Cb already rounded to fp32 and converted, as the
reference stores them in fp32 (``src/core/materials.py:218-250``).

Division. The reference divides by the Python floats dx, dy
(``src/core/fdtd2d.py:232-235, 250-251``). ``division="ieee"`` writes that
expression as the reference does (``x / dy``): on the CPU torch evaluates it
as IEEE ``x / fp32(dy)``. ``division="reciprocal"`` writes
``x * fp32(1/dy)`` (the reciprocal of the double, rounded once, as a 0-dim
tensor of the field dtype): the value torch forms for the reference's
``x / dy`` on a GPU, restated so the CPU can compute it. ``to(device)`` moves
the oracle's state, so the ``"ieee"`` form can also be evaluated by torch on
a GPU, where it must equal the ``"reciprocal"`` form on the CPU.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

# Constants exactly as the reference defines them.
C0 = 299_792_458.0              # src/core/grid.py:23, boundaries.py:23
EPS0 = 8.8541878128e-12         # src/core/fdtd2d.py:40
MU0 = 1.2566370614e-6           # src/core/fdtd2d.py:41

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class OracleDivergedError(RuntimeError):
    """Oracle counterpart of ``SimulationDivergedError``
    (``src/core/fdtd2d.py:48-62``)."""


def stable_dt(dx: float, dy: float, courant: float = 0.98) -> float:
    """``compute_stable_dt`` for 2D (``src/core/grid.py:120-169``)."""
    inv_sq_sum = (1.0 / dx) ** 2 + (1.0 / dy) ** 2
    return courant / (C0 * math.sqrt(inv_sq_sum))


def mur_coefficients(dt: float, dx: float, dy: float) -> tuple[float, float]:
    """``C_mur = (C0*dt - d)/(C0*dt + d)`` (``src/core/boundaries.py:177-178``)."""
    return (C0 * dt - dx) / (C0 * dt + dx), (C0 * dt - dy) / (C0 * dt + dy)


def material_ca_cb(eps_r: float, sigma: float, dt: float) -> tuple[float, float]:
    """``Material.ca/cb`` in Python fp64 (``src/core/materials.py:61-82``)."""
    alpha = sigma * dt / (2.0 * EPS0 * eps_r)
    ca = (1.0 - alpha) / (1.0 + alpha)
    cb = (dt / (EPS0 * eps_r)) / (1.0 + alpha)
    return ca, cb


def ca_cb_from_maps(eps_r, sigma, dt: float):
    """Per-cell Ca/Cb from eps_r/sigma arrays: the same fp64 formula as
    ``Material.ca/cb`` applied elementwise, stored in fp32 (the reference
    stores Ca/Cb as fp32, ``src/core/materials.py:218-250``)."""
    eps = torch.as_tensor(eps_r, dtype=torch.float64)
    sig = torch.as_tensor(sigma, dtype=torch.float64)
    alpha = sig * dt / (2.0 * EPS0 * eps)
    ca = (1.0 - alpha) / (1.0 + alpha)
    cb = (dt / (EPS0 * eps)) / (1.0 + alpha)
    return ca.to(torch.float32), cb.to(torch.float32)


@dataclass
class OracleSource:
    """One soft source: ``field[..., i, j, 0] += wave[n] * amp`` over its cells.

    ``kind`` is ``"cells"`` (point / line source: ``values = w[n:n+1]*amp``
    then indexed add, ``src/core/sources.py:882-887``) or ``"plane_xz"`` /
    ``"plane_yz"`` / ``"plane_xy"`` (``src/core/sources.py:715-736``: the
    0-dim ``w[n]`` is added to the whole slice, no amplitude multiply).
    """

    component: str
    wave: torch.Tensor                     # (N_steps,) fp32/fp64 CPU
    i_idx: Optional[torch.Tensor] = None   # int64 (n_cells,)
    j_idx: Optional[torch.Tensor] = None
    amp: Optional[torch.Tensor] = None     # (n_cells,)
    kind: str = "cells"
    position: int = 0
    k_idx: Optional[torch.Tensor] = None   # 3D grids: z index of each cell (default 0)

    def inject(self, fields: dict, n: int) -> None:
        target = fields[self.component]
        if self.kind == "cells":
            w_now = self.wave[n:n + 1]
            values = w_now * self.amp
            k_idx = torch.zeros_like(self.i_idx) if self.k_idx is None else self.k_idx
            target[..., self.i_idx, self.j_idx, k_idx] += values
        else:
            w_now = self.wave[n]
            if self.kind == "plane_xy":
                target[..., :, :, self.position] += w_now
            elif self.kind == "plane_xz":
                target[..., :, self.position, :] += w_now
            elif self.kind == "plane_yz":
                target[..., self.position, :, :] += w_now
            else:
                raise ValueError(self.kind)


class OracleFDTD2D:
    """CPU restatement of ``FDTD2D`` + ``MurABC`` + ``SourceCollection``.

    Fields have shape ``(Nx, Ny, 1)`` or ``(B, Nx, Ny, 1)``; Ca/Cb
    ``(Nx, Ny, 1)`` or broadcastable. ``detectors`` is a list of ``(i, j)``
    Hz cells recorded after each step, as the reference's callers do
    (``examples/brain_mimo_imaging.py:196-198``).
    """

    def __init__(
        self,
        Nx: int,
        Ny: int,
        dx: float,
        dy: float,
        dt: float,
        *,
        Ca: Optional[torch.Tensor] = None,
        Cb: Optional[torch.Tensor] = None,
        sources: Sequence[OracleSource] = (),
        detectors: Sequence[tuple[int, int]] = (),
        batch: int = 1,
        dtype: torch.dtype = torch.float32,
        n_check: int = 100,
        stability_threshold: float = 1e10,
        functional: bool = False,
        division: str = "ieee",
    ) -> None:
        if Nx < 4 or Ny < 4:
            raise ValueError("grid must be at least 4x4")
        if (Ca is None) != (Cb is None):
            raise ValueError("Ca and Cb must be provided together or both omitted.")
        self.Nx, self.Ny = Nx, Ny
        self.dx, self.dy, self.dt = dx, dy, dt
        shape = (Nx, Ny, 1) if batch == 1 else (batch, Nx, Ny, 1)
        self.fields = {c: torch.zeros(shape, dtype=dtype) for c in COMPONENTS}
        self.Hz_prev = torch.zeros(shape, dtype=dtype)
        self.Dh = dt / MU0          # fdtd2d.py:158
        self.De = dt / EPS0         # fdtd2d.py:159
        self.cx, self.cy = mur_coefficients(dt, dx, dy)
        self.Ca = None if Ca is None else Ca.to(dtype)
        self.Cb = None if Cb is None else Cb.to(dtype)
        self.sources = list(sources)
        self.detectors = list(detectors)
        self.traces: list[torch.Tensor] = []
        self.n_check = n_check
        self.threshold = stability_threshold
        self.steps_completed = 0
        self.last_field_max = -1.0
        self.n_stability_checks = 0
        self.functional = functional
        if division not in ("ieee", "reciprocal"):
            raise ValueError(f"division must be 'ieee' or 'reciprocal', got {division!r}")
        self.division = division
        self._rdx = torch.tensor(1.0 / dx, dtype=dtype)   # fp32(1/dx) for fp32 fields
        self._rdy = torch.tensor(1.0 / dy, dtype=dtype)

    def _over_dx(self, x: torch.Tensor) -> torch.Tensor:
        return x * self._rdx if self.division == "reciprocal" else x / self.dx

    def _over_dy(self, x: torch.Tensor) -> torch.Tensor:
        return x * self._rdy if self.division == "reciprocal" else x / self.dy

    def to(self, device) -> "OracleFDTD2D":
        """Move fields, snapshot, coefficients and sources to ``device``."""
        for k in self.fields:
            self.fields[k] = self.fields[k].to(device)
        self.Hz_prev = self.Hz_prev.to(device)
        if self.Ca is not None:
            self.Ca, self.Cb = self.Ca.to(device), self.Cb.to(device)
        self._rdx, self._rdy = self._rdx.to(device), self._rdy.to(device)
        for src in self.sources:
            src.wave = src.wave.to(device)
            for name in ("i_idx", "j_idx", "amp", "k_idx"):
                v = getattr(src, name)
                if v is not None:
                    setattr(src, name, v.to(device))
        return self

    # --- convenience accessors -------------------------------------------
    @property
    def Ex(self) -> torch.Tensor:
        return self.fields["Ex"]

    @property
    def Ey(self) -> torch.Tensor:
        return self.fields["Ey"]

    @property
    def Hz(self) -> torch.Tensor:
        return self.fields["Hz"]

    # --- the step ----------------------------------------------------------
    def snapshot(self) -> None:
        """``MurABC.snapshot`` (``src/core/boundaries.py:230-237``)."""
        self.Hz_prev.copy_(self.fields["Hz"])

    def apply_mur(self, Hz: torch.Tensor) -> None:
        """``MurABC.apply`` (``src/core/boundaries.py:239-274``)."""
        Nx, Ny = self.Nx, self.Ny
        cx, cy = self.cx, self.cy
        P = self.Hz_prev
        Hz[..., 0, :, 0] = P[..., 1, :, 0] + cx * (Hz[..., 1, :, 0] - P[..., 0, :, 0])
        Hz[..., Nx - 1, :, 0] = (
            P[..., Nx - 2, :, 0] + cx * (Hz[..., Nx - 2, :, 0] - P[..., Nx - 1, :, 0])
        )
        Hz[..., :, 0, 0] = P[..., :, 1, 0] + cy * (Hz[..., :, 1, 0] - P[..., :, 0, 0])
        Hz[..., :, Ny - 1, 0] = (
            P[..., :, Ny - 2, 0] + cy * (Hz[..., :, Ny - 2, 0] - P[..., :, Ny - 1, 0])
        )

    def step(self) -> None:
        """One leapfrog step (``src/core/fdtd2d.py:187-266``)."""
        if self.functional:
            return self._step_functional()
        n = self.steps_completed
        Ex, Ey, Hz = self.fields["Ex"], self.fields["Ey"], self.fields["Hz"]
        Dh, De = self.Dh, self.De

        self.snapshot()
        for src in self.sources:
            src.inject(self.fields, n)

        Hz[..., :-1, :-1, :] += Dh * (
            self._over_dy(Ex[..., :-1, 1:, :] - Ex[..., :-1, :-1, :])
            - self._over_dx(Ey[..., 1:, :-1, :] - Ey[..., :-1, :-1, :])
        )
        self.apply_mur(Hz)

        dHz_dy = self._over_dy(Hz[..., :, 1:, :] - Hz[..., :, :-1, :])
        dHz_dx = self._over_dx(Hz[..., 1:, :, :] - Hz[..., :-1, :, :])
        if self.Ca is None:
            Ex[..., :, 1:, :] += De * dHz_dy
            Ey[..., 1:, :, :] -= De * dHz_dx
        else:
            Ex[..., :, 1:, :] = self.Ca[..., :, 1:, :] * Ex[..., :, 1:, :] + self.Cb[..., :, 1:, :] * dHz_dy
            Ey[..., 1:, :, :] = self.Ca[..., 1:, :, :] * Ey[..., 1:, :, :] - self.Cb[..., 1:, :, :] * dHz_dx

        self.steps_completed += 1
        if self.detectors:
            ii = torch.tensor([d[0] for d in self.detectors])
            jj = torch.tensor([d[1] for d in self.detectors])
            self.traces.append(Hz[..., ii, jj, 0].clone())
        if self.steps_completed % self.n_check == 0:
            self.check_stability()

    def _step_functional(self) -> None:
        """The same step written without in-place updates, so torch autograd
        can differentiate a whole run (used in fp64 as the gradient reference).
        Same formulas and order; adding the Faraday term as zero-padded full
        arrays and rebuilding the Mur edges by concatenation."""
        n = self.steps_completed
        Ex, Ey, Hz = (self.fields[k][..., 0] for k in ("Ex", "Ey", "Hz"))
        Dh, De = self.Dh, self.De
        P = Hz
        comp = {"Ex": Ex, "Ey": Ey, "Hz": Hz}
        for src in self.sources:
            t = comp[src.component]
            if src.kind != "cells":
                raise ValueError("functional oracle supports point and line sources only")
            vals = src.wave[n:n + 1] * src.amp
            comp[src.component] = t.index_put((src.i_idx, src.j_idx), vals, accumulate=True)
        Ex, Ey, H1 = comp["Ex"], comp["Ey"], comp["Hz"]
        dF = Dh * (self._over_dy(Ex[:-1, 1:] - Ex[:-1, :-1]) - self._over_dx(Ey[1:, :-1] - Ey[:-1, :-1]))
        H2 = H1 + torch.nn.functional.pad(dF, (0, 1, 0, 1))
        cx, cy = self.cx, self.cy
        L = P[1, :] + cx * (H2[1, :] - P[0, :])
        R = P[-2, :] + cx * (H2[-2, :] - P[-1, :])
        H3 = torch.cat([L[None], H2[1:-1], R[None]], 0)
        B = P[:, 1] + cy * (H3[:, 1] - P[:, 0])
        T = P[:, -2] + cy * (H3[:, -2] - P[:, -1])
        H4 = torch.cat([B[:, None], H3[:, 1:-1], T[:, None]], 1)
        gy = self._over_dy(H4[:, 1:] - H4[:, :-1])
        gx = self._over_dx(H4[1:, :] - H4[:-1, :])
        if self.Ca is None:
            Ex2 = torch.cat([Ex[:, :1], Ex[:, 1:] + De * gy], 1)
            Ey2 = torch.cat([Ey[:1, :], Ey[1:, :] - De * gx], 0)
        else:
            Ca, Cb = self.Ca[..., 0], self.Cb[..., 0]
            Ex2 = torch.cat([Ex[:, :1], Ca[:, 1:] * Ex[:, 1:] + Cb[:, 1:] * gy], 1)
            Ey2 = torch.cat([Ey[:1, :], Ca[1:, :] * Ey[1:, :] - Cb[1:, :] * gx], 0)
        self.fields["Ex"], self.fields["Ey"], self.fields["Hz"] = Ex2[..., None], Ey2[..., None], H4[..., None]
        self.steps_completed += 1
        if self.detectors:
            ii = torch.tensor([d[0] for d in self.detectors])
            jj = torch.tensor([d[1] for d in self.detectors])
            self.traces.append(H4[ii, jj])

    def check_stability(self) -> None:
        """``_check_stability`` (``src/core/fdtd2d.py:268-295``)."""
        self.n_stability_checks += 1
        for name in ("Ex", "Ey", "Hz"):
            val = float(self.fields[name].abs().max().item())
            if val > self.last_field_max:
                self.last_field_max = val
            if val > self.threshold:
                raise OracleDivergedError(
                    f"Simulation diverged at step {self.steps_completed}: "
                    f"{name} max = {val:.3e} exceeds threshold {self.threshold:.3e}."
                )

    def run(self, n_steps: int) -> None:
        for _ in range(n_steps):
            self.step()

    def trace_array(self) -> torch.Tensor:
        """Recorded detector samples as ``(..., N_det, T)``."""
        return torch.stack(self.traces, dim=-1)
