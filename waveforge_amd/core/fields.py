"""SoA field container (drop-in for the reference ``core.fields``).

Mirrors ``src/core/fields.py``: ``FIELD_COMPONENTS`` (:24), ``YEE_OFFSETS``
(:62-69) and ``FieldSet`` (:85-564). Each component is one contiguous fp32
tensor of shape ``(Nx, Ny, 1)`` (or ``(B, Nx, Ny, 1)``), ``j`` fastest, which
is also the layout the HIP kernels read: a fixed-``i`` row is the coalesced
unit. The tensors are the live simulation state (the stepper never keeps a
shadow copy), so callers may read or write them between steps.
"""

from __future__ import annotations

from typing import NamedTuple

import torch

FIELD_COMPONENTS: tuple[str, ...] = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class StaggerOffset(NamedTuple):
    component: str
    ox: float
    oy: float
    oz: float


YEE_OFFSETS: dict[str, StaggerOffset] = {
    "Ex": StaggerOffset("Ex", 0.5, 0.0, 0.0),
    "Ey": StaggerOffset("Ey", 0.0, 0.5, 0.0),
    "Ez": StaggerOffset("Ez", 0.0, 0.0, 0.5),
    "Hx": StaggerOffset("Hx", 0.0, 0.5, 0.5),
    "Hy": StaggerOffset("Hy", 0.5, 0.0, 0.5),
    "Hz": StaggerOffset("Hz", 0.5, 0.5, 0.0),
}

# Energy constants of the reference FieldSet (src/core/fields.py:77-78).
_EPS0: float = 8.854e-12
_MU0: float = 1.2566e-6


def _component_property(name: str) -> property:
    def getter(self) -> torch.Tensor:
        return self._fields[name]

    def setter(self, t: torch.Tensor) -> None:
        self._validate_tensor(name, t)
        self._fields[name] = t

    return property(getter, setter, doc=f"{name} field tensor (live state).")


class FieldSet:
    """Six-component Yee field storage (``src/core/fields.py:85-564``)."""

    def __init__(self, grid, *, batch_size: int = 1, dtype: torch.dtype | None = None) -> None:
        if batch_size < 1:
            raise ValueError(f"batch_size must be >= 1, got {batch_size!r}")
        self.grid = grid
        self.batch_size = batch_size
        self.dtype = grid.dtype if dtype is None else dtype
        self.offsets = YEE_OFFSETS
        shape = self._tensor_shape()
        self._fields: dict[str, torch.Tensor] = {
            c: torch.zeros(shape, device=grid.device, dtype=self.dtype) for c in FIELD_COMPONENTS
        }

    def _tensor_shape(self) -> tuple[int, ...]:
        Nx, Ny, Nz = self.grid.Nx, self.grid.Ny, self.grid.Nz
        return (Nx, Ny, Nz) if self.batch_size == 1 else (self.batch_size, Nx, Ny, Nz)

    def _validate_tensor(self, name: str, t: torch.Tensor) -> None:
        expected = self._tensor_shape()
        if tuple(t.shape) != expected:
            raise ValueError(f"Component '{name}' expected shape {expected}, got {tuple(t.shape)}")
        if t.dtype != self.dtype:
            raise ValueError(f"Component '{name}' expected dtype {self.dtype}, got {t.dtype}")

    Ex = _component_property("Ex")
    Ey = _component_property("Ey")
    Ez = _component_property("Ez")
    Hx = _component_property("Hx")
    Hy = _component_property("Hy")
    Hz = _component_property("Hz")

    @property
    def shape(self) -> tuple[int, ...]:
        return self._tensor_shape()

    @property
    def is_batched(self) -> bool:
        return self.batch_size > 1

    @property
    def device(self) -> torch.device:
        return self.grid.device

    def zero_(self) -> "FieldSet":
        for c in FIELD_COMPONENTS:
            self._fields[c].zero_()
        return self

    def clone(self) -> "FieldSet":
        new = FieldSet(self.grid, batch_size=self.batch_size, dtype=self.dtype)
        for c in FIELD_COMPONENTS:
            new._fields[c] = self._fields[c].clone()
        return new

    def to(self, device) -> "FieldSet":
        g = self.grid.to(device)
        new = FieldSet(g, batch_size=self.batch_size, dtype=self.dtype)
        for c in FIELD_COMPONENTS:
            new._fields[c] = self._fields[c].to(g.device)
        return new

    def copy_from_(self, other: "FieldSet") -> None:
        if other.shape != self.shape:
            raise ValueError(
                f"Cannot copy_from_: shape mismatch — self.shape={self.shape}, other.shape={other.shape}"
            )
        for c in FIELD_COMPONENTS:
            self._fields[c].copy_(other._fields[c])

    def as_dict(self) -> dict[str, torch.Tensor]:
        return dict(self._fields)

    def as_tensor(self) -> torch.Tensor:
        return torch.stack([self._fields[c] for c in FIELD_COMPONENTS], dim=1 if self.is_batched else 0)

    def from_tensor(self, t: torch.Tensor) -> None:
        Nx, Ny, Nz = self.grid.Nx, self.grid.Ny, self.grid.Nz
        unbatched = (6, Nx, Ny, Nz)
        batched = (self.batch_size, 6, Nx, Ny, Nz)
        if tuple(t.shape) == unbatched:
            for k, c in enumerate(FIELD_COMPONENTS):
                self._fields[c] = t[k]
        elif tuple(t.shape) == batched:
            for k, c in enumerate(FIELD_COMPONENTS):
                self._fields[c] = t[:, k, :, :, :]
        else:
            raise ValueError(
                f"from_tensor: unexpected shape {tuple(t.shape)}. Expected {unbatched} (unbatched) or "
                f"{batched} (batched)."
            )

    def electric(self):
        return self._fields["Ex"], self._fields["Ey"], self._fields["Ez"]

    def magnetic(self):
        return self._fields["Hx"], self._fields["Hy"], self._fields["Hz"]

    def energy_density(self) -> torch.Tensor:
        """``eps0/2 |E|^2 + mu0/2 |H|^2`` per cell (``src/core/fields.py:466-489``)."""
        acc = torch.zeros_like(self._fields["Ex"])
        for c in ("Ex", "Ey", "Ez"):
            acc = acc + self._fields[c] ** 2
        acc = acc * (_EPS0 * 0.5)
        mag = torch.zeros_like(self._fields["Hx"])
        for c in ("Hx", "Hy", "Hz"):
            mag = mag + self._fields[c] ** 2
        mag = mag * (_MU0 * 0.5)
        return acc + mag

    def total_energy(self) -> float:
        return float(self.energy_density().sum().item()) * self.grid.cell_volume

    def max_field(self) -> float:
        best = 0.0
        for c in FIELD_COMPONENTS:
            v = float(self._fields[c].abs().max().item())
            if v > best:
                best = v
        return best

    def norm(self, component: str) -> float:
        if component not in self._fields:
            raise KeyError(f"Unknown field component {component!r}. Valid names: {FIELD_COMPONENTS}")
        return float(torch.linalg.norm(self._fields[component].float()).item())

    def __repr__(self) -> str:
        return (
            f"FieldSet(shape={self.shape}, batch_size={self.batch_size}, dtype={self.dtype}, "
            f"device={self.device})"
        )
