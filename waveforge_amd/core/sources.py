"""Waveforms and soft sources (drop-in for the reference ``core.sources``).

Mirrors ``src/core/sources.py``: the ``Waveform`` ABC (:64-155) and its six
subclasses with the same parameters, defaults, validation and formulas
(Gaussian :238-241, sinusoid :316-321, Ricker :399-403, modulated Gaussian
:465-468, chirp :524-531, UWB monocycle :628-633), and the ``PointSource``
(:774-887), ``LineSource`` (:895-1035), ``PlaneSource`` (:671-766) and
``SourceCollection`` (:1043-1090) injectors with the same attributes.

Two deliberate differences, both invisible to results:

* ``Waveform.build`` evaluates the table on the CPU in the requested dtype
  and then moves it to the device, so a table is bit-identical whichever
  device runs the simulation (the reference evaluates on the target device,
  whose exp/sin may differ from the CPU's in the last ulp).
* The sources' ``step()`` remains available for direct calls (tests call it),
  but ``FDTD2D`` does not call it: the HIP kernels read the flattened source
  table built from the same attributes (cells, amplitudes, waveform) and
  apply ``field[cell] += wave[n] * amp`` in collection order.

``GaussianDerivative`` is an addition: the Gaussian-derivative pulse of the
256x256 slab workload, specified in steps of t0 and width.
"""

from __future__ import annotations

import math
import warnings
from abc import ABC, abstractmethod
from typing import Union

import numpy as np
import torch

VALID_COMPONENTS: tuple[str, ...] = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class Waveform(ABC):
    """Pre-computed source waveform (``src/core/sources.py:64-155``)."""

    @abstractmethod
    def _compute(self, t: torch.Tensor) -> torch.Tensor: ...

    @property
    @abstractmethod
    def peak_time(self) -> float | None: ...

    @property
    @abstractmethod
    def bandwidth(self) -> float: ...

    @property
    @abstractmethod
    def amplitude(self) -> float: ...

    def build(self, N_steps: int, dt: float, device, dtype: torch.dtype = torch.float32) -> torch.Tensor:
        """``f(arange(N) * dt)``, shape ``(N_steps,)``, evaluated on the CPU
        (``t`` in ``dtype``, as the reference) and placed on ``device``."""
        t = torch.arange(N_steps, dtype=dtype) * dt
        return self._compute(t).to(device)

    def is_causal(self, dt: float) -> bool:
        f0 = float(self._compute(torch.tensor([0.0], dtype=torch.float32)).item())
        return abs(f0) / abs(self.amplitude) < 1e-4


def _nonzero_amplitude(amplitude: float) -> float:
    if amplitude == 0:
        raise ValueError("amplitude must be non-zero.")
    return float(amplitude)


class GaussianPulse(Waveform):
    """``A exp(-(t - t0)^2 / (2 sigma^2))``; ``t0`` defaults to ``5 sigma``."""

    def __init__(self, amplitude: float = 1.0, t0: float | None = None, sigma: float | None = None, *,
                 freq: float | None = None) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if freq is not None and sigma is not None:
            raise ValueError("Provide either freq or sigma, not both.")
        if freq is not None:
            if freq <= 0:
                raise ValueError(f"freq must be > 0, got {freq!r}")
            sigma = 1.0 / (2.0 * math.pi * freq / 3.0)
        if sigma is None:
            raise ValueError("Either sigma or freq must be supplied.")
        if sigma <= 0:
            raise ValueError(f"sigma must be > 0, got {sigma!r}")
        if t0 is None:
            t0 = 5.0 * sigma
        if t0 <= 0:
            raise ValueError(f"t0 must be > 0, got {t0!r}")
        self._amplitude, self._sigma, self._t0 = amplitude, float(sigma), float(t0)
        if not self.is_causal(0.0):
            warnings.warn(
                f"GaussianPulse: |f(0)|/|A| >= 1e-4 — waveform is not causal. Consider increasing t0 "
                f"(currently {self._t0:.4g} s) relative to sigma ({self._sigma:.4g} s).",
                UserWarning,
                stacklevel=2,
            )

    def _compute(self, t):
        exponent = -((t - self._t0) ** 2) / (2.0 * self._sigma ** 2)
        return self._amplitude * torch.exp(exponent)

    peak_time = property(lambda self: self._t0)
    bandwidth = property(lambda self: 1.0 / (2.0 * math.pi * self._sigma))
    amplitude = property(lambda self: self._amplitude)


class SinusoidalSource(Waveform):
    """``A sin(2 pi f0 t + phi) (1 - exp(-t / tau))``, ``tau = 3/(2 pi f0)``."""

    def __init__(self, amplitude: float = 1.0, frequency: float = 1e9, phase: float = 0.0) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if frequency <= 0:
            raise ValueError(f"frequency must be > 0, got {frequency!r}")
        self._amplitude, self._frequency, self._phase = amplitude, float(frequency), float(phase)
        self._tau_ramp = 3.0 / (2.0 * math.pi * self._frequency)
        if phase != 0.0 and abs(math.sin(phase)) >= 1e-4:
            warnings.warn(
                f"SinusoidalSource: phase={phase:.4g} rad gives |sin(phi)|={abs(math.sin(phase)):.4g} >= 1e-4 — "
                "source is not causal at t=0.",
                UserWarning,
                stacklevel=2,
            )

    def _compute(self, t):
        omega = 2.0 * math.pi * self._frequency
        carrier = torch.sin(omega * t + self._phase)
        ramp = 1.0 - torch.exp(-t / self._tau_ramp)
        return self._amplitude * carrier * ramp

    peak_time = property(lambda self: -1.0)
    bandwidth = property(lambda self: 0.0)
    amplitude = property(lambda self: self._amplitude)


class RickerWavelet(Waveform):
    """``A (1 - 2u) exp(-u)``, ``u = (pi fp (t - t0))^2``; ``t0`` defaults to ``1.5/fp``."""

    def __init__(self, amplitude: float = 1.0, peak_freq: float = 1e9, t0: float | None = None) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if peak_freq <= 0:
            raise ValueError(f"peak_freq must be > 0, got {peak_freq!r}")
        if t0 is None:
            t0 = 1.5 / peak_freq
        if t0 <= 0:
            raise ValueError(f"t0 must be > 0, got {t0!r}")
        self._amplitude, self._peak_freq, self._t0 = amplitude, float(peak_freq), float(t0)
        if not self.is_causal(0.0):
            warnings.warn(
                f"RickerWavelet: |f(0)|/|A| >= 1e-4 — waveform is not causal. Consider increasing t0 "
                f"(currently {self._t0:.4g} s) relative to 1/peak_freq ({1.0 / self._peak_freq:.4g} s).",
                UserWarning,
                stacklevel=2,
            )

    def _compute(self, t):
        pi2_fp2 = (math.pi * self._peak_freq) ** 2
        u = pi2_fp2 * (t - self._t0) ** 2
        return self._amplitude * (1.0 - 2.0 * u) * torch.exp(-u)

    peak_time = property(lambda self: self._t0)
    bandwidth = property(lambda self: self._peak_freq)
    amplitude = property(lambda self: self._amplitude)


class ModulatedGaussian(Waveform):
    """``A sin(2 pi fc t + phi) exp(-(t - t0)^2 / (2 sigma^2))``."""

    def __init__(self, amplitude: float = 1.0, fc: float = 1e9, sigma: float | None = None,
                 t0: float | None = None, phi: float = 0.0) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if fc <= 0:
            raise ValueError(f"fc must be > 0, got {fc!r}")
        if sigma is None:
            raise ValueError("sigma must be supplied.")
        if sigma <= 0:
            raise ValueError(f"sigma must be > 0, got {sigma!r}")
        if t0 is None:
            t0 = 5.0 * sigma
        if t0 <= 0:
            raise ValueError(f"t0 must be > 0, got {t0!r}")
        self._amplitude, self._fc, self._sigma = amplitude, float(fc), float(sigma)
        self._t0, self._phi = float(t0), float(phi)
        if not self.is_causal(0.0):
            warnings.warn(
                f"ModulatedGaussian: |f(0)|/|A| >= 1e-4 — waveform is not causal. With phi={self._phi:.4g} rad, "
                f"consider increasing t0 (currently {self._t0:.4g} s) to >= 5*sigma ({5.0 * self._sigma:.4g} s).",
                UserWarning,
                stacklevel=2,
            )

    def _compute(self, t):
        carrier = torch.sin(2.0 * math.pi * self._fc * t + self._phi)
        envelope = torch.exp(-((t - self._t0) ** 2) / (2.0 * self._sigma ** 2))
        return self._amplitude * carrier * envelope

    peak_time = property(lambda self: self._t0)
    bandwidth = property(lambda self: 1.0 / (2.0 * math.pi * self._sigma))
    amplitude = property(lambda self: self._amplitude)


class Chirp(Waveform):
    """Hann-windowed linear sweep ``f_start -> f_end`` over ``[t_start, t_end]``."""

    def __init__(self, amplitude: float = 1.0, f_start: float = 1e8, f_end: float = 1e9, t_start: float = 0.0,
                 t_end: float | None = None, phi: float = 0.0) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if f_start <= 0:
            raise ValueError(f"f_start must be > 0, got {f_start!r}")
        if f_end <= 0:
            raise ValueError(f"f_end must be > 0, got {f_end!r}")
        if t_start < 0:
            raise ValueError(f"t_start must be >= 0, got {t_start!r}")
        if t_end is None:
            t_end = t_start + 10.0 / f_start
        if t_end <= t_start:
            raise ValueError(f"t_end ({t_end!r}) must be > t_start ({t_start!r})")
        self._amplitude, self._f_start, self._f_end = amplitude, float(f_start), float(f_end)
        self._t_start, self._t_end, self._phi = float(t_start), float(t_end), float(phi)
        self._duration = self._t_end - self._t_start

    def _compute(self, t):
        in_window = (t >= self._t_start) & (t <= self._t_end)
        tau = t - self._t_start
        k = (self._f_end - self._f_start) / (2.0 * self._duration)
        inst_phase = 2.0 * math.pi * (self._f_start * tau + k * tau ** 2) + self._phi
        hann = 0.5 * (1.0 - torch.cos(2.0 * math.pi * tau / self._duration))
        result = self._amplitude * torch.sin(inst_phase) * hann
        return torch.where(in_window, result, torch.zeros_like(result))

    peak_time = property(lambda self: (self._t_start + self._t_end) / 2.0)
    bandwidth = property(lambda self: abs(self._f_end - self._f_start))
    amplitude = property(lambda self: self._amplitude)


class UWBPulse(Waveform):
    """Gaussian monocycle normalised to peak ``A``; ``sigma = 1/(sqrt2 pi fc)``."""

    def __init__(self, amplitude: float = 1.0, f_low: float = 1e9, f_high: float = 3e9,
                 t0: float | None = None) -> None:
        amplitude = _nonzero_amplitude(amplitude)
        if f_low <= 0:
            raise ValueError(f"f_low must be > 0, got {f_low!r}")
        if f_high <= f_low:
            raise ValueError(f"f_high must be > f_low, got f_low={f_low!r}, f_high={f_high!r}")
        self._amplitude, self._f_low, self._f_high = amplitude, float(f_low), float(f_high)
        self._fc = (f_low + f_high) / 2.0
        self._sigma = 1.0 / (math.sqrt(2.0) * math.pi * self._fc)
        self._norm = self._sigma * math.sqrt(math.e)
        self._t0 = float(t0) if t0 is not None else 6.0 * self._sigma
        if not self.is_causal(0.0):
            warnings.warn(
                f"UWBPulse: |f(0)|/|A| >= 1e-4 — waveform may not be causal. Consider increasing t0 "
                f"(currently {self._t0:.4g} s) to >= 5*sigma ({5.0 * self._sigma:.4g} s).",
                UserWarning,
                stacklevel=2,
            )

    def _compute(self, t):
        tau = t - self._t0
        g = torch.exp(-(tau ** 2) / (2.0 * self._sigma ** 2))
        raw = -(tau / self._sigma ** 2) * g
        return self._amplitude * self._norm * raw

    fc = property(lambda self: self._fc)
    f_low = property(lambda self: self._f_low)
    f_high = property(lambda self: self._f_high)
    peak_time = property(lambda self: self._t0)
    bandwidth = property(lambda self: self._f_high - self._f_low)
    amplitude = property(lambda self: self._amplitude)


class GaussianDerivative(Waveform):
    """``-A ((t - t0)/w) exp(-(t - t0)^2 / (2 w^2))`` with ``t0``/``w`` in seconds.

    Use :meth:`in_steps` to give them in time steps (the 256x256 slab workload
    uses ``t0 = 60`` steps, ``w = 15`` steps)."""

    def __init__(self, amplitude: float = 1.0, t0: float = 1e-10, width: float = 2.5e-11) -> None:
        self._amplitude = _nonzero_amplitude(amplitude)
        if width <= 0:
            raise ValueError(f"width must be > 0, got {width!r}")
        if t0 <= 0:
            raise ValueError(f"t0 must be > 0, got {t0!r}")
        self._t0, self._width = float(t0), float(width)

    @classmethod
    def in_steps(cls, dt: float, t0_steps: float, width_steps: float, amplitude: float = 1.0):
        return cls(amplitude, t0=t0_steps * dt, width=width_steps * dt)

    def _compute(self, t):
        tau = t - self._t0
        return self._amplitude * (-(tau / self._width)) * torch.exp(-(tau ** 2) / (2.0 * self._width ** 2))

    peak_time = property(lambda self: self._t0 - self._width)
    bandwidth = property(lambda self: 1.0 / (2.0 * math.pi * self._width))
    amplitude = property(lambda self: self._amplitude)


# ---------------------------------------------------------------------------
# Injectors
# ---------------------------------------------------------------------------


def _check_component(component: str) -> None:
    if component not in VALID_COMPONENTS:
        raise ValueError(f"component must be one of {VALID_COMPONENTS}, got {component!r}")


class PlaneSource:
    """Uniform soft source over an ``'xy'``, ``'xz'`` or ``'yz'`` plane."""

    def __init__(self, waveform: Waveform, plane: str, position: int, component: str = "Hz", *, grid,
                 N_steps: int) -> None:
        _check_component(component)
        if plane not in ("xy", "xz", "yz"):
            raise ValueError(f"plane must be 'xy', 'xz', or 'yz', got {plane!r}")
        limit, n_cells = {
            "xy": (grid.Nz, grid.Nx * grid.Ny),
            "xz": (grid.Ny, grid.Nx * grid.Nz),
            "yz": (grid.Nx, grid.Ny * grid.Nz),
        }[plane]
        if not (0 <= position < limit):
            raise ValueError(f"position={position} out of range [0, {limit}) for plane={plane!r}")
        self._component, self._plane, self._position = component, plane, position
        self._N_steps, self._n_cells = N_steps, n_cells
        self._waveform = waveform
        self._waveform_tensor = waveform.build(N_steps, grid.dt, grid.device, grid.dtype)

    def step(self, fields_dict, n: int) -> None:
        w_now = self._waveform_tensor[n]
        target = fields_dict[self._component]
        if self._plane == "xy":
            target[..., :, :, self._position] += w_now
        elif self._plane == "xz":
            target[..., :, self._position, :] += w_now
        else:
            target[..., self._position, :, :] += w_now

    def cells(self, Nx: int, Ny: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """2D cells ``(i, j)`` and amplitudes (1.0) this plane touches."""
        if self._plane == "xy":
            I, J = torch.meshgrid(torch.arange(Nx), torch.arange(Ny), indexing="ij")
            i, j = I.reshape(-1), J.reshape(-1)
        elif self._plane == "xz":
            i = torch.arange(Nx)
            j = torch.full((Nx,), self._position)
        else:
            i = torch.full((Ny,), self._position)
            j = torch.arange(Ny)
        return i, j, torch.ones(i.numel(), dtype=torch.float32)

    def cells3d(self, Nx: int, Ny: int, Nz: int):
        """Cells ``(i, j, k)`` and amplitudes (1.0) this plane touches on an
        ``(Nx, Ny, Nz)`` grid (``xy``: all (i, j) at k = position; ``xz``: all
        (i, k) at j; ``yz``: all (j, k) at i)."""
        p = self._position
        if self._plane == "xy":
            I, J = torch.meshgrid(torch.arange(Nx), torch.arange(Ny), indexing="ij")
            i, j = I.reshape(-1), J.reshape(-1)
            k = torch.full_like(i, p)
        elif self._plane == "xz":
            I, K = torch.meshgrid(torch.arange(Nx), torch.arange(Nz), indexing="ij")
            i, k = I.reshape(-1), K.reshape(-1)
            j = torch.full_like(i, p)
        else:
            J, K = torch.meshgrid(torch.arange(Ny), torch.arange(Nz), indexing="ij")
            j, k = J.reshape(-1), K.reshape(-1)
            i = torch.full_like(j, p)
        return i, j, k, torch.ones(i.numel(), dtype=torch.float32)

    n_cells = property(lambda self: self._n_cells)
    waveform_memory_bytes = property(lambda self: self._N_steps * self._waveform_tensor.element_size())
    is_precomputed = property(lambda self: True)
    component = property(lambda self: self._component)
    plane = property(lambda self: self._plane)
    position = property(lambda self: self._position)


def remember_host_cells(src, i, j, k, amp) -> None:
    """Keep host copies of a source's cell indices and amplitudes with the
    device tensors they equal (``_i_idx``, ``_j_idx``, ``_k_idx``, ``_amp``)
    and those tensors' version counters: the kernels' source tables are built
    from these copies, with no read-back, while the tensors are the same
    objects and have not been written in place since."""
    ts = (src._i_idx, src._j_idx, src._k_idx, src._amp)
    src._host_cells = (ts, tuple(t._version for t in ts),
                       tuple(np.asarray(v) for v in (i, j, k, amp)))


def host_cells(src):
    """``(i, j, k, amp)`` of a source as NumPy arrays: the remembered host
    copies when still valid, else read from the device tensors (and kept)."""
    ts = (src._i_idx, src._j_idx, getattr(src, "_k_idx", None), src._amp)
    rec = getattr(src, "_host_cells", None)
    if rec is not None and all(a is b for a, b in zip(rec[0], ts)) and \
            tuple(t._version for t in ts) == rec[1]:
        return rec[2]
    i, j, a = ts[0].cpu().numpy(), ts[1].cpu().numpy(), ts[3].detach().cpu().numpy()
    k = np.zeros_like(i) if ts[2] is None else ts[2].cpu().numpy()
    if ts[2] is not None:
        src._host_cells = (ts, tuple(t._version for t in ts), (i, j, k, a))
    return i, j, k, a


class PointSource:
    """Single-cell soft source (``src/core/sources.py:774-887``)."""

    def __init__(self, waveform: Waveform, i: int, j: int, component: str = "Hz", *, k: int = 0, grid,
                 N_steps: int) -> None:
        _check_component(component)
        if not (0 <= i < grid.Nx):
            raise ValueError(f"i={i} out of range [0, {grid.Nx})")
        if not (0 <= j < grid.Ny):
            raise ValueError(f"j={j} out of range [0, {grid.Ny})")
        if not (0 <= k < grid.Nz):
            raise ValueError(f"k={k} out of range [0, {grid.Nz})")
        self._component, self._N_steps = component, N_steps
        self._waveform = waveform
        self._waveform_tensor = waveform.build(N_steps, grid.dt, grid.device, grid.dtype)
        dev = grid.device
        self._i_idx = torch.tensor([i], dtype=torch.long, device=dev)
        self._j_idx = torch.tensor([j], dtype=torch.long, device=dev)
        self._k_idx = torch.tensor([k], dtype=torch.long, device=dev)
        self._amp = torch.tensor([1.0], dtype=grid.dtype, device=dev)
        remember_host_cells(self, np.array([i]), np.array([j]), np.array([k]), np.array([1.0]))

    n_cells = property(lambda self: 1)
    waveform_memory_bytes = property(lambda self: self._N_steps * self._waveform_tensor.element_size())
    is_precomputed = property(lambda self: True)
    component = property(lambda self: self._component)

    def step(self, fields_dict, n: int) -> None:
        values = self._waveform_tensor[n:n + 1] * self._amp
        target = fields_dict[self._component]
        target[..., self._i_idx, self._j_idx, self._k_idx] += values


def _validate_range(name: str, start: int, stop: int, limit: int) -> None:
    if start < 0 or stop > limit:
        raise ValueError(f"{name} range [{start}, {stop}) is out of grid bounds [0, {limit})")


def _validate_scalar_index(name: str, index: int, limit: int) -> None:
    if not (0 <= index < limit):
        raise ValueError(f"{name}={index} is out of grid bounds [0, {limit})")


class LineSource:
    """Soft source along a grid axis (``src/core/sources.py:895-1035``)."""

    def __init__(self, waveform: Waveform, axis: str, position: int, start: int, stop: int,
                 component: str = "Hz", *, grid, N_steps: int,
                 amplitude_profile: Union[torch.Tensor, list, None] = None) -> None:
        if axis not in ("x", "y"):
            raise ValueError(f"axis must be 'x' or 'y', got {axis!r}")
        _check_component(component)
        if start >= stop:
            raise ValueError(f"start ({start}) must be less than stop ({stop})")
        n_cells = stop - start
        dev = grid.device
        if axis == "x":
            _validate_range("i", start, stop, grid.Nx)
            _validate_scalar_index("j (position)", position, grid.Ny)
            i_vals = torch.arange(start, stop, dtype=torch.long, device=dev)
            j_vals = torch.full((n_cells,), position, dtype=torch.long, device=dev)
        else:
            _validate_range("j", start, stop, grid.Ny)
            _validate_scalar_index("i (position)", position, grid.Nx)
            i_vals = torch.full((n_cells,), position, dtype=torch.long, device=dev)
            j_vals = torch.arange(start, stop, dtype=torch.long, device=dev)
        self._component, self._N_steps, self._n_cells = component, N_steps, n_cells
        self._waveform = waveform
        self._waveform_tensor = waveform.build(N_steps, grid.dt, grid.device, grid.dtype)
        self._i_idx, self._j_idx = i_vals, j_vals
        self._k_idx = torch.zeros(n_cells, dtype=torch.long, device=dev)
        if amplitude_profile is None:
            self._amp = torch.ones(n_cells, dtype=grid.dtype, device=dev)
        else:
            amp = torch.as_tensor(amplitude_profile, dtype=grid.dtype, device=dev)
            if amp.shape != (n_cells,):
                raise ValueError(f"amplitude_profile must have length {n_cells}, got shape {tuple(amp.shape)}")
            self._amp = amp
        span = np.arange(start, stop)
        fixed = np.full(n_cells, position)
        hi, hj = (span, fixed) if axis == "x" else (fixed, span)
        if amplitude_profile is None:
            remember_host_cells(self, hi, hj, np.zeros(n_cells, dtype=np.int64), np.ones(n_cells))
        elif not isinstance(amplitude_profile, torch.Tensor) or amplitude_profile.device.type == "cpu":
            # the profile's values as the amplitude tensor holds them (rounded to the grid dtype)
            a_host = torch.as_tensor(amplitude_profile, dtype=grid.dtype).numpy()
            remember_host_cells(self, hi, hj, np.zeros(n_cells, dtype=np.int64), a_host)

    n_cells = property(lambda self: self._n_cells)
    waveform_memory_bytes = property(lambda self: self._N_steps * self._waveform_tensor.element_size())
    is_precomputed = property(lambda self: True)
    component = property(lambda self: self._component)

    def step(self, fields_dict, n: int) -> None:
        values = self._waveform_tensor[n:n + 1] * self._amp
        target = fields_dict[self._component]
        target[..., self._i_idx, self._j_idx, self._k_idx] += values


class SourceCollection:
    """Ordered sources, stepped together (``src/core/sources.py:1043-1090``)."""

    def __init__(self, sources: list) -> None:
        self._sources = list(sources)

    def step(self, fields_dict, n: int) -> None:
        for src in self._sources:
            src.step(fields_dict, n)

    @property
    def total_vram_bytes(self) -> int:
        return sum(s.waveform_memory_bytes for s in self._sources)

    @property
    def sources_graph_compatible(self) -> bool:
        return all(s.is_precomputed for s in self._sources)

    @property
    def n_injection_cells_total(self) -> int:
        return sum(s.n_cells for s in self._sources)
