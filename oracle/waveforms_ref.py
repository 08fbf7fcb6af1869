"""Restatement of the reference waveform tables (the oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``build(N, dt)`` follows ``Waveform.build`` (``src/core/sources.py:112-138``):
``t = arange(N, fp32) * dt`` in fp32, then each class's ``_compute`` with
Python-float scalars and fp32 torch ops, on the CPU.
"""

from __future__ import annotations

import math

import torch


def time_axis(n_steps: int, dt: float, dtype=torch.float32) -> torch.Tensor:
    """``torch.arange(N, dtype) * dt`` (``src/core/sources.py:137``)."""
    return torch.arange(n_steps, dtype=dtype) * dt


def gaussian(t, amplitude, t0, sigma):
    """``GaussianPulse._compute`` (``src/core/sources.py:238-241``)."""
    exponent = -((t - t0) ** 2) / (2.0 * sigma ** 2)
    return amplitude * torch.exp(exponent)


def sinusoidal(t, amplitude, frequency, phase=0.0):
    """``SinusoidalSource._compute`` (``src/core/sources.py:316-321``)."""
    tau_ramp = 3.0 / (2.0 * math.pi * frequency)
    omega = 2.0 * math.pi * frequency
    carrier = torch.sin(omega * t + phase)
    ramp = 1.0 - torch.exp(-t / tau_ramp)
    return amplitude * carrier * ramp


def ricker(t, amplitude, peak_freq, t0):
    """``RickerWavelet._compute`` (``src/core/sources.py:399-403``)."""
    pi2_fp2 = (math.pi * peak_freq) ** 2
    u = pi2_fp2 * (t - t0) ** 2
    return amplitude * (1.0 - 2.0 * u) * torch.exp(-u)


def modulated_gaussian(t, amplitude, fc, sigma, t0, phi=0.0):
    """``ModulatedGaussian._compute`` (``src/core/sources.py:465-468``)."""
    carrier = torch.sin(2.0 * math.pi * fc * t + phi)
    envelope = torch.exp(-((t - t0) ** 2) / (2.0 * sigma ** 2))
    return amplitude * carrier * envelope


def chirp(t, amplitude, f_start, f_end, t_start, t_end, phi=0.0):
    """``Chirp._compute`` (``src/core/sources.py:524-531``)."""
    duration = t_end - t_start
    in_window = (t >= t_start) & (t <= t_end)
    tau = t - t_start
    k = (f_end - f_start) / (2.0 * duration)
    inst_phase = 2.0 * math.pi * (f_start * tau + k * tau ** 2) + phi
    hann = 0.5 * (1.0 - torch.cos(2.0 * math.pi * tau / duration))
    result = amplitude * torch.sin(inst_phase) * hann
    return torch.where(in_window, result, torch.zeros_like(result))


def uwb(t, amplitude, f_low, f_high, t0=None):
    """``UWBPulse._compute`` (``src/core/sources.py:628-633``; sigma and
    norm from ``:609-612``)."""
    fc = (f_low + f_high) / 2.0
    sigma = 1.0 / (math.sqrt(2.0) * math.pi * fc)
    norm = sigma * math.sqrt(math.e)
    if t0 is None:
        t0 = 6.0 * sigma
    tau = t - t0
    g = torch.exp(-(tau ** 2) / (2.0 * sigma ** 2))
    raw = -(tau / sigma ** 2) * g
    return amplitude * norm * raw


def gaussian_derivative(t, amplitude, t0, width):
    """Gaussian-derivative pulse used by the 256x256 slab workload:
    ``-A*((t-t0)/w)*exp(-(t-t0)^2/(2 w^2))``. The reference has no class with
    this exact normalisation; the product adds one as a ``Waveform`` subclass
    (the ABC is open for extension, ``src/core/sources.py:64-106``)."""
    tau = t - t0
    return amplitude * (-(tau / width)) * torch.exp(-(tau ** 2) / (2.0 * width ** 2))
