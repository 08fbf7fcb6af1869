"""MIMO array geometry and DAS backprojection on the GPU.

The reference keeps imaging in its example scripts; this module brings the
same functions into ``core``:

* ``get_antenna_positions`` — ``examples/brain_mimo_imaging.py:109-142``
  (ring positions with Python ``round``, clamped to ``[0, N-1]``); ``clamp``
  selects the ``[1, N-2]`` clamp of ``src/datasets/brain/antenna.py:95-96``.
* ``das_backprojection`` — ``examples/brain_mimo_imaging.py:211-266`` with the
  variants as keywords: ``gate_start`` (time gate,
  ``examples/09_brain_clot_dataset_sample.py:176-178``), ``v_medium`` (other
  effective permittivity, ``examples/10_breast_tumor_mimo.py:148-165``) and
  ``out_of_range='zero'`` (drop out-of-window samples,
  ``src/datasets/brain/antenna.py:333-336``). It runs the ``wf_das2d`` HIP
  kernel and returns the same fp64 power image, bit for bit.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _native
from ..engine import require_gpu_tensor
from .grid import C0


def get_antenna_positions(n: int, radius: float, center, grid, clamp: str = "grid"):
    """``n`` integer cells on a circle of ``radius`` cells around ``center``,
    clamped to ``[0, N-1]`` (``clamp="grid"``) or ``[1, N-2]`` (``"inner"``)."""
    if clamp not in ("grid", "inner"):
        raise ValueError(f"clamp must be 'grid' or 'inner', got {clamp!r}")
    m = 0 if clamp == "grid" else 1
    lo, hi, lo_j, hi_j = m, grid.Nx - 1 - m, m, grid.Ny - 1 - m
    out = []
    for k in range(n):
        angle = 2.0 * math.pi * k / n
        i = int(round(center[0] + radius * math.cos(angle)))
        j = int(round(center[1] + radius * math.sin(angle)))
        out.append((max(lo, min(i, hi)), max(lo_j, min(j, hi_j))))
    return out


def das_backprojection(
    delta_S,
    antenna_positions,
    grid,
    n_steps: int,
    image_shape: tuple[int, int] | None = None,
    *,
    v_medium: float | None = None,
    gate_start: int = 0,
    out_of_range: str = "clip",
    device=None,
    return_tensor: bool = False,
):
    """Delay-and-sum power image ``(Nx_img, Ny_img)`` in fp64.

    ``delta_S`` is ``(N_tx, N_rx, T)`` fp32 (NumPy or tensor); pixels are at
    ``u * (Nx*dx/Nx_img)``, antennas at ``(i*dx, j*dx)``; the default speed is
    ``C0/sqrt(40)`` as in the reference brain example.
    """
    if out_of_range not in ("clip", "zero"):
        raise ValueError(f"out_of_range must be 'clip' or 'zero', got {out_of_range!r}")
    Nx_img, Ny_img = (grid.Nx, grid.Ny) if image_shape is None else image_shape
    if v_medium is None:
        v_medium = C0 / math.sqrt(40.0)
    dev = torch.device(device) if device is not None else (
        delta_S.device if isinstance(delta_S, torch.Tensor) else grid.device)
    S = torch.as_tensor(delta_S, dtype=torch.float32).to(dev).contiguous()
    require_gpu_tensor(S, "traces")
    n_tx, n_rx, n_t = S.shape
    if n_t != n_steps:
        S = S[:, :, :n_steps].contiguous()
        n_t = S.shape[2]
    dx = grid.dx
    ant = np.asarray(antenna_positions, dtype=np.float64)
    if ant.shape[0] < max(n_tx, n_rx):
        raise ValueError("fewer antenna positions than TX/RX channels")
    ax = torch.tensor([p[0] * dx for p in antenna_positions], dtype=torch.float64, device=dev)
    ay = torch.tensor([p[1] * dx for p in antenna_positions], dtype=torch.float64, device=dev)
    image = torch.empty((Nx_img, Ny_img), dtype=torch.float64, device=dev)
    run_das(S, ax, ay, image, pitch_x=grid.Nx * dx / Nx_img, pitch_y=grid.Ny * dx / Ny_img,
            v_medium=v_medium, dt=grid.dt, gate_start=gate_start, clip=out_of_range == "clip")
    if return_tensor:
        return image
    return image.cpu().numpy()


def run_das(S: torch.Tensor, ax: torch.Tensor, ay: torch.Tensor, image: torch.Tensor, *, pitch_x: float,
            pitch_y: float, v_medium: float, dt: float, gate_start: int = 0, clip: bool = True) -> None:
    """Launch ``wf_das2d`` on device tensors (traces fp32 ``(n_tx, n_rx, n_t)``,
    antenna coordinates fp64, image fp64) on the current stream."""
    a = das_args(S, ax, ay, image, pitch_x=pitch_x, pitch_y=pitch_y, v_medium=v_medium, dt=dt,
                 gate_start=gate_start, clip=clip)
    lib = _native.load()
    stream = _native.stream_handle(S.device)
    _native.check(lib.wf_das2d(ctypes.byref(a), ctypes.c_void_p(stream)), "wf_das2d")


def das_args(S: torch.Tensor, ax: torch.Tensor, ay: torch.Tensor, image: torch.Tensor, *, pitch_x: float,
             pitch_y: float, v_medium: float, dt: float, gate_start: int = 0, clip: bool = True):
    """The ``wf_das2d_args`` of one image (device tensors as ``run_das`` takes)."""
    n_tx, n_rx, n_t = S.shape
    a = _native.WfDas2dArgs()
    a.traces, a.ant_x, a.ant_y, a.image = S.data_ptr(), ax.data_ptr(), ay.data_ptr(), image.data_ptr()
    a.n_tx, a.n_rx, a.n_t = n_tx, n_rx, n_t
    a.nx_img, a.ny_img = image.shape
    a.gate_start = int(gate_start)
    a.clip = 1 if clip else 0
    a.pitch_x, a.pitch_y = float(pitch_x), float(pitch_y)
    a.v_medium, a.dt = float(v_medium), float(dt)
    return a


class DasBatch:
    """Several DAS images formed together (``wf_das2d_batch``, up to 32 images
    a launch): ``add`` takes the arguments of ``run_das``, ``launch`` forms
    every image added on the current stream. Each image equals what
    ``run_das`` gives for the same arguments, bit for bit. The batch holds
    the added tensors until ``launch`` (the descriptors carry only their
    addresses), so a caller may drop its own references in between."""

    def __init__(self, device) -> None:
        self.device = torch.device(device)
        self._args = []
        self._keep = []

    def add(self, S, ax, ay, image, **kw) -> None:
        self._args.append(das_args(S, ax, ay, image, **kw))
        self._keep.append((S, ax, ay, image))

    def __len__(self) -> int:
        return len(self._args)

    def launch(self) -> None:
        if not self._args:
            return
        arr = (_native.WfDas2dArgs * len(self._args))(*self._args)
        lib = _native.load()
        stream = _native.stream_handle(self.device)
        _native.check(lib.wf_das2d_batch(arr, len(self._args), ctypes.c_void_p(stream)), "wf_das2d_batch")
        # launched on the current stream: memory freed from here on is reused only by later work on it
        self._args, self._keep = [], []
