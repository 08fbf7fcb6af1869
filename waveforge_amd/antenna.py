"""Antenna ring array and its envelope DAS image (2D), on the HIP kernels.

Mirrors ``AntennaRing`` of ``src/datasets/brain/antenna.py`` (constructor and
validation :60-106, positions :113-123, ``build_sources`` :129-161,
``reset``/``record``/``get_signals`` :167-212) for 2D grids (``z_plane`` 0,
``Nz`` 1), with ``compute_das_image`` (:262-333) running on the GPU through
``wf_das_envelope``: Hilbert envelope of every trace, the layered-medium
(or free-space) travel-time table, and the envelope delay-and-sum. The
result is the reference's fp32 ``(image_size, image_size)`` power image.

The analytic signal's imaginary part is a circular convolution with a
kernel that depends only on the record length (SciPy's spectrum mask
1, 2, ..., 2, [1], 0, ...); ``hilbert_kernel`` builds it once per length on
the host in fp64 and the kernel applies it in fp64. SciPy computes the same
envelope with a single-precision FFT, so envelopes agree to fp32 rounding,
not bit for bit; the travel-time table and the accumulation are the
reference's own arithmetic.
"""

from __future__ import annotations

import ctypes
import math
from functools import lru_cache
from typing import Optional

import numpy as np
import torch

from . import _native
from .core.sources import PointSource, SourceCollection, Waveform


@lru_cache(maxsize=32)
def _hilbert_kernel_np(n: int) -> np.ndarray:
    h = np.zeros(n)
    if n % 2 == 0:
        h[0] = h[n // 2] = 1.0
        h[1:n // 2] = 2.0
    else:
        h[0] = 1.0
        h[1:(n + 1) // 2] = 2.0
    k = np.fft.ifft(h).imag.copy()
    k.setflags(write=False)
    return k


def hilbert_kernel(n: int, device) -> torch.Tensor:
    """fp64 ``k`` with ``Im(hilbert(x)) = x (circularly convolved with) k`` for length ``n``."""
    return torch.from_numpy(np.array(_hilbert_kernel_np(int(n)))).to(device)


class EnvelopeDAS:
    """Envelope DAS for one array and image geometry: the travel-time table is
    computed once (stage 2) and reused for every signal set.

    ``positions`` are antenna cells ``[(i, j), ...]`` of a grid ``(Nx, Ny)``;
    ``head_geometry`` is ``{'cx_cells', 'cy_cells', 'layers': [(r_cells,
    eps_r), ...]}`` outermost first, as ``AntennaRing.compute_das_image``
    takes it; None means free space."""

    def __init__(self, positions, grid_shape, dx: float, dt: float, image_size: int = 40,
                 head_geometry: Optional[dict] = None, device=None) -> None:
        dev = torch.device(device) if device is not None else torch.device("cuda")
        if dev.type != "cuda":
            raise RuntimeError("waveforge_amd: the envelope DAS runs on the ROCm GPU; there is no CPU path")
        _native.load()
        self.device, self.dt, self.image_size = dev, float(dt), int(image_size)
        self.n_ant = len(positions)
        Nx, Ny = grid_shape
        # host geometry exactly as the reference forms it (mm)
        x_px = np.linspace(0, (Nx - 1) * dx * 1e3, image_size)
        y_px = np.linspace(0, (Ny - 1) * dx * 1e3, image_size)
        ants = [(i * dx * 1e3, j * dx * 1e3) for i, j in positions]
        f64 = dict(dtype=torch.float64, device=dev)
        self._ax = torch.tensor([a[0] for a in ants], **f64)
        self._ay = torch.tensor([a[1] for a in ants], **f64)
        self._px = torch.from_numpy(x_px).to(dev)
        self._py = torch.from_numpy(y_px).to(dev)
        a = _native.WfDasEnvArgs()
        self._layers = ()
        if head_geometry is not None:
            a.cx = head_geometry["cx_cells"] * dx * 1e3
            a.cy = head_geometry["cy_cells"] * dx * 1e3
            lr = torch.tensor([r * dx * 1e3 for r, _ in head_geometry["layers"]], **f64)
            le = torch.tensor([float(e) for _, e in head_geometry["layers"]], **f64)
            self._layers = (lr, le)
            a.layer_r, a.layer_eps, a.n_layers = lr.data_ptr(), le.data_ptr(), len(head_geometry["layers"])
        self.tt = torch.empty((self.n_ant, image_size, image_size), dtype=torch.float32, device=dev)
        a.tt = self.tt.data_ptr()
        a.ant_x, a.ant_y, a.px, a.py = (t.data_ptr() for t in (self._ax, self._ay, self._px, self._py))
        a.dt = self.dt
        a.n_ant, a.n_t, a.img_w, a.img_h = self.n_ant, 1, image_size, image_size
        self._args = a
        self._launch(2)

    def _launch(self, stages: int) -> None:
        self._args.stages = stages
        lib = _native.load()
        stream = ctypes.c_void_p(_native.stream_handle(self.device))
        _native.check(lib.wf_das_envelope(ctypes.byref(self._args), stream), "wf_das_envelope")

    def _signals(self, signals) -> torch.Tensor:
        S = torch.as_tensor(signals).to(device=self.device, dtype=torch.float32).contiguous()
        if S.dim() != 3 or S.shape[0] != self.n_ant or S.shape[1] != self.n_ant or S.shape[2] < 1:
            raise ValueError(f"signals must be (N_ant, N_ant, N_steps) with N_ant = {self.n_ant}, "
                             f"got {tuple(S.shape)}")
        return S

    def envelope(self, signals) -> torch.Tensor:
        """``|hilbert(signals, axis=-1)|`` fp32 on the device (stage 1)."""
        S = self._signals(signals)
        kern = hilbert_kernel(S.shape[2], self.device)
        env = torch.empty_like(S)
        a = self._args
        a.traces, a.hilbert_k, a.env, a.n_t = S.data_ptr(), kern.data_ptr(), env.data_ptr(), S.shape[2]
        self._launch(1)
        self._keep = (S, kern)
        return env

    def image_from_envelope(self, env: torch.Tensor) -> torch.Tensor:
        """The delay-and-sum power image of an envelope set (stage 4)."""
        env = self._signals(env)
        img = torch.empty((self.image_size, self.image_size), dtype=torch.float32, device=self.device)
        a = self._args
        a.env, a.image, a.n_t = env.data_ptr(), img.data_ptr(), env.shape[2]
        self._launch(4)
        self._keep_img = env
        return img

    def __call__(self, signals) -> torch.Tensor:
        S = self._signals(signals)
        kern = hilbert_kernel(S.shape[2], self.device)
        env = torch.empty_like(S)
        img = torch.empty((self.image_size, self.image_size), dtype=torch.float32, device=self.device)
        a = self._args
        a.traces, a.hilbert_k, a.env, a.image, a.n_t = (S.data_ptr(), kern.data_ptr(), env.data_ptr(),
                                                        img.data_ptr(), S.shape[2])
        self._launch(1 | 4)
        self._keep = (S, kern, env)
        return img


def envelope_das(signals, positions, grid_shape, dx: float, dt: float, image_size: int = 40,
                 head_geometry: Optional[dict] = None, device=None, return_tensor: bool = False):
    """Envelope DAS power image ``(image_size, image_size)`` fp32 of
    ``signals`` ``(N_ant, N_ant, N_steps)`` (one-shot form of ``EnvelopeDAS``)."""
    dev = device if device is not None else (signals.device if isinstance(signals, torch.Tensor) else "cuda")
    img = EnvelopeDAS(positions, grid_shape, dx, dt, image_size, head_geometry, device=dev)(signals)
    return img if return_tensor else img.cpu().numpy()


class AntennaRing:
    """Uniform circular array of ``n_elements`` transmit/receive elements (2D)."""

    def __init__(self, n_elements: int, ring_radius_cells: int, z_plane: int, grid, centre=None,
                 component: str = "Ez") -> None:
        if n_elements < 2:
            raise ValueError(f"n_elements must be >= 2, got {n_elements}")
        if not (0 < ring_radius_cells < min(grid.Nx, grid.Ny) // 2):
            raise ValueError(f"ring_radius_cells={ring_radius_cells} out of range for grid ({grid.Nx}x{grid.Ny})")
        if not (0 <= z_plane < grid.Nz):
            raise ValueError(f"z_plane={z_plane} out of range [0, {grid.Nz})")
        self._n, self._r, self._z = n_elements, ring_radius_cells, z_plane
        self._grid, self._component = grid, component
        cx = centre[0] if centre else grid.Nx // 2
        cy = centre[1] if centre else grid.Ny // 2
        self._cx, self._cy = cx, cy
        self._positions: list[tuple[int, int, int]] = []
        for k in range(n_elements):
            angle = 2.0 * math.pi * k / n_elements
            i = int(round(cx + ring_radius_cells * math.cos(angle)))
            j = int(round(cy + ring_radius_cells * math.sin(angle)))
            i = max(1, min(grid.Nx - 2, i))
            j = max(1, min(grid.Ny - 2, j))
            self._positions.append((i, j, z_plane))
        self._N_steps = 0
        self._signals: Optional[np.ndarray] = None

    n_elements = property(lambda self: self._n)
    positions = property(lambda self: list(self._positions))

    @property
    def positions_mm(self):
        d = self._grid.dx * 1e3
        return [(i * d, j * d, k * d) for i, j, k in self._positions]

    def build_sources(self, waveform: Waveform, tx_idx: int, N_steps: int) -> SourceCollection:
        if not (0 <= tx_idx < self._n):
            raise ValueError(f"tx_idx={tx_idx} out of range [0, {self._n})")
        self._N_steps = N_steps
        if self._signals is None:
            self._signals = np.zeros((self._n, self._n, N_steps), dtype=np.float32)
        i, j, k = self._positions[tx_idx]
        return SourceCollection([PointSource(waveform, i, j, self._component, k=k, grid=self._grid,
                                             N_steps=N_steps)])

    def reset(self) -> None:
        if self._signals is not None:
            self._signals[:] = 0.0

    def record(self, field: torch.Tensor, step: int, tx_idx: int) -> None:
        """Record the field at every RX position for ``step`` (one gather per call)."""
        if self._signals is None:
            raise RuntimeError("Call build_sources() before record().")
        if step >= self._N_steps:
            return
        ii = torch.tensor([p[0] for p in self._positions], device=field.device)
        jj = torch.tensor([p[1] for p in self._positions], device=field.device)
        kk = torch.tensor([p[2] for p in self._positions], device=field.device)
        self._signals[tx_idx, :, step] = field[ii, jj, kk].detach().cpu().numpy()

    def set_signals(self, signals) -> None:
        """Install a whole ``(N_tx, N_rx, N_steps)`` signal set (e.g. from ``run_mimo``)."""
        s = signals.detach().cpu().numpy() if isinstance(signals, torch.Tensor) else np.asarray(signals)
        if s.shape[:2] != (self._n, self._n):
            raise ValueError(f"signals must be ({self._n}, {self._n}, N_steps), got {s.shape}")
        self._signals = s.astype(np.float32, copy=True)
        self._N_steps = s.shape[2]

    def get_signals(self) -> np.ndarray:
        if self._signals is None:
            raise RuntimeError("No signals recorded yet.")
        return self._signals.copy()

    def compute_das_image(self, signals, image_size: int = 40, head_geometry: Optional[dict] = None) -> np.ndarray:
        """Envelope DAS power image (``antenna.py:262-333``), on the GPU."""
        if signals.shape[-1] != self._N_steps:
            raise ValueError(f"signals have {signals.shape[-1]} samples, the ring records {self._N_steps}")
        return envelope_das(signals, [(i, j) for i, j, _ in self._positions], (self._grid.Nx, self._grid.Ny),
                            self._grid.dx, self._grid.dt, image_size, head_geometry,
                            device=self._grid.device)

    def __repr__(self) -> str:
        return (f"AntennaRing(n={self._n}, r={self._r} cells, centre=({self._cx},{self._cy}), "
                f"z={self._z}, component={self._component!r})")
