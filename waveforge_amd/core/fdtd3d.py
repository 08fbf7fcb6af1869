"""3D full-vector Yee stepper on HIP kernels (drop-in for ``core.fdtd3d``).

Same constructor, validation, CFL warning, attributes, telemetry, errors and
properties as ``src/core/fdtd3d.py`` (``FDTD3D`` :118-580,
``SimulationDivergedError`` :95-112). ``step()`` is one reference step as a
short sequence of HIP launches (``wf_fdtd3d_step``: Mur snapshot planes,
sources, Faraday, Mur faces in the reference's order, Ampere, detectors);
``run(n)`` advances in chunks that end on the ``n_check`` steps, where the
six-component stability check runs exactly as the reference's. The steps of
a chunk take the fused path by default (``wf_fdtd3d_fused_step``: one
streaming pass per step that reads the ``FieldSet`` tensors' state and writes
a second state buffer, then the reverse, same values bit for bit), with one
in-place step first when the chunk is odd, so results always land in the
``FieldSet`` tensors; it moves about 68 B per cell-update against the
per-stage pair's 90 and runs about 1.4 times as fast at 256^3. It holds a
second set of six field buffers; ``fused=False`` keeps every step in place.

The ``FieldSet`` tensors are the live state (read and written in place);
``sim._sources`` is read at every call. ``compile_step`` exists for API
compatibility and does nothing: the step is already a compiled kernel
sequence.
"""

from __future__ import annotations

import ctypes
import math
import time
import warnings
from typing import Optional

import numpy as np
import torch

from .. import _native
from ..engine import SourceTable, check_grid_cells, f32, maxabs, require_gpu_tensor
from .detectors import DetectorSet
from .fields import FieldSet
from .grid import YeeGrid

EPS0: float = 8.8541878128e-12
MU0: float = 1.2566370614e-6
_C0: float = 1.0 / math.sqrt(EPS0 * MU0)


def snapshot_is_state(table: SourceTable, Nx: int, Ny: int, Nz: int) -> bool:
    """True when no H-component source cell lies on a Mur plane of its
    component (index 0, 1, n-2 or n-1 across either tangential axis: Hx y and
    z, Hy x and z, Hz x and y; ``boundaries.py:495-504``). The snapshot, taken
    before the sources are added, then equals the state the fused pass reads,
    and the fused step needs no snapshot copy."""
    if not table.n_grp:
        return True
    comp = table.host_inj["comp"]
    cell = table.host_inj["cell"].astype(np.int64)
    k = cell % Nz
    j = (cell // Nz) % Ny
    i = cell // (Ny * Nz)

    def on(v, n):
        return (v <= 1) | (v >= n - 2)

    zf = on(k, Nz) if Nz > 1 else np.zeros(cell.shape, dtype=bool)
    c = _native.WF_COMP3
    bad = ((comp == c["Hx"]) & (on(j, Ny) | zf)) | ((comp == c["Hy"]) & (on(i, Nx) | zf)) | \
          ((comp == c["Hz"]) & (on(i, Nx) | on(j, Ny)))
    return not bool(bad.any())


class SimulationDivergedError(RuntimeError):
    """Raised when a field component exceeds the stability threshold."""

    def __init__(self, message: str) -> None:
        super().__init__(message)


class FDTD3D:
    """3D Yee leapfrog stepper ({Ex, Ey, Ez, Hx, Hy, Hz}) on gfx950 HIP kernels."""

    def __init__(self, grid: YeeGrid, fields: FieldSet, boundary: object, sources: Optional[object] = None, *,
                 Ca: Optional[torch.Tensor] = None, Cb: Optional[torch.Tensor] = None,
                 stability_threshold: float = 1e10, n_check: int = 100,
                 detectors: Optional[DetectorSet] = None, fused: bool = True) -> None:
        if grid.Nx < 4:
            raise ValueError(f"grid.Nx must be >= 4 for a valid 3D Yee stencil, got {grid.Nx}")
        if grid.Ny < 4:
            raise ValueError(f"grid.Ny must be >= 4 for a valid 3D Yee stencil, got {grid.Ny}")
        if grid.Nz > 1 and grid.Nz < 4:
            raise ValueError(
                f"grid.Nz must be >= 4 when using 3D mode, got {grid.Nz}. Use Nz=1 for 2D-equivalent mode."
            )
        check_grid_cells(grid.Nx * grid.Ny, grid.Nz)
        self._grid, self._fields, self._boundary, self._sources = grid, fields, boundary, sources
        self._Dh: float = grid.dt / MU0
        self._De: float = grid.dt / EPS0
        self._dx, self._dy, self._dz = grid.dx, grid.dy, grid.dz
        if (Ca is None) != (Cb is None):
            raise ValueError("Ca and Cb must be provided together or both omitted.")
        self._Ca, self._Cb = Ca, Cb
        self._has_materials = Ca is not None
        dx, dy, dz = grid.dx, grid.dy, grid.dz
        cfl_limit = 0.99 / (_C0 * math.sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)))
        if grid.dt > cfl_limit:
            warnings.warn(
                f"grid.dt={grid.dt:.6e} exceeds 0.99 × CFL limit {cfl_limit:.6e}. Simulation may be unstable.",
                UserWarning,
                stacklevel=2,
            )
        self._threshold = stability_threshold
        self._n_check = n_check
        self._detectors = detectors
        self._fused = fused
        self._second = None
        self.steps_completed: int = 0
        self.elapsed_time: float = 0.0
        self.mcells_per_second: float = 0.0
        self.last_field_max: float = -1.0
        self.n_stability_checks: int = 0
        self._key = None
        self._structs = None
        self._src_table = None

    # ------------------------------------------------------------------
    def _source_list(self):
        if self._sources is None:
            return []
        return list(getattr(self._sources, "_sources", self._sources))

    def _members(self, t: torch.Tensor) -> list[torch.Tensor]:
        g = self._grid
        if t.dim() == 4:
            return [t[b] for b in range(t.shape[0])]
        return [t.reshape(g.Nx, g.Ny, g.Nz)]

    def _build(self):
        f = self._fields
        comps = (f.Ex, f.Ey, f.Ez, f.Hx, f.Hy, f.Hz)
        for t in comps:
            require_gpu_tensor(t, "field")
        srcs = self._source_list()
        det = self._detectors
        key = (tuple(t.data_ptr() for t in comps), None if self._Ca is None else (self._Ca.data_ptr(), self._Cb.data_ptr()),
               id(self._sources), tuple(id(s) for s in srcs), id(det))
        if key == self._key:
            return self._structs
        _native.load()
        g = self._grid
        dev = f.Hz.device
        if self._Ca is not None:
            require_gpu_tensor(self._Ca, "Ca")
            require_gpu_tensor(self._Cb, "Cb")
        self._src_table = SourceTable(srcs, g.Nx, g.Ny, dev, Nz=g.Nz, comps_map=_native.WF_COMP3)
        members = [self._members(t) for t in comps]
        B = len(members[0])
        for m in members:
            for t in m:
                if not t.is_contiguous():
                    raise ValueError("waveforge_amd: each field (batch member) must be contiguous")
        ca_m = cb_m = [None] * B
        if self._Ca is not None:
            ca, cb = self._Ca.contiguous(), self._Cb.contiguous()
            ca_m = self._members(ca) if ca.dim() == 4 else [ca.reshape(g.Nx, g.Ny, g.Nz)] * B
            cb_m = self._members(cb) if cb.dim() == 4 else [cb.reshape(g.Nx, g.Ny, g.Nz)] * B
            self._coef_keep = (ca, cb)
        lib = _native.load()
        n_strip = int(lib.wf_fdtd3d_strip_floats(g.Nx, g.Ny, g.Nz))
        second = None
        if self._fused:
            if self._second is None or self._second[0].shape != f.Hz.shape:
                self._second = tuple(torch.empty_like(f.Hz) for _ in range(6))
            second = [self._members(t) for t in self._second]
        b_ = self._boundary
        flags = _native.WF_RUN3D_SNAP_FROM_STATE if snapshot_is_state(self._src_table, g.Nx, g.Ny, g.Nz) else 0
        self._strips, structs = [], []
        for b in range(B):
            r = _native.WfRun3d()
            r.ex, r.ey, r.ez, r.hx, r.hy, r.hz = (m[b].data_ptr() for m in members)
            if second is not None:
                r.ex2, r.ey2, r.ez2, r.hx2, r.hy2, r.hz2 = (m[b].data_ptr() for m in second)
            if ca_m[b] is not None:
                r.ca, r.cb = ca_m[b].data_ptr(), cb_m[b].data_ptr()
            strip = torch.empty(n_strip, dtype=torch.float32, device=dev)
            self._strips.append(strip)
            r.pstrip = strip.data_ptr()
            t = self._src_table
            if t.n_grp:
                r.inj, r.grp_start = t.ptrs()
                r.n_grp, r.n_inj = t.n_grp, t.n_inj
            if det is not None and len(det):
                r.det_cell, r.det_comp = det.cell_index.data_ptr(), det.component_index.data_ptr()
                tr = det.traces if det.traces.dim() == 2 else det.traces[b]
                r.traces, r.trace_stride = tr.data_ptr(), tr.stride(0)
                r.n_det, r.trace_n0, r.trace_len = len(det), det.start_step, tr.shape[-1]
            r.nx, r.ny, r.nz = g.Nx, g.Ny, g.Nz
            r.dh, r.de = f32(g.dt / MU0), f32(g.dt / EPS0)
            r.dx, r.dy, r.dz = f32(g.dx), f32(g.dy), f32(g.dz)
            r.cmx = f32(getattr(b_, "C_mur_x", (_C0_REF * g.dt - g.dx) / (_C0_REF * g.dt + g.dx)))
            r.cmy = f32(getattr(b_, "C_mur_y", (_C0_REF * g.dt - g.dy) / (_C0_REF * g.dt + g.dy)))
            r.cmz = f32(getattr(b_, "C_mur_z", (_C0_REF * g.dt - g.dz) / (_C0_REF * g.dt + g.dz)))
            r.flags = flags  # read by the fused path only
            structs.append(r)
        self._structs = structs
        self._key = key
        return structs

    def _stream(self) -> int:
        return _native.stream_handle(self._fields.Hz.device)

    def _advance(self, n_steps: int) -> None:
        if n_steps <= 0:
            return
        structs = self._build()
        limit = self._src_table.min_wave_len
        if limit is not None and self.steps_completed + n_steps > limit:
            raise RuntimeError(
                f"source waveform table has {limit} samples; cannot inject at step "
                f"{max(self.steps_completed, limit)} (build sources with a larger N_steps)"
            )
        lib = _native.load()
        stream = ctypes.c_void_p(self._stream())
        first = self.steps_completed
        if self._fused and n_steps >= 2 and n_steps % 2:
            for r in structs:
                r.step = first
            arr = (_native.WfRun3d * len(structs))(*structs)
            _native.check(lib.wf_fdtd3d_step(arr, len(structs), 1, stream), "wf_fdtd3d_step")
            first += 1
            n_steps -= 1
        for r in structs:
            r.step = first
        arr = (_native.WfRun3d * len(structs))(*structs)
        if self._fused and n_steps >= 2:
            _native.check(lib.wf_fdtd3d_fused_step(arr, len(structs), int(n_steps), stream), "wf_fdtd3d_fused_step")
        else:
            _native.check(lib.wf_fdtd3d_step(arr, len(structs), int(n_steps), stream), "wf_fdtd3d_step")

    def step(self) -> None:
        """Advance all six components by one time step in place (``fdtd3d.py:262-373``)."""
        self._advance(1)
        self.steps_completed += 1
        if self.steps_completed % self._n_check == 0:
            self._check_stability()

    def compile_step(self, *, backend: str = "inductor", mode: str = "reduce-overhead") -> None:
        """No-op: the step is already a compiled HIP kernel sequence."""
        return None

    def _check_stability(self) -> None:
        """Max-abs of Ex, Ey, Ez, Hx, Hy, Hz in that order (``fdtd3d.py:390-425``)."""
        f = self._fields
        s = self._stream()
        vals = torch.cat([maxabs(f.Ex, f.Ey, f.Ez, s), maxabs(f.Hx, f.Hy, f.Hz, s)]).cpu().tolist()
        self.n_stability_checks += 1
        for name, val in zip(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"), vals):
            if val > self.last_field_max:
                self.last_field_max = val
            if val > self._threshold:
                raise SimulationDivergedError(
                    f"Simulation diverged at step {self.steps_completed}: {name} max = {val:.3e} exceeds "
                    f"threshold {self._threshold:.3e}. Check CFL condition (dt={self._grid.dt:.3e}) and "
                    f"source amplitude."
                )

    def run(self, n_steps: int, *, verbose: bool = False) -> None:
        """Run ``n_steps`` steps (``fdtd3d.py:443-515``), same telemetry."""
        is_cuda = str(self._grid.device).startswith("cuda")
        report = max(1, n_steps // 10)
        if is_cuda:
            torch.cuda.synchronize()
        t_start = time.perf_counter()
        start = self.steps_completed
        g = self._grid
        done = 0
        try:
            while done < n_steps:
                chunk = min(n_steps - done, self._n_check - self.steps_completed % self._n_check)
                if verbose:
                    chunk = min(chunk, report - done % report)
                self._build()  # the source table (cached per set-up)
                limit = self._src_table.min_wave_len
                if limit is not None and self.steps_completed + chunk > limit:
                    # step up to the table's last sample, then raise (as the
                    # reference's per-step loop does at the first step past it)
                    chunk = limit - self.steps_completed
                    if chunk <= 0:
                        self._advance(1)  # raises with the step named
                self._advance(chunk)
                self.steps_completed += chunk
                done += chunk
                if self.steps_completed % self._n_check == 0:
                    self._check_stability()
                if verbose and done % report == 0:
                    if is_cuda:
                        torch.cuda.synchronize()
                    elapsed = time.perf_counter() - t_start
                    if elapsed > 0.0:
                        self.mcells_per_second = (self.steps_completed - start) * g.Nx * g.Ny * g.Nz / elapsed / 1e6
                    self.elapsed_time = elapsed
                    print(f"  {100 * done // n_steps:3d}%  step {self.steps_completed}  "
                          f"field_max={self.last_field_max:.3e}  {self.mcells_per_second:.1f} Mcells/s")
        finally:
            if is_cuda:
                torch.cuda.synchronize()
            total = time.perf_counter() - t_start
            self.elapsed_time += total
            if total > 0.0:
                self.mcells_per_second = (self.steps_completed - start) * g.Nx * g.Ny * g.Nz / total / 1e6

    def reset(self) -> None:
        self._fields.zero_()
        self.steps_completed = 0
        self.elapsed_time = 0.0
        self.mcells_per_second = 0.0
        self.last_field_max = -1.0
        self.n_stability_checks = 0

    grid = property(lambda self: self._grid)
    fields = property(lambda self: self._fields)
    dt = property(lambda self: self._grid.dt)
    time = property(lambda self: self.steps_completed * self._grid.dt)
    throughput = property(lambda self: self.mcells_per_second)

    def __repr__(self) -> str:
        g = self._grid
        return (f"FDTD3D(shape=({g.Nx}, {g.Ny}, {g.Nz}), steps={self.steps_completed}, time={self.time:.4e} s, "
                f"throughput={self.mcells_per_second:.2f} Mcells/s, device={g.device})")


_C0_REF = 299_792_458.0  # the Mur coefficients' speed of light (src/core/boundaries.py:23)
