"""2D TM FDTD stepper on HIP kernels (drop-in for the reference ``core.fdtd2d``).

Same constructor, attributes, telemetry and errors as ``src/core/fdtd2d.py``
(``FDTD2D`` :69-423, ``SimulationDivergedError`` :48-62). What changes is
where the work runs:

* ``step()`` runs one reference step as three HIP launches (snapshot+sources,
  Faraday+Mur, Ampere+detectors) instead of ~38 eager tensor ops;
* ``run(n)`` advances in blocks of ``depth`` steps per launch with the
  temporally blocked LDS kernel, split at the ``n_check`` boundaries so the
  stability check sees the same fields at the same steps as the reference.

Division. The reference divides float32 field differences by the Python
floats dx, dy (``src/core/fdtd2d.py:232-235, 250-251``); torch forms that
quotient as ``x * fp32(1/dx)`` on a GPU and as IEEE ``x / fp32(dx)`` on the
CPU. ``division="reciprocal"`` (the default: this stepper runs on the GPU)
reproduces the reference as it runs on a GPU, ``division="ieee"`` as it runs
on the CPU; both bit for bit.

Sources are read from ``self._sources`` at every call (tests reassign it), and
the field tensors of the ``FieldSet`` are the live state, read and written in
place, so callers may inspect or modify them between calls. Optional
``detectors`` (a ``DetectorSet``) record Hz inside the kernels after every
step, replacing the callers' ``buf[:, step] = Hz[rx_i, rx_j, 0]`` loop.
"""

from __future__ import annotations

import ctypes
import time
from typing import Optional

import numpy as np
import torch

from .. import _native
from .._native import WfF64Args, WfRun2dF64
from ..engine import (
    RunState,
    check_division,
    iter_launches,
    judge,
    launch_parity,
    schedule_segments,
    SourceTable,
    StepCoefficients,
    TBPlan,
    maxabs,
    maxabs64,
    per_step_advance,
    require_gpu_tensor,
    split_launches,
    tb_limits_ok,
)
from .boundaries import MurABC
from .detectors import DetectorSet
from .fields import FieldSet
from .grid import YeeGrid

EPS0: float = 8.8541878128e-12
MU0: float = 1.2566370614e-6

DEFAULT_TILE_ROWS = 48
DEFAULT_DEPTH = 8
DEFAULT_CPL = 1
SMALL_SET_TILE_ROWS = 16
SMALL_SET_MAX_TILES = 64
TINY_SET_TILE_ROWS = 8
TINY_SET_MAX_TILES = 288
MID_SET_TILE_ROWS = 32
MID_SET_MAX_TILES = 1000
LARGE_SET_TILE_ROWS = 64
HUGE_RUN_CELLS = 1 << 23
FALLBACK_SHAPES = ((32, 6, 1), (16, 6, 1))


def tile_shape(tile) -> tuple[int, int]:
    """``(tile rows, columns per lane)`` of a tile argument: an int is a tile
    height with one column per lane, a pair names both."""
    if isinstance(tile, (tuple, list)) and len(tile) == 2 and all(isinstance(v, int) for v in tile):
        return int(tile[0]), int(tile[1])
    if isinstance(tile, int) and not isinstance(tile, bool):
        return int(tile), 1
    raise ValueError(f"tile must be an int (rows), a (rows, columns per lane) pair, None or 'auto'; got {tile!r}")


def default_launch_shape(runs) -> tuple[int, int, int]:
    """(tile rows, depth, columns per lane) when none is given: always 8
    steps deep, the tile height from the number n48 of 48-row tiles (48
    columns) the run set makes, following the autotuner's tables on the
    MI355X (single grids of 128^2 to 4096^2, 16 to 64 runs of 150^2 to
    384^2):

    * at most 288 tiles of 8 rows (one grid up to about 300^2): 8-row tiles,
      which spread a small grid over the most CUs; else, n48 < 64: 16 rows;
    * n48 < 1000 (one grid of 512^2 to about 1400^2, 16 runs of 256^2): 32
      rows, 10-35 % ahead of 48;
    * larger sets: 64 rows (fewer partial and edge tiles per grid: 7-40 %
      ahead of 48 on 2048^2, 16 x 384^2, 64 x 256^2), except when a run holds
      2^23 cells or more (4096^2), where 48 and 64 rows tie at depth 8 and 48
      rows are ahead at depth 7.

    If a run has more sources or detectors in one tile than the kernel
    stages, smaller tiles are tried before the caller falls back to the
    per-step path."""
    tj = 64 * DEFAULT_CPL - 2 * DEFAULT_DEPTH
    n = sum(-(-r.nx // DEFAULT_TILE_ROWS) * -(-r.ny // tj) for r in runs)
    n_tiny = sum(-(-r.nx // TINY_SET_TILE_ROWS) * -(-r.ny // tj) for r in runs)
    if n_tiny <= TINY_SET_MAX_TILES:
        first = (TINY_SET_TILE_ROWS, DEFAULT_DEPTH, 1)
    elif n < SMALL_SET_MAX_TILES:
        first = (SMALL_SET_TILE_ROWS, DEFAULT_DEPTH, 1)
    elif n < MID_SET_MAX_TILES:
        first = (MID_SET_TILE_ROWS, DEFAULT_DEPTH, 1)
    elif max(r.nx * r.ny for r in runs) >= HUGE_RUN_CELLS:
        first = (DEFAULT_TILE_ROWS, DEFAULT_DEPTH, DEFAULT_CPL)
    else:
        first = (LARGE_SET_TILE_ROWS, DEFAULT_DEPTH, 1)
    for tile, depth, cpl in (first,) + FALLBACK_SHAPES:
        if all(tb_limits_ok(r, tile, d, cpl) for r in runs for d in range(1, depth + 1)):
            return tile, depth, cpl
    return first


class SimulationDivergedError(RuntimeError):
    """Raised when a field component exceeds the stability threshold."""

    def __init__(self, message: str) -> None:
        super().__init__(message)


class FDTD2D:
    """2D TM Yee leapfrog stepper ({Ex, Ey, Hz}) running on gfx950 HIP kernels."""

    def __init__(
        self,
        grid: YeeGrid,
        fields: FieldSet,
        boundary: MurABC,
        sources=None,
        *,
        Ca: Optional[torch.Tensor] = None,
        Cb: Optional[torch.Tensor] = None,
        stability_threshold: float = 1e10,
        n_check: int = 100,
        detectors: Optional[DetectorSet] = None,
        depth=None,
        tile=None,
        fast_div: bool = True,
        division: str = "reciprocal",
    ) -> None:
        if grid.Nx < 4:
            raise ValueError(f"grid.Nx must be >= 4 for a valid Yee stencil, got {grid.Nx}")
        if grid.Ny < 4:
            raise ValueError(f"grid.Ny must be >= 4 for a valid Yee stencil, got {grid.Ny}")
        if boundary._Nx != grid.Nx or boundary._Ny != grid.Ny:
            raise ValueError(
                f"boundary was built for a grid of shape ({boundary._Nx}, {boundary._Ny}) but grid has shape "
                f"({grid.Nx}, {grid.Ny}). They must match."
            )
        if (Ca is None) != (Cb is None):
            raise ValueError("Ca and Cb must be provided together or both omitted.")
        if detectors is not None and getattr(detectors, "component", "Hz") != "Hz":
            raise ValueError("the 2D TM stepper records Hz detectors only")
        self._grid = grid
        self._fields = fields
        self._boundary = boundary
        self._sources = sources
        self._Dh: float = grid.dt / MU0
        self._De: float = grid.dt / EPS0
        self._dx: float = grid.dx
        self._dy: float = grid.dy
        self._Ca = Ca
        self._Cb = Cb
        self._has_materials = Ca is not None
        self._threshold = stability_threshold
        self._n_check = n_check
        self._detectors = detectors
        if depth is not None and depth != "auto" and not isinstance(depth, int):
            raise ValueError(f"depth must be an int, None or 'auto', got {depth!r}")
        if isinstance(depth, int) and depth not in range(1, _native.TB_MAX_DEPTH + 1):
            raise ValueError(f"depth must be in 1..{_native.TB_MAX_DEPTH}, got {depth}")
        if tile is not None and tile != "auto":
            tile = tile_shape(tile)
        self._depth = depth
        self._tile = tile
        self.launch_shape = None
        self._fast_div = fast_div
        self._division = check_division(division)
        self.steps_completed: int = 0
        self.elapsed_time: float = 0.0
        self.mcells_per_second: float = 0.0
        self.last_field_max: float = -1.0
        self.n_stability_checks: int = 0
        self._cache_key = None
        self._runs: list[RunState] = []
        self._src_table: Optional[SourceTable] = None
        self._src_key = None
        self._second = None
        self._third = None
        self._tb_plan: Optional[TBPlan] = None
        self._tb_key = None
        self._shape_key = None
        self._status: Optional[torch.Tensor] = None   # device word the kernels raise WF_STATUS_* bits in
        self._slots: Optional[torch.Tensor] = None    # check slots (+ a copy of the status word), reused
        self._slots_dirty = 0                          # leading slot words the last call left nonzero
        self._shadows: dict = {}     # contiguous stand-ins for strided field tensors
        self._shadowed: list = []    # (FieldSet tensor, shadow) pairs of the current call
        self._live = None            # (Ex, Ey, Hz) the kernels step
        self._checked_key = None
        self._f64_key = None
        self._stop: Optional[torch.Tensor] = None     # per-run stop flags, all zero between calls

    # ------------------------------------------------------------------
    # device state
    # ------------------------------------------------------------------
    def _device_check(self) -> None:
        """Fields on the GPU, float32 (the fp32 kernels) or all float64 (the
        fp64 kernels, as the reference runs a float64 grid); Ca/Cb float32, or
        float32/float64 with float64 fields. Checked when the tensors change."""
        f = self._fields
        key = (id(f.Ex), id(f.Ey), id(f.Hz), id(self._Ca), id(self._Cb))
        if key == self._checked_key:
            return
        f64 = f.Hz.dtype == torch.float64
        kinds = (torch.float64,) if f64 else (torch.float32,)
        for name in ("Ex", "Ey", "Hz"):
            require_gpu_tensor(getattr(f, name), f"fields.{name}", kinds)
        if self._Ca is not None:
            ckinds = (torch.float32, torch.float64) if f64 else (torch.float32,)
            require_gpu_tensor(self._Ca, "Ca", ckinds)
            require_gpu_tensor(self._Cb, "Cb", ckinds)
        _native.load()
        self._checked_key = key

    def _is_f64(self) -> bool:
        return self._fields.Hz.dtype == torch.float64

    def _live_fields(self) -> tuple:
        """(Ex, Ey, Hz) the kernels step: the FieldSet tensors, or, for one not
        laid out as the kernels read it (a strided view, as
        ``FieldSet.from_tensor`` can give, ``src/core/fields.py:427-430``), a
        contiguous stand-in that every call copies in before it steps and
        back after, so the FieldSet tensors stay the live state."""
        f = self._fields
        out, shadowed = [], []
        for name in ("Ex", "Ey", "Hz"):
            t = getattr(f, name)
            if t.is_contiguous() or (t.dim() == 4 and all(t[b].is_contiguous() for b in range(t.shape[0]))):
                out.append(t)
                continue
            sh = self._shadows.get(name)
            if sh is None or sh.shape != t.shape or sh.dtype != t.dtype or sh.device != t.device:
                sh = torch.empty(t.shape, dtype=t.dtype, device=t.device)
                self._shadows[name] = sh
            shadowed.append((t, sh))
            out.append(sh)
        self._shadowed = shadowed
        self._live = tuple(out)
        return self._live

    def _copy_in(self) -> None:
        for t, sh in self._shadowed:
            sh.copy_(t)

    def _copy_out(self) -> None:
        for t, sh in self._shadowed:
            t.copy_(sh)

    def _stream(self) -> int:
        return _native.stream_handle(self._fields.Hz.device)

    def _source_list(self):
        if self._sources is None:
            return []
        return list(getattr(self._sources, "_sources", self._sources))

    def _source_table(self) -> SourceTable:
        srcs = self._source_list()
        key = (id(self._sources), tuple(id(s) for s in srcs))
        if key != self._src_key:
            self._src_table = SourceTable(srcs, self._grid.Nx, self._grid.Ny, self._fields.Hz.device)
            self._src_key = key
            self._tb_key = None
        return self._src_table

    def _members(self, t: torch.Tensor) -> list[torch.Tensor]:
        """Per-batch-member (Nx, Ny) views of a field or coefficient tensor."""
        Nx, Ny = self._grid.Nx, self._grid.Ny
        if t.dim() == 4:
            return [t[b].reshape(Nx, Ny) for b in range(t.shape[0])]
        return [t.reshape(Nx, Ny)]

    def _build_runs(self) -> list[RunState]:
        tensors = self._live_fields()
        srcs = self._source_table()
        det = self._detectors
        key = (
            tuple(t.data_ptr() for t in tensors),
            None if self._Ca is None else (self._Ca.data_ptr(), self._Cb.data_ptr()),
            self._src_key,
            id(det),
        )
        if key == self._cache_key:
            return self._runs
        exs, eys, hzs = (self._members(t) for t in tensors)
        B = len(hzs)
        cas = cbs = [None] * B
        if self._Ca is not None:
            ca_m = self._members(self._Ca.contiguous())
            cb_m = self._members(self._Cb.contiguous())
            cas = ca_m if len(ca_m) == B else ca_m * B
            cbs = cb_m if len(cb_m) == B else cb_m * B
        coef = StepCoefficients.from_grid(self._grid.dt, self._grid.dx, self._grid.dy,
                                          self._boundary.C_mur_x, self._boundary.C_mur_y)
        Nx, Ny = self._grid.Nx, self._grid.Ny
        if self._status is None or self._status.device != hzs[0].device:
            self._status = torch.zeros(1, dtype=torch.int32, device=hzs[0].device)
        runs = []
        for b in range(B):
            rs = RunState(exs[b], eys[b], hzs[b], cas[b], cbs[b], Nx, Ny, coef, sources=srcs, status=self._status,
                          reciprocal=self._division == "reciprocal")
            rs.pstrip = torch.empty(4 * (Nx + Ny), dtype=torch.float32, device=hzs[b].device)
            if det is not None and len(det):
                rs.det_cell = det.cell_index
                rs.traces = det.traces if det.traces.dim() == 2 else det.traces[b]
                rs.trace_n0 = det.start_step
                host = getattr(det, "_cell_index_host", None)
                if host is not None and host.size == det.cell_index.numel():
                    rs.extra["det_cells_np"] = host  # the tile lists need no read-back
            runs.append(rs)
        self._runs = runs
        self._cache_key = key
        self._tb_key = None
        return runs

    # ------------------------------------------------------------------
    # stepping
    # ------------------------------------------------------------------
    def _apply_inert(self, n0: int, n_steps: int) -> None:
        """Sources on components the 2D TM update never reads (Ez, Hx, Hy)
        still add into their tensors, as the reference does."""
        inert = self._src_table.inert if self._src_table is not None else []
        if not inert:
            return
        d = self._fields.as_dict()
        for n in range(n0, n0 + n_steps):
            for src in inert:
                src.step(d, n)

    def _wave_limit(self) -> Optional[int]:
        t = self._f64_table()[2] if self._is_f64() else self._source_table()
        return t.min_wave_len

    def _advance(self, n_steps: int) -> None:
        """Advance ``n_steps`` steps starting at ``steps_completed`` (no checks)."""
        if n_steps <= 0:
            return
        if self._is_f64():
            self._f64_advance(n_steps)
            self._apply_inert(self.steps_completed, n_steps)
            return
        setup = self._tb_setup()
        runs = self._build_runs()
        stream = self._stream()
        if n_steps >= 2 and setup is not None:
            _, tile, depth = setup
            first = self.steps_completed
            if split_launches(n_steps, depth, odd_ok=True) is None:
                # no split into two or more launches: one in-place step first
                for r in runs:
                    r.step = first
                per_step_advance(runs, 1, stream)
                first += 1
                n_steps -= 1
            if n_steps > 0:
                plan = self._tb_plan_for(runs, tile, depth, set(split_launches(n_steps, depth, odd_ok=True)))
                plan.advance(n_steps, depth, stream, base_step=first)
                plan.prepare(depth)  # the other depths, while the GPU runs these launches
        else:
            for r in runs:
                r.step = self.steps_completed
            per_step_advance(runs, n_steps, stream)
            self._apply_inert(self.steps_completed, n_steps)

    def _tb_plan_for(self, runs, tile: tuple[int, int], depth: int, depths=None) -> TBPlan:
        """The blocked-kernel plan for these runs, with its tile tables and
        launch arguments built for ``depths`` (the depths the caller will
        launch; every depth up to ``depth`` if None), in one upload. It uses a
        second and a third state buffer of the fields' shape, allocated
        together on first use and kept, so launch counts of either parity end
        in the live tensors."""
        shape = self._fields.Hz.shape
        dev = self._fields.Hz.device
        if self._second is None or self._second[0].shape != shape:
            self._second = tuple(torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(3))
            self._third = tuple(torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(3))
            self._tb_key = None
        key = (self._cache_key, tile, self._fast_div, depth)
        if key != self._tb_key:
            ex2, ey2, hz2 = (self._members(t) for t in self._second)
            ex3 = ey3 = hz3 = None
            if self._third is not None:
                ex3, ey3, hz3 = (self._members(t) for t in self._third)
            for b, r in enumerate(runs):
                r.second = (ex2[b], ey2[b], hz2[b])
                r.third = None if ex3 is None else (ex3[b], ey3[b], hz3[b])
                r.step = 0  # launches pass the global step as step_offset
            self._tb_plan = TBPlan(runs, tile[0], dev, fast_div=self._fast_div, cpl=tile[1])
            self._tb_key = key
        self._tb_plan.prepare(depth, depths)
        return self._tb_plan

    def step(self) -> None:
        """Advance all fields by one time step in place (``src/core/fdtd2d.py:187-266``)."""
        n = self.steps_completed
        limit = self._wave_limit()
        if limit is not None and n >= limit:
            raise RuntimeError(
                f"source waveform table has {limit} samples; cannot inject at step {n} "
                "(build sources with a larger N_steps)"
            )
        self._device_check()
        if self._is_f64():
            self._live_fields()
            self._copy_in()
            try:
                self._f64_advance(1)
            finally:
                self._copy_out()
        else:
            runs = self._build_runs()
            self._copy_in()
            for r in runs:
                r.step = n
            try:
                per_step_advance(runs, 1, self._stream())
            finally:
                self._copy_out()
        self._apply_inert(n, 1)
        self.steps_completed += 1
        if self.steps_completed % self._n_check == 0:
            self._check_stability()

    def _check_stability(self) -> None:
        """Max-abs check of Ex, Ey, Hz (``src/core/fdtd2d.py:268-295``)."""
        ex, ey, hz = self._live if self._live is not None else self._live_fields()
        reduce = maxabs64 if ex.dtype == torch.float64 else maxabs
        self._judge_values(reduce(ex, ey, hz, self._stream()).cpu().tolist())

    # ------------------------------------------------------------------
    # float64 grids: the fp64 kernels, per step
    # ------------------------------------------------------------------
    def _f64_table(self):
        """Device table of fp64 run descriptors over the live float64 fields
        (one per batch member), rebuilt when tensors or sources change. The
        update and Mur coefficients are the reference's Python doubles
        (``src/core/fdtd2d.py:158-159``, ``boundaries.py:177-178``); Ca/Cb are
        widened to float64 exactly, as torch promotes them."""
        from ..fp64 import SourceTable64
        from ..engine import EPS0 as _E0, MU0 as _M0, check_grid_cells

        ex, ey, hz = self._live_fields()
        srcs = self._source_list()
        det = self._detectors
        key = (ex.data_ptr(), ey.data_ptr(), hz.data_ptr(),
               None if self._Ca is None else (self._Ca.data_ptr(), self._Cb.data_ptr()),
               id(self._sources), tuple(id(q) for q in srcs), id(det))
        if key == self._f64_key:
            return self._f64_state
        Nx, Ny = self._grid.Nx, self._grid.Ny
        check_grid_cells(Nx, Ny)
        dev = hz.device
        exs, eys, hzs = (self._members(t) for t in (ex, ey, hz))
        B = len(hzs)
        keep = []
        cas = cbs = [None] * B
        if self._Ca is not None:
            ca_m = self._members(self._Ca.to(torch.float64).contiguous())
            cb_m = self._members(self._Cb.to(torch.float64).contiguous())
            keep += ca_m + cb_m
            cas = ca_m if len(ca_m) == B else ca_m * B
            cbs = cb_m if len(cb_m) == B else cb_m * B
        table = SourceTable64(srcs, Nx, Ny, dev)
        if self._status is None or self._status.device != dev:
            self._status = torch.zeros(1, dtype=torch.int32, device=dev)
        structs = []
        g = self._grid
        for b in range(B):
            r = WfRun2dF64()
            r.ex, r.ey, r.hz = exs[b].data_ptr(), eys[b].data_ptr(), hzs[b].data_ptr()
            if cas[b] is not None:
                r.ca, r.cb = cas[b].data_ptr(), cbs[b].data_ptr()
            strip = torch.empty(4 * (Nx + Ny), dtype=torch.float64, device=dev)
            keep.append(strip)
            r.pstrip = strip.data_ptr()
            if table.n_grp:
                r.inj, r.grp_start, r.n_grp, r.n_inj = (table.inj.data_ptr(), table.grp_start.data_ptr(),
                                                         table.n_grp, table.n_inj)
            if det is not None and len(det):
                tr = det.traces if det.traces.dim() == 2 else det.traces[b]
                if tr.dtype != torch.float64:
                    raise ValueError("detectors of a float64 simulation record float64 traces "
                                     "(build the DetectorSet with the float64 grid)")
                r.det_cell, r.traces = det.cell_index.data_ptr(), tr.data_ptr()
                r.trace_stride, r.n_det, r.trace_n0, r.trace_len = tr.stride(0), len(det), det.start_step, tr.shape[-1]
            r.nx, r.ny, r.step = Nx, Ny, 0
            r.dh, r.de, r.dx, r.dy = g.dt / _M0, g.dt / _E0, g.dx, g.dy
            r.cmx, r.cmy = self._boundary.C_mur_x, self._boundary.C_mur_y
            r.status = self._status.data_ptr()
            structs.append(r)
        arr = (WfRun2dF64 * B)(*structs)
        dev_table = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(dev)
        self._f64_state = (dev_table, B, table, keep)
        self._f64_key = key
        self._src_table = table   # its `inert` and `min_wave_len` serve the shared host logic
        self._src_key = (id(self._sources), tuple(id(q) for q in srcs))
        return self._f64_state

    def _f64_advance(self, n_steps: int) -> None:
        """``n_steps`` fp64 steps in place from ``steps_completed``."""
        dev_table, B, _, _ = self._f64_table()
        a = WfF64Args()
        a.runs, a.n_runs = dev_table.data_ptr(), B
        a.max_nx, a.max_ny = self._grid.Nx, self._grid.Ny
        a.n_steps, a.step_offset = int(n_steps), self.steps_completed
        lib = _native.load()
        _native.check(lib.wf_fdtd2d_step_f64(ctypes.byref(a), ctypes.c_void_p(self._stream())), "wf_fdtd2d_step_f64")

    def _tb_setup(self):
        """(runs, tile, depth) if the temporally blocked kernel can run this
        simulation (no inert-component sources, per-tile limits met)."""
        self._device_check()
        runs = self._build_runs()
        inert = self._src_table.inert if self._src_table is not None else []
        key = (self._cache_key, self._tile, self._depth, self._fast_div)
        if self._shape_key != key:
            tile, depth = self._launch_shape(runs, inert)
            ok = not inert and all(tb_limits_ok(r, tile[0], d, tile[1]) for r in runs for d in range(1, depth + 1))
            self._shape_cached = (tile, depth, ok)
            self._shape_key = key
        tile, depth, ok = self._shape_cached
        self.launch_shape = (tile[0], depth, tile[1])
        return (runs, tile, depth) if ok else None

    def _launch_shape(self, runs, inert) -> tuple[tuple[int, int], int]:
        """((tile rows, columns per lane), depth): as given, the defaults, or,
        for ``"auto"``, the autotuner's choice for these grids (timed once per
        shape set)."""
        tile0, depth0, cpl0 = default_launch_shape(runs)
        depth = self._depth if isinstance(self._depth, int) else depth0
        tile = self._tile if isinstance(self._tile, tuple) else (tile0, cpl0)
        if "auto" in (self._tile, self._depth) and not inert:
            if getattr(self, "_tuned_key", None) != self._cache_key:
                from .. import autotune

                res = autotune.tune_runs(
                    runs, fast_div=self._fast_div,
                    tile_rows=None if self._tile == "auto" else [tile],
                    depths=autotune.DEPTHS if self._depth == "auto" else [depth])
                self._tuned, self._tuned_key = res, self._cache_key
            tile, depth = (self._tuned.tile_rows, self._tuned.cpl), self._tuned.depth
        return tile, depth

    def _run_async(self, n_steps: int, setup) -> None:
        """``n_steps`` steps with the stability checks judged on the device.

        The whole run is enqueued without waiting for the host: at each check
        step the launch ending there folds max|Ex|, |Ey|, |Hz| into a slot, a
        one-thread kernel compares it with the threshold and raises a device
        flag, and every later launch carries the flagged state through
        unchanged. One synchronisation at the end reads the slots (and the
        kernels' status word); a failed check raises
        ``SimulationDivergedError`` with the fields and ``steps_completed``
        of that check step, as the reference does. Slots, stop flags and the
        launch arguments persist across calls, so a call costs the launches
        and little else."""
        runs, tile, depth = setup
        start = self.steps_completed
        limit = self._wave_limit()
        n_exec = n_steps if limit is None else max(0, min(n_steps, limit - start))
        segs, checks = schedule_segments(start, n_exec, depth, self._n_check, odd_ok=True)
        if n_exec and segs is None:
            return self._run_sync(n_steps)
        if n_exec:
            stream = self._stream()
            dev = self._fields.Hz.device
            # the depths this call launches: each segment's launches are as equal as possible
            used = {q for seg, n_l, _ in segs for q in (seg // n_l, -(-seg // n_l))}
            plan = self._tb_plan_for(runs, tile, depth, used)
            n_runs = len(runs)
            if self._stop is None or self._stop.numel() != n_runs or self._stop.device != dev:
                self._stop = torch.zeros(n_runs, dtype=torch.int32, device=dev)
            need = max(1, len(checks)) * n_runs * 3 + 1
            if self._slots is None or self._slots.numel() < need or self._slots.device != dev:
                self._slots = torch.zeros(need, dtype=torch.int32, device=dev)
                self._slots_dirty = 0
            if checks and self._slots_dirty:
                self._slots[:self._slots_dirty].zero_()  # what the previous call folded into
                self._slots_dirty = 0
            n_launch = sum(n for _, n, _ in segs)
            offset = start
            for k, (d, slot) in enumerate(iter_launches(segs)):
                view = self._slots[slot * n_runs * 3:(slot + 1) * n_runs * 3] if slot >= 0 else None
                plan.launch(d, launch_parity(k, n_launch), offset, stream, maxabs_out=view, stop=self._stop)
                if slot >= 0:
                    judge(view, self._threshold, self._stop, True, stream)
                offset += d
            # the other depths' tables, built while the GPU runs the launches above,
            # so a later call with another step count finds them ready
            plan.prepare(depth)
            if checks:
                used = len(checks) * n_runs * 3
                self._slots[used:used + 1].copy_(self._status)
                host = self._slots[:used + 1].cpu().numpy()
                self._slots_dirty = used + 1
                status = int(host[used])
                vals = host[:used].view(np.float32).reshape(len(checks), n_runs, 3).max(axis=1).tolist()
                if status:
                    self._status.zero_()
                    raise RuntimeError(f"waveforge_amd: the stepping kernels reported status 0x{status:x} "
                                       "(a source past its waveform table, or a tile over its source or "
                                       "detector limit)")
                try:
                    for step_m, (vx, vy, vh) in zip(checks, vals):
                        self.steps_completed = step_m
                        self._judge_values((vx, vy, vh))
                except SimulationDivergedError:
                    self._stop.zero_()
                    raise
        self.steps_completed = start + n_exec
        if n_exec < n_steps:
            raise RuntimeError(
                f"source waveform table has {limit} samples; cannot inject at step {self.steps_completed} "
                "(build sources with a larger N_steps)"
            )

    def _judge_values(self, vals) -> None:
        self.n_stability_checks += 1
        for name, val in zip(("Ex", "Ey", "Hz"), vals):
            if val > self.last_field_max:
                self.last_field_max = val
            if val > self._threshold:
                raise SimulationDivergedError(
                    f"Simulation diverged at step {self.steps_completed}: {name} max = {val:.3e} exceeds "
                    f"threshold {self._threshold:.3e}. Check CFL condition (dt={self._grid.dt:.3e}) and "
                    f"source amplitude."
                )

    def run(self, n_steps: int, *, verbose: bool = False) -> None:
        """Run ``n_steps`` steps (``src/core/fdtd2d.py:301-359``), same telemetry."""
        is_cuda = str(self._grid.device).startswith("cuda")
        if is_cuda:
            torch.cuda.synchronize()
        t_start = time.perf_counter()
        start = self.steps_completed
        self._device_check()
        f64 = self._is_f64()
        setup = self._tb_setup() if (n_steps >= 2 and not verbose and not f64) else None
        self._live_fields()
        self._copy_in()
        try:
            if setup is not None:
                self._run_async(n_steps, setup)
            else:
                self._run_sync(n_steps, verbose=verbose, t_start=t_start)
        finally:
            self._copy_out()
            if is_cuda:
                torch.cuda.synchronize()
            total = time.perf_counter() - t_start
            self.elapsed_time += total
            if total > 0.0:
                self.mcells_per_second = (
                    (self.steps_completed - start) * self._grid.Nx * self._grid.Ny / total / 1e6
                )

    def _run_sync(self, n_steps: int, verbose: bool = False, t_start: float | None = None) -> None:
        """Chunked run with a host-side check at every ``n_check`` step (used
        for verbose runs and whenever the temporally blocked kernel cannot run)."""
        is_cuda = str(self._grid.device).startswith("cuda")
        report = max(1, n_steps // 10)
        t_start = time.perf_counter() if t_start is None else t_start
        start = self.steps_completed
        done = 0
        while done < n_steps:
            chunk = n_steps - done
            chunk = min(chunk, self._n_check - (self.steps_completed % self._n_check))
            if verbose:
                chunk = min(chunk, report - (done % report))
            limit = self._wave_limit()
            if limit is not None and self.steps_completed + chunk > limit:
                chunk = limit - self.steps_completed
                if chunk <= 0:
                    raise RuntimeError(
                        f"source waveform table has {limit} samples; cannot inject at step "
                        f"{self.steps_completed} (build sources with a larger N_steps)"
                    )
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
                    self.mcells_per_second = (
                        (self.steps_completed - start) * self._grid.Nx * self._grid.Ny / elapsed / 1e6
                    )
                self.elapsed_time = elapsed
                print(
                    f"  {100 * done // n_steps:3d}%  step {self.steps_completed}  "
                    f"field_max={self.last_field_max:.3e}  {self.mcells_per_second:.1f} Mcells/s"
                )

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
        return (
            f"FDTD2D(shape=({g.Nx}, {g.Ny}, {g.Nz}), steps={self.steps_completed}, time={self.time:.4e} s, "
            f"throughput={self.mcells_per_second:.2f} Mcells/s, device={g.device})"
        )
