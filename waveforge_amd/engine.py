"""Host driver of the HIP stepping kernels.

Builds the device tables the C ABI reads (run descriptors, flattened source
contributions, detector cells, tile lists) from the reference-style objects
and launches either

* the per-step path (``wf_fdtd2d_step``): three launches per step, in place,
  used for single steps and as the general path; or
* the temporally blocked path (``wf_fdtd2d_tb_advance``): ``depth`` steps per
  launch over LDS tiles, ping-ponging between the live tensors and a second
  state buffer, with an even number of launches so the result always lands in
  the live tensors.

Both paths compute the same fp32 values (the kernels share the reference's
operation order), so the choice is a performance decision only.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native
from ._native import WfInj2d, WfRun2d, WfTbArgs

EPS0 = 8.8541878128e-12
MU0 = 1.2566370614e-6
C0 = 299_792_458.0

INJ_DTYPE = np.dtype(
    {
        "names": ["wave", "wave_len", "comp", "cell", "amp"],
        "formats": ["<u8", "<i4", "<i4", "<i4", "<f4"],
        "offsets": [0, 8, 12, 16, 20],
        "itemsize": 24,
    }
)
assert ctypes.sizeof(WfInj2d) == INJ_DTYPE.itemsize

TB_MAX_GROUPS_PER_TILE = 64
TB_MAX_DETS_PER_TILE = 64
TB_MAX_GROUPS = 4096


def to_device_async(a, device, dtype=None) -> torch.Tensor:
    """A host array (NumPy or a list) as a device tensor, without waiting for
    the device: the values are staged in a newly allocated pinned buffer and
    copied asynchronously on the current stream (torch's pinned allocator
    keeps the buffer until the copy has run), so a host that builds tables
    while the GPU works is not held back by a synchronous copy."""
    t = torch.as_tensor(np.asarray(a)) if dtype is None else torch.as_tensor(np.asarray(a), dtype=dtype)
    device = torch.device(device)
    if device.type != "cuda":
        return t.to(device)
    pinned = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    pinned.copy_(t)
    return pinned.to(device, non_blocking=True)


_TORCH_OF = {np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64, np.dtype(np.float32): torch.float32,
             np.dtype(np.float64): torch.float64, np.dtype(np.uint8): torch.uint8}


def upload_packed(arrays: Sequence[np.ndarray], device) -> list[torch.Tensor]:
    """Several host arrays in one asynchronous copy (``to_device_async`` of
    their bytes, each at a 16-byte boundary) and a device view of each, of its
    shape and dtype (structured records come back as bytes). One copy instead
    of one per array: each copy costs a queue round trip before the kernels
    that need it can start."""
    parts, spans, off = [], [], 0
    for x in arrays:
        x = np.ascontiguousarray(x)
        b = x.view(np.uint8).reshape(-1)
        spans.append((off, b.size, x))
        parts.append(b)
        pad = (-b.size) % 16
        if pad:
            parts.append(np.zeros(pad, dtype=np.uint8))
        off += b.size + pad
    buf = to_device_async(np.concatenate(parts) if parts else np.zeros(16, np.uint8), device)
    out = []
    for o, n, x in spans:
        v = buf[o:o + n]
        td = _TORCH_OF.get(x.dtype)
        out.append(v if td is None or td == torch.uint8 else v.view(td).view(x.shape))
    return out


def f32(x: float) -> float:
    """Round a Python float to fp32, as torch does for a scalar operand."""
    return float(np.float32(x))


def require_gpu_tensor(t: torch.Tensor, what: str, dtypes=(torch.float32,)) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError(
            f"waveforge_amd: {what} must live on the ROCm GPU (device 'cuda'); got "
            f"{getattr(t, 'device', type(t))}. The HIP kernels are the only compute path."
        )
    if t.dtype not in dtypes:
        names = " or ".join(str(d).replace("torch.", "") for d in dtypes)
        raise RuntimeError(f"waveforge_amd: {what} must be {names}, got {t.dtype}")


DIVISIONS = ("reciprocal", "ieee")


def check_division(division: str) -> str:
    """How a run forms the reference's quotients ``x / dx``, ``x / dy`` of a
    float32 tensor by a Python float (``src/core/fdtd2d.py:232-235, 250-251``):
    ``"reciprocal"``, ``x * fp32(1/dx)`` (the reciprocal of the double rounded
    once, then one fp32 product), which is what torch computes for that
    expression on a GPU; ``"ieee"``, IEEE ``x / fp32(dx)``, which is what torch
    computes on the CPU."""
    if division not in DIVISIONS:
        raise ValueError(f"division must be one of {DIVISIONS}, got {division!r}")
    return division


@dataclass
class StepCoefficients:
    """fp32 scalars of one grid, rounded exactly as the reference's torch ops round them."""

    dh: float
    de: float
    dx: float
    dy: float
    cmx: float
    cmy: float
    rdx: float = 0.0   # fp32(1/dx) of the double dx: the multiplier of the "reciprocal" division
    rdy: float = 0.0

    @classmethod
    def from_grid(cls, dt: float, dx: float, dy: float, cmx: float | None = None, cmy: float | None = None):
        if cmx is None:
            cmx = (C0 * dt - dx) / (C0 * dt + dx)
        if cmy is None:
            cmy = (C0 * dt - dy) / (C0 * dt + dy)
        return cls(f32(dt / MU0), f32(dt / EPS0), f32(dx), f32(dy), f32(cmx), f32(cmy),
                   f32(1.0 / dx), f32(1.0 / dy))


class CellSource:
    """Minimal soft source ``field[i, j] += wave[n] * amp`` for drivers that do
    not build reference-style source objects (same attributes the table reads)."""

    def __init__(self, component: str, wave: torch.Tensor, i: int, j: int, amp: float = 1.0) -> None:
        self.component = component
        self._waveform_tensor = wave
        self._i_idx = torch.tensor([int(i)])
        self._j_idx = torch.tensor([int(j)])
        self._amp = torch.tensor([float(amp)], dtype=torch.float32)


class SourceTable:
    """Flattened soft-source contributions of a source list, on the device.

    Contributions are listed in collection order and cell order, then grouped
    by (component, cell) with a stable sort, so a cell hit by several sources
    receives them in the reference's order. Components other than Ex/Ey/Hz
    (inert in 2D TM) are kept aside in ``inert`` for the host to apply.
    """

    def __init__(self, sources: Sequence, Nx: int, Ny: int, device: torch.device, Nz: int = 1,
                 comps_map: Optional[dict] = None, upload: bool = True) -> None:
        """``comps_map`` maps the components the kernels update to their codes
        (2D TM: Ex, Ey, Hz; 3D: all six); cells are ``(i*Ny + j)*Nz + k``."""
        from .core.sources import PlaneSource

        cmap = _native.WF_COMP if comps_map is None else comps_map
        comps, cells, amps, waves, lens = [], [], [], [], []
        self.inert = []
        self.keepalive = []
        self.min_wave_len = None
        for order, src in enumerate(sources):
            comp = src.component
            wave = src._waveform_tensor
            if comp not in cmap:
                self.inert.append(src)
                self.min_wave_len = len(wave) if self.min_wave_len is None else min(self.min_wave_len, len(wave))
                continue
            require_gpu_tensor(wave, "source waveform")
            wave = wave.contiguous()
            self.keepalive.append(wave)
            if isinstance(src, PlaneSource):
                i, j, k, a = src.cells3d(Nx, Ny, Nz)
            else:
                from .core.sources import host_cells

                i, j, k, a = (torch.as_tensor(np.asarray(v)) for v in host_cells(src))
            n = i.numel()
            comps.append(np.full(n, cmap[comp], dtype=np.int32))
            flat = (i.numpy().astype(np.int64) * Ny + j.numpy().astype(np.int64)) * Nz + k.numpy().astype(np.int64)
            cells.append(flat.astype(np.int32))
            amps.append(a.numpy().astype(np.float32))
            waves.append(np.full(n, wave.data_ptr(), dtype=np.uint64))
            lens.append(np.full(n, wave.numel(), dtype=np.int32))
            self.min_wave_len = wave.numel() if self.min_wave_len is None else min(self.min_wave_len, wave.numel())
        if comps:
            comp = np.concatenate(comps)
            cell = np.concatenate(cells)
            key = comp.astype(np.int64) * (int(Nx) * int(Ny) * int(Nz)) + cell
            perm = np.argsort(key, kind="stable")
            rec = np.zeros(perm.size, dtype=INJ_DTYPE)
            rec["wave"] = np.concatenate(waves)[perm]
            rec["wave_len"] = np.concatenate(lens)[perm]
            rec["comp"] = comp[perm]
            rec["cell"] = cell[perm]
            rec["amp"] = np.concatenate(amps)[perm]
            skey = key[perm]
            starts = np.flatnonzero(np.r_[True, skey[1:] != skey[:-1]])
            grp = np.r_[starts, perm.size].astype(np.int32)
            self.n_inj = int(perm.size)
            self.n_grp = int(starts.size)
            self.group_cells = rec["cell"][starts].copy()
            self.host_inj, self.host_grp = rec, grp
            if upload:
                self.inj, self.grp_start = upload_packed([rec.view(np.uint8), grp], device)
            else:
                self.inj = self.grp_start = None
        else:
            self.n_inj = self.n_grp = 0
            self.group_cells = np.zeros(0, dtype=np.int32)
            self.host_inj = self.host_grp = None
            self.inj = None
            self.grp_start = None

    @classmethod
    def cell(cls, component: str, wave: torch.Tensor, i: int, j: int, amp: float, Nx: int, Ny: int) -> "SourceTable":
        """The table of one soft source at one cell of a 2D TM grid, built
        directly (no source object, no sort): the same records as
        ``SourceTable([CellSource(component, wave, i, j, amp)], Nx, Ny, device,
        upload=False)``, for drivers that build hundreds of runs. Upload with
        ``upload_many``."""
        if component not in _native.WF_COMP:
            return cls([CellSource(component, wave, i, j, amp)], Nx, Ny, wave.device, upload=False)
        require_gpu_tensor(wave, "source waveform")
        wave = wave.contiguous()
        t = cls.__new__(cls)
        rec = np.zeros(1, dtype=INJ_DTYPE)
        rec["wave"] = wave.data_ptr()
        rec["wave_len"] = wave.numel()
        rec["comp"] = _native.WF_COMP[component]
        rec["cell"] = int(i) * int(Ny) + int(j)
        rec["amp"] = np.float32(amp)
        t.inert, t.keepalive, t.min_wave_len = [], [wave], wave.numel()
        t.n_inj = t.n_grp = 1
        t.group_cells = rec["cell"].copy()
        t.host_inj, t.host_grp = rec, np.array([0, 1], dtype=np.int32)
        t.inj = t.grp_start = None
        return t

    @staticmethod
    def upload_many(tables: Sequence["SourceTable"], device: torch.device) -> None:
        """Upload the host records of tables built with ``upload=False`` in one
        copy each for contributions and group offsets; every table then points
        into the shared buffers."""
        live = [t for t in tables if t.n_grp]
        if not live:
            return
        inj, grp = upload_packed([np.concatenate([t.host_inj.view(np.uint8) for t in live]),
                                  np.concatenate([t.host_grp for t in live])], device)
        o_i = o_g = 0
        for t in live:
            n_b, n_g = t.host_inj.nbytes, t.host_grp.size
            t.inj, t.grp_start = inj[o_i:o_i + n_b], grp[o_g:o_g + n_g]
            t.keepalive += [inj, grp]
            o_i += n_b
            o_g += n_g

    def ptrs(self) -> tuple[int, int]:
        return (self.inj.data_ptr() if self.inj is not None else 0,
                self.grp_start.data_ptr() if self.grp_start is not None else 0)


MAX_GRID_CELLS = 2 ** 31 - 1


def check_grid_cells(nx: int, ny: int) -> None:
    """The kernels' source and detector tables hold cells as int32 (i*ny + j;
    in 3D, called with nx = Nx*Ny and ny = Nz)."""
    if nx * ny > MAX_GRID_CELLS:
        raise ValueError(f"a grid of {nx * ny} cells is beyond the 2^31-cell index range of the "
                         f"stepping kernels' source and detector tables")


@dataclass
class RunState:
    """One simulation's device buffers and scalars, as read by the kernels."""

    ex: torch.Tensor
    ey: torch.Tensor
    hz: torch.Tensor
    ca: Optional[torch.Tensor]
    cb: Optional[torch.Tensor]
    nx: int
    ny: int
    coef: StepCoefficients
    sources: Optional[SourceTable] = None
    det_cell: Optional[torch.Tensor] = None
    traces: Optional[torch.Tensor] = None
    trace_n0: int = 0
    step: int = 0
    pstrip: Optional[torch.Tensor] = None
    second: Optional[tuple] = None   # (ex2, ey2, hz2)
    third: Optional[tuple] = None    # (ex3, ey3, hz3): lets a blocked run use an odd launch count
    status: Optional[torch.Tensor] = None  # device uint32 word: WF_STATUS_* bits the kernels raise
    step_end: int = 0                      # 0, or the step the blocked kernel stops the run at
    reciprocal: bool = False               # quotients as x * fp32(1/d) (division "reciprocal"), else IEEE
    extra: dict = field(default_factory=dict)

    def __post_init__(self) -> None:
        check_grid_cells(self.nx, self.ny)

    def struct(self) -> WfRun2d:
        r = WfRun2d()
        r.ex, r.ey, r.hz = self.ex.data_ptr(), self.ey.data_ptr(), self.hz.data_ptr()
        if self.second is not None:
            r.ex2, r.ey2, r.hz2 = (t.data_ptr() for t in self.second)
        if self.third is not None:
            r.ex3, r.ey3, r.hz3 = (t.data_ptr() for t in self.third)
        if self.ca is not None:
            r.ca, r.cb = self.ca.data_ptr(), self.cb.data_ptr()
        if self.pstrip is not None:
            r.pstrip = self.pstrip.data_ptr()
        if self.sources is not None and self.sources.n_grp:
            r.inj, r.grp_start = self.sources.ptrs()
            r.n_grp, r.n_inj = self.sources.n_grp, self.sources.n_inj
        if self.det_cell is not None and self.det_cell.numel():
            r.det_cell = self.det_cell.data_ptr()
            r.traces = self.traces.data_ptr()
            r.trace_stride = self.traces.stride(-2)
            r.n_det = self.det_cell.numel()
            r.trace_n0 = self.trace_n0
            r.trace_len = self.traces.shape[-1]
        r.nx, r.ny = self.nx, self.ny
        c = self.coef
        r.dh, r.de, r.dx, r.dy, r.cmx, r.cmy = c.dh, c.de, c.dx, c.dy, c.cmx, c.cmy
        if self.reciprocal:
            if not (c.rdx > 0.0 and c.rdy > 0.0):
                raise ValueError("a run with reciprocal division needs its coefficients' rdx, rdy")
            r.rdx, r.rdy = c.rdx, c.rdy
            r.flags = _native.WF_RUN2D_RECIPROCAL
        r.step = self.step
        if self.status is not None:
            r.status = self.status.data_ptr()
        r.step_end = int(self.step_end)
        return r


def per_step_advance(runs: Sequence[RunState], n_steps: int, stream: int) -> None:
    """``n_steps`` reference-exact steps of every run, in place (per-step path)."""
    if n_steps <= 0:
        return
    lib = _native.load()
    arr = (WfRun2d * len(runs))(*[r.struct() for r in runs])
    _native.check(lib.wf_fdtd2d_step(arr, len(runs), int(n_steps), ctypes.c_void_p(stream)), "wf_fdtd2d_step")


def maxabs(ex: torch.Tensor, ey: torch.Tensor, hz: torch.Tensor, stream: int) -> torch.Tensor:
    """Device tensor ``[max|ex|, max|ey|, max|hz|]`` via the HIP reduction."""
    lib = _native.load()
    n = ex.numel()
    n_partial = int(max(1, min(1024, (n + 255) // 256)))
    partial = torch.empty(3 * n_partial, dtype=torch.float32, device=ex.device)
    out = torch.empty(3, dtype=torch.float32, device=ex.device)
    _native.check(
        lib.wf_fdtd2d_maxabs(ex.data_ptr(), ey.data_ptr(), hz.data_ptr(), n, partial.data_ptr(), n_partial,
                             out.data_ptr(), ctypes.c_void_p(stream)),
        "wf_fdtd2d_maxabs",
    )
    return out


def maxabs64(ex: torch.Tensor, ey: torch.Tensor, hz: torch.Tensor, stream: int) -> torch.Tensor:
    """Device tensor ``[max|ex|, max|ey|, max|hz|]`` of float64 fields (``wf_fdtd2d_maxabs_f64``)."""
    lib = _native.load()
    n = ex.numel()
    n_partial = int(max(1, min(1024, (n + 255) // 256)))
    partial = torch.empty(3 * n_partial, dtype=torch.float64, device=ex.device)
    out = torch.empty(3, dtype=torch.float64, device=ex.device)
    _native.check(
        lib.wf_fdtd2d_maxabs_f64(ex.data_ptr(), ey.data_ptr(), hz.data_ptr(), n, partial.data_ptr(), n_partial,
                                 out.data_ptr(), ctypes.c_void_p(stream)),
        "wf_fdtd2d_maxabs_f64",
    )
    return out


def tile_count(nx: int, ny: int, tile_i: int, depth: int, cpl: int = 1) -> int:
    tj = _native.tile_cols(depth, cpl)
    return ((nx + tile_i - 1) // tile_i) * ((ny + tj - 1) // tj)


def tb_fits_any_tile(run: RunState) -> bool:
    """True if ``run`` meets the blocked kernel's per-tile limits at every tile
    shape and depth: it has no more source groups and detectors in all than
    one tile may hold (and no sources the kernel does not apply)."""
    n_grp = run.sources.n_grp if run.sources is not None else 0
    n_det = run.det_cell.numel() if run.det_cell is not None else 0
    return n_grp <= TB_MAX_GROUPS_PER_TILE and n_det <= TB_MAX_DETS_PER_TILE and \
        not (run.sources is not None and run.sources.inert)


def tb_limits_ok(run: RunState, tile_i: int, depth: int, cpl: int = 1) -> bool:
    """True if every tile region of ``run`` holds at most the per-tile source
    groups and core detectors the temporally blocked kernel can stage."""
    TI, TJ = tile_i, _native.tile_cols(depth, cpl)
    if tb_fits_any_tile(run):
        return True  # no tile can hold more than the whole run has
    nti = (run.nx + TI - 1) // TI
    ntj = (run.ny + TJ - 1) // TJ
    if run.sources is not None and run.sources.n_grp:
        if run.sources.n_grp > TB_MAX_GROUPS or run.sources.inert:
            return False
        cells = run.sources.group_cells.astype(np.int64)
        gi, gj = cells // run.ny, cells % run.ny
        counts = np.zeros((nti, ntj), dtype=np.int64)
        lo_i = np.maximum(0, -(-(gi - TI - depth + 1) // TI))
        hi_i = np.minimum(nti - 1, (gi + depth) // TI)
        lo_j = np.maximum(0, -(-(gj - TJ - depth + 1) // TJ))
        hi_j = np.minimum(ntj - 1, (gj + depth) // TJ)
        for a, b, c, d in zip(lo_i, hi_i, lo_j, hi_j):
            counts[a:b + 1, c:d + 1] += 1
        if counts.max() > TB_MAX_GROUPS_PER_TILE:
            return False
    if run.det_cell is not None and run.det_cell.numel():
        cells = run.extra.get("det_cells_np")
        if cells is None:
            cells = run.det_cell.cpu().numpy().astype(np.int64)
            run.extra["det_cells_np"] = cells
        ti, tj = (cells // run.ny) // TI, (cells % run.ny) // TJ
        _, cnt = np.unique(ti * ntj + tj, return_counts=True)
        if cnt.max() > TB_MAX_DETS_PER_TILE:
            return False
    return True


def _f32_bits(x: float) -> int:
    return int(np.array(x, dtype=np.float32).view(np.uint32))


def division_range(d: float) -> Optional[tuple[int, int]]:
    """Bit range ``[0, span)`` of ``|x|`` on which the fast quotient (the
    reciprocal product with its residual corrected at the scale that puts
    ``d`` at 2^31) is checked against IEEE ``x/d``: every ``|x| < 2^(96+e)``
    for ``d = m*2^e``, below where ``x*2^(31-e)`` would overflow. None for a
    divisor outside ``-95 <= e <= 31`` (no fast path)."""
    bits = _f32_bits(d)
    ed = ((bits >> 23) & 0xFF) - 127
    if not (-95 <= ed <= 31) or (bits >> 31):
        return None
    return 0, (96 + ed + 127) << 23


_DIV_CACHE: dict = {}


def verified_division(d: float, device: torch.device) -> Optional[tuple[float, int, int]]:
    """``(1/d, lo, span)`` if the device proved the fast quotient exact for
    divisor ``fp32(d)`` on that range (exhaustively, once per process), else None."""
    df = f32(d)
    key = (df, str(device))
    if key in _DIV_CACHE:
        return _DIV_CACHE[key]
    rng = division_range(df) if df > 0.0 else None
    if rng is None:
        _DIV_CACHE[key] = None
        return None
    r = float(np.float32(1.0) / np.float32(df))
    lo, span = rng
    bad = torch.zeros(1, dtype=torch.int32, device=device)
    lib = _native.load()
    stream = _native.stream_handle(device)
    _native.check(lib.wf_div_verify(df, r, lo, span, bad.data_ptr(), ctypes.c_void_p(stream)), "wf_div_verify")
    ok = int(bad.item()) == 0
    _DIV_CACHE[key] = (r, lo, span) if ok else None
    return _DIV_CACHE[key]


def split_launches(n_steps: int, depth: int, odd_ok: bool = False) -> Optional[list[int]]:
    """Depths (each in ``[1, depth]``, as equal as possible) of the fewest
    launches covering ``n_steps`` steps; the count is even unless ``odd_ok``
    (a third state buffer) allows an odd count of 3 or more. ``None`` when no
    such split exists (``n_steps < 2``, or odd ``n_steps`` at ``depth == 1``
    without ``odd_ok``)."""
    if n_steps < 2:
        return None
    n_launch = max(2, -(-n_steps // depth))
    if n_launch % 2 and not odd_ok:
        if n_launch == n_steps:
            return None
        n_launch += 1
    base, rem = divmod(n_steps, n_launch)
    return [base + (1 if k < rem else 0) for k in range(n_launch)]


def launch_parity(k: int, n_launch: int) -> int:
    """State-buffer code of launch ``k`` of ``n_launch`` (``launch_parities``)."""
    if n_launch % 2 == 0:
        return k % 2
    return (2, 3, 1)[k] if k < 3 else (k - 3) % 2


def launch_parities(n_launch: int) -> list[int]:
    """State-buffer codes of ``n_launch`` consecutive launches that start and
    end in the live buffers A: 0 (A->B), 1 (B->A) alternating for an even
    count; 2 (A->C), 3 (C->B), 1 (B->A) and then alternating for an odd
    count of 3 or more (needs the third buffer C)."""
    if n_launch % 2 == 0:
        return [k % 2 for k in range(n_launch)]
    if n_launch < 3:
        raise ValueError("a single launch cannot end in the buffer it reads")
    return [2, 3, 1] + [k % 2 for k in range(n_launch - 3)]


class TBPlan:
    """Device run table + per-depth tile lists for the temporally blocked kernel.

    ``TBPlan.event_log``: when set to a list, every launch appends
    ``(depth, n_cells, start_event, end_event)`` with HIP events recorded on the
    launch stream around it (used by the benchmark to time the kernel live);
    with ``event_stride`` k > 1 only every k-th launch is bracketed, which
    keeps the event records' own cost out of the timed stream. ``event_pool``
    (a list of timing events, consumed from the end) supplies the events, so
    a caller can create and first-record them outside its timed region."""

    event_log: Optional[list] = None
    event_stride: int = 1
    event_pool: Optional[list] = None
    _event_count: int = 0

    def __init__(self, runs: Sequence[RunState], tile_i: int, device: torch.device, fast_div: bool = True,
                 cpl: int = 1) -> None:
        self.tile_i = int(tile_i)
        self.cpl = int(cpl)
        self.odd_ok = all(r.third is not None for r in runs)
        self._args: dict = {}
        self._lib = _native.load()
        self._streams: dict = {}
        self.runs = list(runs)
        self.device = device
        mats = {r.ca is not None for r in runs}
        if len(mats) != 1:
            raise ValueError("temporally blocked launch needs all runs with or all without materials")
        self.has_materials = int(mats.pop())
        arr = (WfRun2d * len(runs))(*[r.struct() for r in runs])
        self._run_raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        self._run_table: Optional[torch.Tensor] = None  # uploaded with the first tables (prepare) or on use
        self._tiles: dict[int, tuple[torch.Tensor, int]] = {}
        self._lists: dict = {}
        self.div = None
        recips = {bool(r.reciprocal) for r in runs}
        if len(recips) != 1:
            raise ValueError("temporally blocked launch needs all runs with the same division mode")
        self.reciprocal = recips.pop()
        self.div_mode = _native.WF_DIV_RECIPROCAL if self.reciprocal else _native.WF_DIV_IEEE
        if fast_div and not self.reciprocal:
            dxs = {r.coef.dx for r in runs}
            dys = {r.coef.dy for r in runs}
            if len(dxs) == 1 and len(dys) == 1:
                vx = verified_division(dxs.pop(), device)
                vy = verified_division(dys.pop(), device)
                if vx is not None and vy is not None:
                    self.div = (vx, vy)
                    self.div_mode = _native.WF_DIV_FAST
        self.n_cells = sum(r.nx * r.ny for r in runs)

    @property
    def run_table(self) -> torch.Tensor:
        """The runs' descriptors on the device (``wf_run2d`` array)."""
        if self._run_table is None:
            self._run_table = to_device_async(self._run_raw, self.device)
        return self._run_table

    def prepare(self, max_depth: int, depths: Optional[Sequence[int]] = None) -> None:
        """Upload the tile tables and lists and build the launch arguments of
        ``depths`` (every depth up to ``max_depth`` by default) and every
        state-buffer code now, so no later launch waits on a table upload.
        The tables and lists of all those depths go up in one copy (one
        int32 buffer, every array at a 16-byte boundary) instead of four
        small copies per depth, each of which costs a queue round trip."""
        ds = list(depths if depths is not None else range(1, max_depth + 1))
        todo = [d for d in ds if d not in self._tiles or d not in self._lists]
        if todo:
            host = [a for d in todo for a in ((self._tiles_host(d),) + self._lists_host(d))]
            with_runs = self._run_table is None
            views = upload_packed(([self._run_raw] if with_runs else []) + host, self.device)
            if with_runs:
                self._run_table = views.pop(0)
            for k, d in enumerate(todo):
                t, lists, grp, det = views[4 * k:4 * k + 4]
                self._tiles[d] = (t, int(t.shape[0]))
                self._lists[d] = (lists, grp, det)
        for d in ds:
            for par in ((0, 1, 2, 3) if self.odd_ok else (0, 1)):
                self._base_args(d, par)

    def _tiles_host(self, depth: int) -> np.ndarray:
        """The (run, tile) pairs of a launch of ``depth`` steps, on the host."""
        nts = np.array([tile_count(r.nx, r.ny, self.tile_i, depth, self.cpl) for r in self.runs], dtype=np.int64)
        total = int(nts.sum())
        starts = np.cumsum(nts) - nts
        t = np.empty((total, 2), dtype=np.int32)
        t[:, 0] = np.repeat(np.arange(len(self.runs)), nts)
        t[:, 1] = np.arange(total) - np.repeat(starts, nts)
        return t

    def tiles(self, depth: int) -> tuple[torch.Tensor, int]:
        if depth not in self._tiles:
            t = self._tiles_host(depth)
            self._tiles[depth] = (to_device_async(t, self.device), int(t.shape[0]))
        return self._tiles[depth]

    @staticmethod
    def _det_cells_host(r: RunState) -> np.ndarray:
        """A run's detector cells on the host (kept with the run; read back once if absent)."""
        dc = r.extra.get("det_cells_np")
        if dc is None:
            dc = r.det_cell.cpu().numpy().astype(np.int64)
            r.extra["det_cells_np"] = dc
        return dc

    def tile_lists(self, depth: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """Per-tile source-group and detector lists of a launch of ``depth``
        steps (``wf_tb_args.tile_lists``, ``grp_list``, ``det_list``): the
        groups whose cell lies in a tile's region (its core plus ``depth``
        cells around it) and the detectors in its core, so a tile reads its
        own few entries instead of testing every group and detector."""
        if depth not in self._lists:
            self._lists[depth] = tuple(to_device_async(x, self.device) for x in self._lists_host(depth))
        return self._lists[depth]

    def _lists_host(self, depth: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """``tile_lists``' three arrays on the host (int32), for every run of
        the plan in one vectorised pass: a source group enters every tile
        whose region (core plus ``depth`` cells around it) holds its cell, a
        detector the tile whose core holds it; within a tile both keep their
        run's order. Tiles are numbered as in ``tiles`` (run after run, each
        run's tiles row-major), and list entries are indices into the run's
        own group and detector arrays."""
        TI, TJ = self.tile_i, _native.tile_cols(depth, self.cpl)
        runs = self.runs
        nx = np.array([r.nx for r in runs], dtype=np.int64)
        ny = np.array([r.ny for r in runs], dtype=np.int64)
        nti, ntj = -(-nx // TI), -(-ny // TJ)
        nt = nti * ntj
        toff = np.cumsum(nt) - nt
        total = int(nt.sum())
        # source groups: (run, cell, index in the run) of every group
        g_parts = [(k, r.sources.group_cells) for k, r in enumerate(runs)
                   if r.sources is not None and r.sources.n_grp]
        if g_parts:
            g_n = np.array([c.size for _, c in g_parts], dtype=np.int64)
            g_run = np.repeat(np.array([k for k, _ in g_parts], dtype=np.int64), g_n)
            cells = np.concatenate([c for _, c in g_parts]).astype(np.int64)
            g_loc = np.arange(cells.size) - np.repeat(np.cumsum(g_n) - g_n, g_n)
            gi, gj = cells // ny[g_run], cells % ny[g_run]
            lo_i = np.maximum(0, -(-(gi - TI - depth + 1) // TI))
            hi_i = np.minimum(nti[g_run] - 1, (gi + depth) // TI)
            lo_j = np.maximum(0, -(-(gj - TJ - depth + 1) // TJ))
            hi_j = np.minimum(ntj[g_run] - 1, (gj + depth) // TJ)
            nj = hi_j - lo_j + 1
            cnt = (hi_i - lo_i + 1) * nj
            within = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            tile = (np.repeat(toff[g_run], cnt)
                    + (np.repeat(lo_i, cnt) + within // np.repeat(nj, cnt)) * np.repeat(ntj[g_run], cnt)
                    + np.repeat(lo_j, cnt) + within % np.repeat(nj, cnt))
            order = np.argsort(tile, kind="stable")
            grp = np.repeat(g_loc, cnt)[order]
            g_cnt = np.bincount(tile, minlength=total)
        else:
            grp = np.zeros(0, dtype=np.int64)
            g_cnt = np.zeros(total, dtype=np.int64)
        # detectors: the tile whose core holds the cell
        d_parts = [(k, self._det_cells_host(r)) for k, r in enumerate(runs)
                   if r.det_cell is not None and r.det_cell.numel()]
        if d_parts:
            d_n = np.array([c.size for _, c in d_parts], dtype=np.int64)
            d_run = np.repeat(np.array([k for k, _ in d_parts], dtype=np.int64), d_n)
            dc = np.concatenate([c for _, c in d_parts]).astype(np.int64)
            d_loc = np.arange(dc.size) - np.repeat(np.cumsum(d_n) - d_n, d_n)
            dtile = toff[d_run] + (dc // ny[d_run] // TI) * ntj[d_run] + (dc % ny[d_run]) // TJ
            det = d_loc[np.argsort(dtile, kind="stable")]
            d_cnt = np.bincount(dtile, minlength=total)
        else:
            det = np.zeros(0, dtype=np.int64)
            d_cnt = np.zeros(total, dtype=np.int64)
        g_end, d_end = np.cumsum(g_cnt), np.cumsum(d_cnt)
        lists = np.ascontiguousarray(np.stack([g_end - g_cnt, g_end, d_end - d_cnt, d_end], axis=1).astype(np.int32))
        grp = np.concatenate([grp, np.zeros(1, dtype=np.int64)]).astype(np.int32)
        det = np.concatenate([det, np.zeros(1, dtype=np.int64)]).astype(np.int32)
        return lists, grp, det

    def _base_args(self, depth: int, parity: int) -> WfTbArgs:
        """The launch arguments that do not change from launch to launch."""
        key = (depth, parity)
        a = self._args.get(key)
        if a is None:
            tiles, n_tiles = self.tiles(depth)
            a = WfTbArgs()
            a.runs = self.run_table.data_ptr()
            a.tiles = tiles.data_ptr()
            a.n_runs = len(self.runs)
            a.n_tiles = n_tiles
            a.tile_i = self.tile_i
            a.tile_j = _native.tile_cols(depth, self.cpl)
            a.depth = int(depth)
            a.parity = int(parity)
            a.has_materials = self.has_materials
            lists, grp, det = self.tile_lists(depth)
            a.tile_lists, a.grp_list, a.det_list = lists.data_ptr(), grp.data_ptr(), det.data_ptr()
            if self.div is not None:
                (a.rdx, a.div_lo_x, a.div_span_x), (a.rdy, a.div_lo_y, a.div_span_y) = self.div
            a.fast_div = self.div_mode
            self._args[key] = a
        return a

    def launch(self, depth: int, parity: int, step_offset: int, stream: int,
               maxabs_out: Optional[torch.Tensor] = None, stop: Optional[torch.Tensor] = None) -> None:
        a = self._base_args(depth, parity)
        a.step_offset = int(step_offset)
        a.maxabs_out = maxabs_out.data_ptr() if maxabs_out is not None else 0
        a.stop = stop.data_ptr() if stop is not None else 0
        st = self._streams.get(stream)
        if st is None:
            st = self._streams[stream] = ctypes.c_void_p(stream)
        log = TBPlan.event_log
        if log is not None:
            TBPlan._event_count += 1
            if (TBPlan._event_count - 1) % max(1, TBPlan.event_stride):
                log = None
        if log is not None:
            pool = TBPlan.event_pool
            if pool and len(pool) >= 2:
                ev0, ev1 = pool.pop(), pool.pop()
            else:
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        rc = self._lib.wf_fdtd2d_tb_advance(ctypes.byref(a), st)
        if rc:
            _native.check(rc, "wf_fdtd2d_tb_advance")
        if log is not None:
            ev1.record()
            log.append((int(depth), self.n_cells, ev0, ev1))

    def advance(self, n_steps: int, depth: int, stream: int, base_step: int = 0) -> None:
        """Advance every run ``n_steps`` (>= 2) steps, ending in the live buffers.
        Run ``k`` injects and records step ``table step + base_step + offset``."""
        depths = split_launches(n_steps, depth, odd_ok=self.odd_ok)
        if depths is None:
            raise ValueError(f"cannot split {n_steps} steps into an even number of launches of depth <= {depth}")
        offset = 0
        for d, par in zip(depths, launch_parities(len(depths))):
            self.launch(d, par, base_step + offset, stream)
            offset += d


def judge(maxabs: torch.Tensor, threshold: float, stop: torch.Tensor, all_runs: bool, stream: int) -> None:
    """Device-side stability judgement of one check slot (``wf_fdtd2d_judge``)."""
    lib = _native.load()
    n_runs = stop.numel()
    _native.check(lib.wf_fdtd2d_judge(maxabs.data_ptr(), n_runs, float(threshold), 1 if all_runs else 0,
                                      stop.data_ptr(), ctypes.c_void_p(stream)), "wf_fdtd2d_judge")


def schedule_segments(start: int, n_steps: int, depth: int, check_every: Optional[int], odd_ok: bool = False):
    """The launch plan of ``schedule_launches`` in compact form: returns
    ``(segments, checks)`` with ``segments`` a list of ``[steps, launches,
    slot]`` (or ``None`` if no valid plan exists), cheap to build for long
    runs; ``iter_launches`` expands it."""
    end = start + n_steps
    checks = []
    if check_every:
        m = (start // check_every + 1) * check_every
        while m <= end:
            checks.append(m)
            m += check_every
    segs = []
    prev = start
    for k, b in enumerate(checks + ([end] if not checks or checks[-1] != end else [])):
        if b - prev > 0:
            segs.append([b - prev, -(-(b - prev) // depth), k if k < len(checks) else -1])
        prev = b
    total = sum(n for _, n, _ in segs)
    if total % 2 and not (odd_ok and total >= 3):
        # one more launch in the segment where it costs least: the one whose
        # launches are deepest (its split gets closest to equal depths)
        cand = [k for k, (seg, n, _) in enumerate(segs) if n < seg]
        if not cand:
            return None, checks
        k = max(cand, key=lambda i: (segs[i][0] / (segs[i][1] + 1), segs[i][0]))
        segs[k][1] += 1
    return segs, checks


def iter_launches(segs):
    """``(depth, slot)`` of every launch of a segment plan, in order: each
    segment's launches as equal as possible, the last one feeding the
    segment's check slot (or -1)."""
    for seg, n_l, slot in segs:
        base, rem = divmod(seg, n_l)
        for k in range(n_l):
            yield base + (1 if k < rem else 0), (slot if k == n_l - 1 else -1)


def schedule_launches(start: int, n_steps: int, depth: int, check_every: Optional[int], odd_ok: bool = False):
    """Launch plan for ``n_steps`` steps from step ``start``.

    Returns ``(launches, checks)``: ``launches`` is a list of ``(depth, slot)``
    where ``slot`` (or -1) is the stability-check slot the launch's final
    level feeds, ``checks`` the global step of each slot. Segments end at every
    step ``m`` with ``m % check_every == 0`` and are split into the fewest
    launches of at most ``depth`` steps, as equal as possible. The launch count
    is even (so the state ends in the live buffers) unless ``odd_ok`` (a third
    state buffer) allows an odd count of 3 or more; to make it even, the
    segment with the most launches is split into one launch more. ``None`` is
    returned if that is impossible (every launch of depth 1 and an odd
    count)."""
    segs, checks = schedule_segments(start, n_steps, depth, check_every, odd_ok)
    if segs is None:
        return None, checks
    return list(iter_launches(segs)), checks


def runs_add_step(run_table: torch.Tensor, n_runs: int, delta: int, stream: int) -> None:
    """Add ``delta`` to every device run descriptor's ``step`` (``wf_runs_add_step``);
    lets a captured step loop be replayed for later step ranges."""
    lib = _native.load()
    _native.check(lib.wf_runs_add_step(run_table.data_ptr(), int(n_runs), int(delta), ctypes.c_void_p(stream)),
                  "wf_runs_add_step")
