"""Batched multi-run driver: many independent 2D TM runs in one launch.

The reference images a phantom by looping over transmitters and running one
``FDTD2D`` per TX with a per-step Python loop and a detector gather
(``run_mimo``, ``examples/brain_mimo_imaging.py:149-204``). ``BatchedRuns``
advances any number of runs — different grid sizes, materials, sources and
receivers — together: every launch of the temporally blocked kernel covers
the tiles of all runs, so small grids fill the GPU and the per-step launch
overhead disappears. Each run behaves exactly as if it were stepped alone
(the kernel reads only that run's descriptor and tiles), which the tests
check bit for bit against the oracle.

Stability checks are judged on the device per run every ``n_check`` steps:
a diverged run is flagged and carried through unchanged while the others
continue (failure isolation), and ``status()`` reports it.

``use_graph=True`` captures ``chunk`` steps of launches (ending with a
kernel that advances the runs' step index) in a HIP graph through
``torch.cuda.CUDAGraph`` and replays it. A replayed chunk judges stability
once, at its end (a step that is a multiple of ``chunk``, not of
``n_check``), on the same device flags.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native
from .engine import (
    CellSource,
    RunState,
    SourceTable,
    StepCoefficients,
    TBPlan,
    check_division,
    judge,
    runs_add_step,
    schedule_launches,
    tb_fits_any_tile,
    tb_limits_ok,
)


@dataclass
class RunSpec:
    """One run: grid size, optional per-cell Ca/Cb ``(nx, ny)`` on the device,
    soft point sources ``(component, i, j, wave, amp)`` and Hz receivers."""

    nx: int
    ny: int
    ca: Optional[torch.Tensor]
    cb: Optional[torch.Tensor]
    sources: Sequence[tuple] = ()
    receivers: Sequence[tuple[int, int]] = ()
    n_record: int = 0
    tag: object = None


class BatchedRuns:
    """Device state and launch plan of a set of independent runs.

    ``division`` is how every run divides by dx, dy (``"reciprocal"``, as the
    reference computes on a GPU, or ``"ieee"``, as on the CPU; see
    ``core.fdtd2d``)."""

    def __init__(self, specs: Sequence[RunSpec], *, dx: float, dy: float, dt: float, device=None,
                 tile_rows=None, depth=None, n_check: int = 500, stability_threshold: float = 1e10,
                 fast_div: bool = True, division: str = "reciprocal") -> None:
        if not specs:
            raise ValueError("no runs")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("waveforge_amd: batched runs need the ROCm GPU; there is no CPU path")
        _native.load()
        self.specs = list(specs)
        self.dt, self.dx, self.dy = dt, dx, dy
        self.tile_rows, self.depth = tile_rows, depth
        self.n_check, self.threshold = n_check, stability_threshold
        self.division = check_division(division)
        coef = StepCoefficients.from_grid(dt, dx, dy)
        self.runs: list[RunState] = []
        dev = self.device
        # one zeroed block for every run's two states, one for every trace buffer,
        # one upload for all detector cells and one for all source tables
        for k, sp in enumerate(self.specs):
            if sp.nx < 4 or sp.ny < 4:
                raise ValueError(f"run {k}: grid must be at least 4x4, got {sp.nx}x{sp.ny}")
            for (c, i, j, _w, _a) in sp.sources:
                if not (0 <= i < sp.nx and 0 <= j < sp.ny):
                    raise ValueError(f"run {k}: source cell ({i}, {j}) outside the {sp.nx}x{sp.ny} grid")
            for (i, j) in sp.receivers:
                if not (0 <= i < sp.nx and 0 <= j < sp.ny):
                    raise ValueError(f"run {k}: receiver ({i}, {j}) outside the {sp.nx}x{sp.ny} grid")
        cells = [sp.nx * sp.ny for sp in self.specs]
        self._state = torch.zeros(6 * sum(cells), dtype=torch.float32, device=dev)
        n_rec = [len(sp.receivers) * max(1, sp.n_record) if sp.receivers else 0 for sp in self.specs]
        self._trace_buf = torch.zeros(max(1, sum(n_rec)), dtype=torch.float32, device=dev)
        det_host = [np.array([i * sp.ny + j for i, j in sp.receivers], dtype=np.int32) if sp.receivers
                    else np.zeros(0, dtype=np.int32) for sp in self.specs]
        det_all = torch.from_numpy(np.concatenate(det_host) if det_host else np.zeros(0, np.int32)).to(dev)
        tables = []
        o_s = o_t = o_d = 0
        for k, sp in enumerate(self.specs):
            n = cells[k]
            f = self._state[o_s:o_s + 6 * n].view(6, sp.nx, sp.ny)
            o_s += 6 * n
            if len(sp.sources) == 1:  # the common case, built directly
                c, i, j, w, a = sp.sources[0]
                table = SourceTable.cell(c, w, i, j, a, sp.nx, sp.ny)
            else:
                srcs = [CellSource(c, w, i, j, a) for (c, i, j, w, a) in sp.sources]
                table = SourceTable(srcs, sp.nx, sp.ny, dev, upload=False)
            tables.append(table)
            rs = RunState(f[0], f[1], f[2], sp.ca, sp.cb, sp.nx, sp.ny, coef, sources=table,
                          reciprocal=self.division == "reciprocal")
            rs.second = (f[3], f[4], f[5])
            if sp.receivers:
                nd = len(sp.receivers)
                rs.det_cell = det_all[o_d:o_d + nd]
                rs.extra["det_cells_np"] = det_host[k].astype(np.int64)
                rs.traces = self._trace_buf[o_t:o_t + n_rec[k]].view(nd, max(1, sp.n_record))
                rs.trace_n0 = 0
                o_d += nd
                o_t += n_rec[k]
            rs.step = 0
            self.runs.append(rs)
        SourceTable.upload_many(tables, dev)
        from .core.fdtd2d import default_launch_shape, tile_shape

        t0, d0, c0 = default_launch_shape(self.runs)  # what is not given comes from FDTD2D's default
        cpl = c0
        if tile_rows is None:
            tile_rows = t0
        elif tile_rows != "auto":
            tile_rows, cpl = tile_shape(tile_rows)
        depth = d0 if depth is None else depth
        self.tuned = None
        if "auto" in (tile_rows, depth):
            from . import autotune

            self.tuned = autotune.tune_runs(self.runs, self.device, fast_div=fast_div,
                                            tile_rows=None if tile_rows == "auto" else [(tile_rows, cpl)],
                                            depths=autotune.DEPTHS if depth == "auto" else [depth])
            tile_rows, depth, cpl = self.tuned.tile_rows, self.tuned.depth, self.tuned.cpl
        self.tile_rows, self.depth, self.cpl = tile_rows, depth, cpl
        for k, rs in enumerate(self.runs):
            if tb_fits_any_tile(rs):
                continue
            for d in range(1, depth + 1):
                if not tb_limits_ok(rs, tile_rows, d, cpl):
                    raise ValueError(f"run {k}: too many sources or receivers per tile for the batched kernel")
        self.plan = TBPlan(self.runs, tile_rows, self.device, fast_div=fast_div, cpl=cpl)
        self.steps_completed = 0
        self._dev_step = 0  # value of every descriptor's device-side `step` field
        self.stop = torch.zeros(len(self.runs), dtype=torch.int32, device=self.device)
        self.check_log: list[tuple[int, torch.Tensor]] = []
        self._graph = None
        self._graph_checks = 0

    # ------------------------------------------------------------------
    @property
    def n_cells(self) -> int:
        return sum(r.nx * r.ny for r in self.runs)

    def traces(self, k: int) -> torch.Tensor:
        return self.runs[k].traces

    def fields(self, k: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        r = self.runs[k]
        return r.ex, r.ey, r.hz

    def _stream(self) -> int:
        return _native.stream_handle(self.device)

    def _check_waves(self, n_steps: int) -> None:
        end = self.steps_completed + n_steps
        for k, r in enumerate(self.runs):
            t = r.sources
            if t is not None and t.min_wave_len is not None and end > t.min_wave_len:
                raise RuntimeError(f"run {k}: source waveform has {t.min_wave_len} samples, cannot reach step {end}")

    def advance(self, n_steps: int) -> None:
        """Enqueue ``n_steps`` steps of every run (asynchronous)."""
        if n_steps <= 0:
            return
        self._check_waves(n_steps)
        launches, checks = schedule_launches(self.steps_completed, n_steps, self.depth, self.n_check)
        if launches is None:
            raise ValueError("batched advance needs an even launch split (use n_steps >= 2 or depth >= 2)")
        stream = self._stream()
        slots = None
        if checks:
            slots = torch.zeros((len(checks), len(self.runs), 3), dtype=torch.int32, device=self.device)
        offset = self.steps_completed - self._dev_step
        for k, (d, slot) in enumerate(launches):
            self.plan.launch(d, k % 2, offset, stream, maxabs_out=slots[slot] if slot >= 0 else None,
                             stop=self.stop)
            if slot >= 0:
                judge(slots[slot], self.threshold, self.stop, False, stream)
            offset += d
        for m, step in enumerate(checks):
            self.check_log.append((step, slots[m]))
        self.steps_completed += n_steps

    # ------------------------------------------------------------------
    def capture(self, chunk: int) -> None:
        """Capture ``chunk`` steps in a HIP graph. The chunk's last launch
        folds every run's max |Ex|, |Ey|, |Hz| into a slot that keeps the
        running maximum over all replays (zeroed once, here, outside the
        graph), and the judge flags a run whose maximum passed the threshold,
        so a run diverged at any replayed chunk end is stopped and reported."""
        launches, _ = schedule_launches(0, chunk, self.depth, None)
        if launches is None:
            raise ValueError("graph chunk needs an even launch split")
        # the tile tables and lists of the chunk's depths, uploaded before capture
        self.plan.prepare(max(d for d, _ in launches), {d for d, _ in launches})
        self._graph_slot = torch.zeros((len(self.runs), 3), dtype=torch.int32, device=self.device)
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            # warm the launch path (table uploads) outside the capture
            torch.cuda.synchronize(self.device)
            with torch.cuda.graph(g, stream=s):
                stream = _native.stream_handle(self.device)
                off = 0
                for k, (d, _) in enumerate(launches):
                    last = k == len(launches) - 1
                    self.plan.launch(d, k % 2, off, stream, maxabs_out=self._graph_slot if last else None,
                                     stop=self.stop)
                    off += d
                judge(self._graph_slot, self.threshold, self.stop, False, stream)
                runs_add_step(self.plan.run_table, len(self.runs), chunk, stream)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph = (g, chunk)

    def replay(self, n_chunks: int) -> None:
        """Replay the captured chunk ``n_chunks`` times (device step index advances)."""
        if self._graph is None:
            raise RuntimeError("capture() first")
        g, chunk = self._graph
        self._check_waves(n_chunks * chunk)
        if self.steps_completed != self._dev_step:
            # bring the device step index up to date after eager advances
            runs_add_step(self.plan.run_table, len(self.runs), self.steps_completed - self._dev_step, self._stream())
            self._dev_step = self.steps_completed
        for _ in range(n_chunks):
            g.replay()
        if n_chunks > 0 and not self._graph_checks:
            self.check_log.append((-1, self._graph_slot))  # running maximum over replayed chunk ends
        self._graph_checks += n_chunks
        self.steps_completed += n_chunks * chunk
        self._dev_step += n_chunks * chunk

    # ------------------------------------------------------------------
    def status(self) -> dict:
        """Synchronise and report per-run stability: ``{"diverged": [...],
        "checks": n, "field_max": [...]}`` (field_max is the largest checked
        max|E|, |H| per run, as ``last_field_max``)."""
        stop = self.stop.cpu().numpy().astype(bool)
        fmax = np.full(len(self.runs), -1.0)
        for _, slot in self.check_log:
            v = slot.view(torch.float32).cpu().numpy().max(axis=1)
            fmax = np.maximum(fmax, np.where(np.isnan(v), -1.0, v))
        n_checks = len(self.check_log) - (1 if self._graph_checks else 0) + self._graph_checks
        return {"diverged": stop.tolist(), "checks": n_checks, "field_max": fmax.tolist()}


# ---------------------------------------------------------------------------
# MIMO imaging on top of the batched runs
# ---------------------------------------------------------------------------


def mimo_specs(ca: torch.Tensor, cb: torch.Tensor, antennas, wave: torch.Tensor, n_steps: int, tag=None):
    """One RunSpec per TX antenna: Hz point source ``wave`` at the TX, all
    antennas record Hz (``examples/brain_mimo_imaging.py:176-201``)."""
    nx, ny = ca.shape[0], ca.shape[1]
    ca2, cb2 = ca.reshape(nx, ny), cb.reshape(nx, ny)
    return [RunSpec(nx, ny, ca2, cb2, sources=[("Hz", i, j, wave, 1.0)], receivers=list(antennas),
                    n_record=n_steps, tag=(tag, t)) for t, (i, j) in enumerate(antennas)]


def run_mimo(grid, Ca, Cb, antenna_positions, n_steps: int, waveform=None, *, depth=None,
             tile_rows=None, n_check: int = 500, return_tensor: bool = False, division: str = "reciprocal"):
    """All TX runs of one array in one batched launch sequence.

    Same inputs and output as the reference's ``run_mimo``: returns the
    ``(N_tx, N_rx, n_steps)`` float32 signals (NumPy unless
    ``return_tensor``). The default waveform is the reference's 1 GHz Ricker
    with ``t0 = 1.5/f`` built for ``n_steps`` samples.
    """
    from .core.sources import RickerWavelet

    if waveform is None:
        waveform = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9)
    wave = waveform.build(n_steps, grid.dt, grid.device, torch.float32)
    specs = mimo_specs(Ca, Cb, antenna_positions, wave, n_steps)
    br = BatchedRuns(specs, dx=grid.dx, dy=grid.dy, dt=grid.dt, device=grid.device, tile_rows=tile_rows,
                     depth=depth, n_check=n_check, division=division)
    br.advance(n_steps)
    st = br.status()
    if any(st["diverged"]):
        from .core.fdtd2d import SimulationDivergedError

        bad = [k for k, v in enumerate(st["diverged"]) if v]
        raise SimulationDivergedError(f"TX runs {bad} diverged")
    out = torch.stack([br.traces(k) for k in range(len(specs))])
    return out if return_tensor else out.cpu().numpy()


def image_phantoms(phantoms, n_steps: int, *, dx: float = 1e-3, dt: float, device, depth=None,
                   tile_rows=None, n_check: int = 500, v_medium: float | None = None, gate_start: int = 0,
                   use_graph: bool = False, chunk: int = 240, division: str = "reciprocal"):
    """Batched MIMO imaging of many phantoms: all TX runs of all phantoms in
    one batched launch sequence, then the DAS images of all phantoms together
    (``wf_das2d_batch``; raw traces).

    ``phantoms`` is a list of ``(eps_r, sigma, antennas)`` (NumPy maps in
    cells, antenna cells). Returns ``(images, traces, batched_runs)`` with
    ``images[p]`` a device fp64 ``(n, n)`` image and ``traces[p]`` the
    ``(N_tx, N_rx, n_steps)`` device signals.

    With ``use_graph`` the first ``q`` whole chunks of ``chunk`` steps are
    replayed from a HIP graph and the rest runs eagerly; a remainder of one
    step (which no even launch split covers) joins the last chunk, run
    eagerly. Replayed chunks judge stability at their ends (multiples of
    ``chunk``), eager steps at multiples of ``n_check``.
    """
    from .core.imaging import DasBatch
    from .core.materials import coefficients_from_maps
    from .core.sources import RickerWavelet
    from .engine import to_device_async

    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9).build(n_steps, dt, device, torch.float32)
    specs, owners = [], []
    for p, (eps, sig, ants) in enumerate(phantoms):
        ca, cb = coefficients_from_maps(eps, sig, dt, device=device)
        sp = mimo_specs(ca, cb, ants, wave, n_steps, tag=p)
        specs += sp
        owners += [p] * len(sp)
    br = BatchedRuns(specs, dx=dx, dy=dx, dt=dt, device=device, tile_rows=tile_rows, depth=depth,
                     n_check=n_check, division=division)
    n_chunks, rest = divmod(n_steps, chunk) if use_graph else (0, n_steps)
    if n_chunks and rest == 1:
        n_chunks, rest = n_chunks - 1, rest + chunk
    if n_chunks:
        br.capture(chunk)
        br.replay(n_chunks)
    br.advance(rest)
    if v_medium is None:
        v_medium = 299_792_458.0 / math.sqrt(40.0)
    # antenna coordinates of every phantom in one upload
    xy = to_device_async(np.array([[a[0] * dx, a[1] * dx] for _, _, ants in phantoms for a in ants],
                                  dtype=np.float64).T.copy(), device)
    das = DasBatch(device)
    images, traces = [], []
    k = 0
    for p, (eps, sig, ants) in enumerate(phantoms):
        n_ant = len(ants)
        S = _trace_block(br, k, n_ant)
        nx, ny = eps.shape
        img = torch.empty((nx, ny), dtype=torch.float64, device=device)
        das.add(S, xy[0, k:k + n_ant], xy[1, k:k + n_ant], img, pitch_x=nx * dx / nx, pitch_y=ny * dx / ny,
                v_medium=v_medium, dt=dt, gate_start=gate_start, clip=True)
        k += n_ant
        images.append(img)
        traces.append(S)
    das.launch()
    return images, traces, br


def _trace_block(br: BatchedRuns, k0: int, n: int) -> torch.Tensor:
    """The ``(n, n_rx, n_record)`` traces of runs ``k0 .. k0+n-1``: a view of
    the batch's trace buffer when those runs' traces lie back to back in it
    (as ``BatchedRuns`` lays them out), else a stacked copy."""
    ts = [br.traces(k0 + t) for t in range(n)]
    t0 = ts[0]
    step = t0.numel() * t0.element_size()
    if all(t.shape == t0.shape and t.is_contiguous() and t.data_ptr() == t0.data_ptr() + i * step
           for i, t in enumerate(ts)):
        return t0.as_strided((n,) + tuple(t0.shape), (t0.numel(),) + tuple(t0.stride()), t0.storage_offset())
    return torch.stack(ts)
