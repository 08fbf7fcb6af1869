"""Continuous-batching imaging server.

A request is one MIMO imaging job: a phantom (eps_r, sigma maps), an array of
E antennas (every element transmits once, all receive Hz) and a step count;
its answer is the DAS image of its E x E x T traces. The reference serves
nothing: it images one phantom at a time with a sequential per-TX Python
loop (``examples/brain_mimo_imaging.py:149-204, 211-266``).

The server keeps every admitted request's field state on the device and
advances the active runs together, one quantum of steps per scheduling round,
through the temporally blocked multi-run kernel: requests that arrive while
others are on the GPU are admitted once their HBM footprint fits the budget
and join the rounds in arrival order, as many as fill a tile budget (and
leave as soon as their runs reach their step count), so work of different
requests overlaps inside the same launches. Each run reads only
its own descriptor and tiles, so a request's result does not depend on what
else is in flight (the tests check alone == in the mix, bit for bit).
Stability is judged on the device per run at every round; a diverged request
is reported failed without stopping the others.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import _native
from .core.imaging import DasBatch, run_das
from .core.materials import coefficients_from_maps
from .core.sources import RickerWavelet
from .engine import (
    RunState,
    SourceTable,
    StepCoefficients,
    TBPlan,
    check_division,
    judge,
    runs_add_step,
    tb_fits_any_tile,
    tb_limits_ok,
    tile_count,
    to_device_async,
)

C0 = 299_792_458.0


@dataclass
class ImagingRequest:
    """One imaging job. ``arrival`` is in seconds from the start of the stream."""

    rid: int
    arrival: float
    eps: np.ndarray
    sigma: np.ndarray
    antennas: Sequence[tuple[int, int]]
    n_steps: int
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    image: Optional[torch.Tensor] = None
    traces: Optional[torch.Tensor] = None
    failed: bool = False
    _event: object = None
    _failed_dev: object = None
    _runs: list = field(default_factory=list)
    _state: object = None
    _bytes: int = 0

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival

    @property
    def cell_updates(self) -> int:
        n = self.eps.shape[0] * self.eps.shape[1]
        return n * self.n_steps * len(self.antennas)


def request_footprint(nx: int, ny: int, n_ant: int, n_steps: int, quantum: int = 0) -> int:
    """Device bytes a request holds while active: two field states per TX run,
    shared Ca/Cb, the traces, the waveform and the image."""
    cells = nx * ny
    return (6 * cells * n_ant + 2 * cells) * 4 + n_ant * n_ant * n_steps * 4 + n_steps * 4 + cells * 8


class ImagingServer:
    """Admit, co-schedule and retire imaging requests on one GPU."""

    def __init__(self, *, dt: float, dx: float = 1e-3, device=None, hbm_budget: float = 64e9, quantum: int = 192,
                 depth: int = 8, tile_rows: int = 64, max_requests: Optional[int] = None,
                 stability_threshold: float = 1e10, v_medium: Optional[float] = None,
                 division: str = "reciprocal", admit_slice: float = 2e-3,
                 fcfs_tiles: Optional[int] = 16384) -> None:
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("waveforge_amd: the imaging server needs the ROCm GPU; there is no CPU path")
        self.division = check_division(division)
        _native.load()
        self.dt, self.dx = dt, dx
        self.budget = hbm_budget
        self.quantum = quantum
        self.depth, self.tile_rows = depth, tile_rows
        self.max_requests = max_requests
        self.admit_slice = admit_slice  # host seconds of admissions between two rounds (with requests active)
        # each round advances the oldest admitted requests whose runs make at
        # least this many tiles (enough to fill the GPU many times over), the
        # rest waiting in HBM; None advances every admitted request (processor
        # sharing), under which the largest requests share every round with
        # each newcomer: at 80-90 % load their p99 is about 1.4x as long
        self.fcfs_tiles = fcfs_tiles
        self.threshold = stability_threshold
        self.v_medium = v_medium if v_medium is not None else C0 / math.sqrt(40.0)
        self.coef = StepCoefficients.from_grid(dt, dx, dx)
        self.active: list[ImagingRequest] = []
        self.retiring: list[ImagingRequest] = []
        self.done: list[ImagingRequest] = []
        self.used = 0
        self._plan: Optional[TBPlan] = None
        self._plan_runs: list = []
        self._stop: Optional[torch.Tensor] = None
        self.rounds = 0
        if quantum % 2 or quantum < 2 * depth:
            raise ValueError("quantum must be even and at least 2*depth steps")

    # ------------------------------------------------------------------
    def footprint(self, req: ImagingRequest) -> int:
        nx, ny = req.eps.shape
        return request_footprint(nx, ny, len(req.antennas), req.n_steps, self.quantum)

    def can_admit(self, req: ImagingRequest) -> bool:
        if self.max_requests is not None and len(self.active) >= self.max_requests:
            return False
        fp = self.footprint(req)
        if fp > self.budget:
            raise ValueError(f"request {req.rid} needs {fp/1e9:.2f} GB, above the {self.budget/1e9:.2f} GB budget")
        return self.used + fp <= self.budget

    def admit(self, req: ImagingRequest, now: float) -> None:
        dev = self.device
        nx, ny = req.eps.shape
        # tables built on the host and staged to the device without waiting on it
        ca, cb = coefficients_from_maps(req.eps, req.sigma, self.dt)
        ca = to_device_async(ca.reshape(nx, ny).numpy(), dev)
        cb = to_device_async(cb.reshape(nx, ny).numpy(), dev)
        wave = to_device_async(RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9).build(
            req.n_steps, self.dt, torch.device("cpu"), torch.float32).numpy(), dev)
        if any(not (0 <= i < nx and 0 <= j < ny) for i, j in req.antennas):
            raise ValueError(f"request {req.rid}: antenna cell outside the {nx}x{ny} grid")
        cells_host = np.array([i * ny + j for i, j in req.antennas], dtype=np.int64)
        cells = to_device_async(cells_host.astype(np.int32), dev)
        n_ant = len(req.antennas)
        req.traces = torch.zeros((n_ant, n_ant, req.n_steps), dtype=torch.float32, device=dev)
        runs, tables = [], []
        state = torch.zeros((n_ant, 6, nx, ny), dtype=torch.float32, device=dev)  # both states of every TX run
        for t, (i, j) in enumerate(req.antennas):
            f = state[t]
            table = SourceTable.cell("Hz", wave, i, j, 1.0, nx, ny)
            tables.append(table)
            rs = RunState(f[0], f[1], f[2], ca, cb, nx, ny, self.coef, sources=table, det_cell=cells,
                          traces=req.traces[t], trace_n0=0, step=0, step_end=req.n_steps,
                          reciprocal=self.division == "reciprocal")
            rs.second = (f[3], f[4], f[5])
            rs.extra["req"] = req
            rs.extra["det_cells_np"] = cells_host  # the receivers' cells for host-side tables (no read-back)
            rs.extra["slot"] = -1  # index in the current plan's stop flags (-1: not yet in a plan)
            for d in range(1, self.depth + 1):
                if not tb_fits_any_tile(rs) and not tb_limits_ok(rs, self.tile_rows, d):
                    raise ValueError(f"request {req.rid}: receivers too dense for the batched kernel")
            runs.append(rs)
        SourceTable.upload_many(tables, dev)
        req._runs = runs
        req._state = state
        req._bytes = self.footprint(req)
        req.admitted = now
        self.used += req._bytes
        self.active.append(req)
        self._plan = None

    # ------------------------------------------------------------------
    def _build_plan(self) -> None:
        """Plan over the active runs; their stop flags move to the new plan on
        the device (a gather from the old flags), so no admission or
        retirement waits for the GPU."""
        runs = []
        tiles = 0
        for req in self.active:  # admission order
            runs += req._runs
            if self.fcfs_tiles is not None:
                nx, ny = req.eps.shape
                tiles += len(req._runs) * tile_count(nx, ny, self.tile_rows, self.depth)
                if tiles >= self.fcfs_tiles:
                    break
        if not runs:
            self._plan = self._stop = None
            self._plan_runs = []
            return
        old = self._stop
        keep = [r.extra["slot"] for r in runs]
        if old is not None and any(k >= 0 for k in keep):
            idx = to_device_async([max(k, 0) for k in keep], self.device, torch.long)
            fresh = to_device_async([k < 0 for k in keep], self.device, torch.bool)
            stop = torch.where(fresh, torch.zeros_like(idx, dtype=torch.int32), old.index_select(0, idx))
        else:
            stop = torch.zeros(len(runs), dtype=torch.int32, device=self.device)
        for k, r in enumerate(runs):
            r.extra["slot"] = k
        self._plan_runs = runs
        self._plan = TBPlan(runs, self.tile_rows, self.device)
        self._stop = stop.contiguous()

    def advance_round(self) -> None:
        """Advance every active run by one quantum (asynchronous) and judge it."""
        if not self.active:
            return
        if self._plan is None:
            self._build_plan()
        stream = _native.stream_handle(self.device)
        q, d = self.quantum, self.depth
        n_l = -(-q // d)
        if n_l % 2:
            n_l += 1
        base, rem = divmod(q, n_l)
        depths = [base + (1 if k < rem else 0) for k in range(n_l)]
        slot = torch.zeros((len(self._plan_runs), 3), dtype=torch.int32, device=self.device)
        off = 0
        for k, dd in enumerate(depths):
            last = k == len(depths) - 1
            self._plan.launch(dd, k % 2, off, stream, maxabs_out=slot if last else None, stop=self._stop)
            off += dd
        judge(slot, self.threshold, self._stop, False, stream)
        runs_add_step(self._plan.run_table, len(self._plan_runs), q, stream)
        for r in self._plan_runs:
            r.step += q
        self.rounds += 1

    def retire_finished(self) -> None:
        """Requests whose runs reached their step count leave the active set;
        their DAS image is enqueued and completion is tracked by an event.
        Each run stopped at its own last step inside the round (its step_end),
        so a request is judged on its state at that step, not on steps past it."""
        finished = [req for req in self.active if req._runs[0].step >= req.n_steps]
        if not finished:
            return
        self._finish(finished)
        self.active = [req for req in self.active if req not in finished]
        self._build_plan()

    def _finish(self, reqs: list) -> None:
        """Enqueue the DAS images of every request in ``reqs`` together (one
        ``wf_das2d_batch`` call, antenna coordinates in one upload) and the
        copies of their stop flags, then one event they all complete on."""
        dev = self.device
        xy = to_device_async(np.array([[a[0] * self.dx, a[1] * self.dx] for req in reqs for a in req.antennas],
                                      dtype=np.float64).T.copy(), dev)
        das = DasBatch(dev)
        k = 0
        for req in reqs:
            nx, ny = req.eps.shape
            n = len(req.antennas)
            req.image = torch.empty((nx, ny), dtype=torch.float64, device=dev)
            das.add(req.traces, xy[0, k:k + n], xy[1, k:k + n], req.image, pitch_x=self.dx, pitch_y=self.dx,
                    v_medium=self.v_medium, dt=self.dt, clip=True)
            k += n
        das.launch()
        for req in reqs:
            lo, hi = req._runs[0].extra["slot"], req._runs[-1].extra["slot"] + 1
            req._failed_dev = self._stop[lo:hi].clone()  # read once the image is ready (no wait here)
        ev = torch.cuda.Event()
        ev.record()
        for req in reqs:
            req._event = ev
            self.retiring.append(req)

    def poll(self, now_fn) -> None:
        """Complete retiring requests whose image is ready; free their memory."""
        still = []
        for req in self.retiring:
            if req._event.query():
                req.finished = now_fn()
                req.failed = bool(req._failed_dev.any())
                req._failed_dev = None
                self.used -= req._bytes
                for r in req._runs:
                    r.ex = r.ey = r.hz = r.second = None
                req._runs = []
                req._state = None
                self.done.append(req)
            else:
                still.append(req)
        self.retiring = still

    # ------------------------------------------------------------------
    def serve(self, requests: Sequence[ImagingRequest], *, realtime: bool = True,
              progress: Optional[Callable[[int, int, float], None]] = None, progress_every: float = 20.0) -> dict:
        """Serve a stream of requests (sorted by arrival). With ``realtime``
        the arrivals follow the wall clock from the call; returns latency and
        throughput statistics. ``progress(done, admitted, seconds)`` is
        called about every ``progress_every`` seconds of a long stream."""
        # arrival order; requests that arrive together, the largest first (the
        # longest jobs of a burst start first and the short ones fill the end)
        pending = sorted(requests, key=lambda r: (r.arrival, -r.cell_updates))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        k = 0
        last = 0.0
        ahead: collections.deque = collections.deque()  # an event at the end of each round in flight
        while k < len(pending) or self.active or self.retiring:
            if progress is not None and now() - last >= progress_every:
                last = now()
                progress(len(self.done), k, last)
            t = now()
            while k < len(pending) and (pending[k].arrival <= t or not realtime) and self.can_admit(pending[k]):
                self.admit(pending[k], t)
                k += 1
                # a burst is admitted a slice at a time between rounds, so the
                # GPU starts on the first requests while the host builds the rest
                if self.active and now() - t >= self.admit_slice:
                    break
            if self.active:
                self.advance_round()
                self.retire_finished()
                # keep the host at most two rounds ahead of the device without
                # draining it: wait for the end of the round before the last,
                # so the next round's launches queue while the last one runs
                ev = torch.cuda.Event()
                ev.record()
                ahead.append(ev)
                if len(ahead) > 2:
                    ahead.popleft().synchronize()
            elif self.retiring:
                torch.cuda.current_stream(self.device).synchronize()
            elif k < len(pending):
                wait = pending[k].arrival - now()
                if wait > 0:
                    time.sleep(min(wait, 0.05))
            self.poll(now)
        torch.cuda.synchronize(self.device)
        self.poll(now)
        return _stats(self.done, now(), self.rounds)


def _stats(done, total: float, rounds: int = 0) -> dict:
    lat = np.array([r.latency for r in done if r.latency is not None])
    cu = sum(r.cell_updates for r in done)
    return {
        "requests": len(done),
        "failed": int(sum(r.failed for r in done)),
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
        "mean_ms": float(lat.mean() * 1e3) if lat.size else None,
        "wall_s": total,
        "images_per_s": len(done) / total if total > 0 else None,
        "mcell_updates_per_s": cu / total / 1e6 if total > 0 else None,
        "rounds": rounds,
    }


def serve_plain(requests: Sequence[ImagingRequest], *, dt: float, dx: float = 1e-3, device=None,
                realtime: bool = True, v_medium: Optional[float] = None, n_check: int = 100,
                stability_threshold: float = 1e10, division: str = "reciprocal",
                progress: Optional[Callable[[int, float], None]] = None, progress_every: float = 20.0) -> dict:
    """The plain one-run-per-launch GPU path as a server, for comparison.

    Requests are taken one at a time in arrival order; every TX of a request
    is its own ``core.FDTD2D`` run (the temporally blocked kernel over that
    one grid, one run per launch sequence, stability checked every
    ``n_check`` steps), as a user gets by moving the reference's per-TX
    ``run_mimo`` loop (``examples/brain_mimo_imaging.py:149-204``) onto the
    GPU; then DAS. Fills ``image``, ``traces``, ``finished`` of each request
    and returns the same statistics as ``ImagingServer.serve``;
    ``progress(done, seconds)`` is called about every ``progress_every``
    seconds."""
    from .core import DetectorSet, FDTD2D, FieldSet, MurABC, PointSource, SourceCollection, YeeGrid
    from .core.fdtd2d import SimulationDivergedError

    dev = torch.device(device) if device is not None else torch.device("cuda")
    v = v_medium if v_medium is not None else C0 / math.sqrt(40.0)
    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9)
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    last = 0.0
    for k, req in enumerate(pending):
        if progress is not None and now() - last >= progress_every:
            last = now()
            progress(k, last)
        if realtime:
            wait = req.arrival - now()
            if wait > 0:
                time.sleep(wait)
        req.admitted = now()
        nx, ny = req.eps.shape
        grid = YeeGrid(nx, ny, dx, dx, courant=0.99, dt=dt, device=dev)
        ca, cb = coefficients_from_maps(req.eps, req.sigma, dt, device=dev)
        n_ant = len(req.antennas)
        req.traces = torch.empty((n_ant, n_ant, req.n_steps), dtype=torch.float32, device=dev)
        req.failed = False
        for t, (i, j) in enumerate(req.antennas):
            f = FieldSet(grid)
            det = DetectorSet(req.antennas, req.n_steps, grid=grid)
            src = PointSource(wave, i, j, "Hz", grid=grid, N_steps=req.n_steps)
            sim = FDTD2D(grid, f, MurABC(grid, f.Hz), SourceCollection([src]), Ca=ca, Cb=cb, detectors=det,
                         n_check=n_check, stability_threshold=stability_threshold, division=division)
            try:
                sim.run(req.n_steps)
            except SimulationDivergedError:
                req.failed = True
            req.traces[t].copy_(det.traces)
        ax = torch.tensor([a[0] * dx for a in req.antennas], dtype=torch.float64, device=dev)
        ay = torch.tensor([a[1] * dx for a in req.antennas], dtype=torch.float64, device=dev)
        req.image = torch.empty((nx, ny), dtype=torch.float64, device=dev)
        run_das(req.traces, ax, ay, req.image, pitch_x=dx, pitch_y=dx, v_medium=v, dt=dt, clip=True)
        torch.cuda.current_stream(dev).synchronize()
        req.finished = now()
    return _stats(pending, now())


def poisson_requests(n: int, rate_hz: float, *, n_min: int = 128, n_max: int = 1024, elements=(8, 16, 32),
                     seed: int = 0, steps_per_diag: float = 8.0, dx: float = 1e-3) -> list[ImagingRequest]:
    """The serving stream: Poisson arrivals at ``rate_hz``; grid size
    log-uniform in ``[n_min, n_max]``, E from ``elements``, steps =
    ``steps_per_diag`` x grid diagonal in cells; head phantoms with a clot."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate_hz)
        n_cells = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        e = int(rng.choice(elements))
        out.append(make_request(rid, t, n_cells, e, seed=seed * 100003 + rid, steps_per_diag=steps_per_diag, dx=dx))
    return out


def make_request(rid: int, arrival: float, n_cells: int, elements: int, *, seed: int,
                 steps_per_diag: float = 8.0, dx: float = 1e-3) -> ImagingRequest:
    """One request of the serving stream: a head phantom with a clot on an
    ``n_cells``^2 grid, an ``elements``-element array, steps =
    ``steps_per_diag`` x grid diagonal in cells."""
    from .workloads import head_phantom

    ph = head_phantom(n_cells, elements, seed=seed, clot=True, dx=dx)
    steps = int(math.floor(steps_per_diag * n_cells * math.sqrt(2.0)))
    return ImagingRequest(rid, arrival, ph.eps, ph.sigma, ph.antennas, steps)
