"""Gradients of a loss on recorded traces with respect to per-cell eps_r and sigma.

``TraceProblem`` describes a multi-TX experiment on one phantom: a grid, TX
cells (each a soft Hz point source driven by the same waveform), RX cells
(Hz recorded after every step) and a step count. ``traces(eps_r, sigma)`` is
a ``torch.autograd.Function``: forward runs every TX through the temporally
blocked HIP stepper in one batched launch sequence and keeps the state of
every run every ``checkpoint`` steps; backward walks the segments from last
to first, recomputes each one from its checkpoint (one depth-1 launch per
step writing every intermediate state) and runs the hand-written discrete
adjoint kernel (``wf_adjoint_step``) back through it, accumulating
dL/dCa and dL/dCb; the chain rule to (eps_r, sigma) goes through the fp64
coefficient formula of ``Material.ca/cb`` (``src/core/materials.py:61-82``).
The reference only sketches this (``docs/architecture/DIFFERENTIABLE_PHYSICS.md:33``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native
from ._native import WfAdjArgs, WfRun2d, WfTbArgs
from .engine import EPS0, CellSource, RunState, SourceTable, StepCoefficients, tile_count
from .mimo import BatchedRuns, RunSpec


def coefficients(eps: torch.Tensor, sigma: torch.Tensor, dt: float) -> tuple[torch.Tensor, torch.Tensor]:
    """Differentiable fp64 ``(Ca, Cb)`` from eps_r and sigma (``Material.ca/cb``)."""
    alpha = sigma * dt / (2.0 * EPS0 * eps)
    ca = (1.0 - alpha) / (1.0 + alpha)
    cb = (dt / (EPS0 * eps)) / (1.0 + alpha)
    return ca, cb


@dataclass
class FullHistory:
    """Every step's state of a forward: ``states[k]`` is u_k, ``(n_tx, 3, nx, ny)``."""

    states: torch.Tensor


@dataclass
class TraceProblem:
    nx: int
    ny: int
    dx: float
    dt: float
    tx: Sequence[tuple[int, int]]
    rx: Sequence[tuple[int, int]]
    wave: torch.Tensor          # fp32 on the device, >= n_steps samples
    n_steps: int
    checkpoint: int = 100
    depth: object = 6           # forward launch depth (or "auto": the autotuner's)
    tile_rows: object = None    # forward tile rows (None: FDTD2D's default; "auto": the autotuner's)
    receiver_map: bool = True   # False: one record launch per adjoint step (atomic adds)
    two_pass: bool = False      # True: two launches per adjoint step (same results)
    one_step: bool = False      # True: one launch per adjoint step, never two steps a launch (same results)
    fast_div: bool = True       # two-step launches divide by dx, dy with the device-verified fast quotient
    history: str = "auto"       # "full": the forward keeps every step's state and backward runs straight
                                # through it; "checkpoint": every `checkpoint`-th state, each segment
                                # recomputed in backward; "auto": full when it fits in history_budget
    history_budget: Optional[float] = None  # bytes for a full history (None: half the free device memory)

    def __post_init__(self):
        self.device = self.wave.device
        if self.device.type != "cuda":
            raise RuntimeError("waveforge_amd: the adjoint runs on the ROCm GPU; there is no CPU path")
        _native.load()
        self.coef = StepCoefficients.from_grid(self.dt, self.dx, self.dx)
        cells = [i * self.ny + j for i, j in self.rx]
        self.rx_cells = torch.tensor(cells, dtype=torch.int32, device=self.device)
        # receiver map: receivers ordered by cell (index order within a cell) and,
        # per cell, the position of its first receiver in that order (-1: none)
        order = sorted(range(len(cells)), key=lambda k: (cells[k], k))
        first = torch.full((self.nx * self.ny,), -1, dtype=torch.int32)
        for p in range(len(order) - 1, -1, -1):
            first[cells[order[p]]] = p
        self.rx_order = torch.tensor(order, dtype=torch.int32, device=self.device)
        self.rx_first = first.to(self.device)

    # ------------------------------------------------------------------
    def history_bytes(self) -> int:
        """Device bytes of the full state history (every step's Ex, Ey, Hz of every TX)."""
        return (self.n_steps + 1) * len(self.tx) * 3 * self.nx * self.ny * 4

    def full_history(self) -> bool:
        """Whether ``forward`` keeps the full state history (``history`` mode)."""
        if self.history in ("full", "checkpoint"):
            return self.history == "full"
        if self.history != "auto":
            raise ValueError(f"history must be 'auto', 'full' or 'checkpoint', got {self.history!r}")
        budget = self.history_budget
        if budget is None:
            budget = 0.5 * torch.cuda.mem_get_info(self.device)[0]
        return self.history_bytes() <= budget

    def _specs(self, ca, cb):
        return [RunSpec(self.nx, self.ny, ca, cb, sources=[("Hz", i, j, self.wave, 1.0)], receivers=list(self.rx),
                        n_record=self.n_steps) for (i, j) in self.tx]

    def forward(self, ca: torch.Tensor, cb: torch.Tensor):
        """Traces ``(n_tx, n_rx, n_steps)`` and what backward needs: with the
        full history, ``FullHistory`` (every step's state, written by the same
        launches that record the traces); otherwise the list of checkpoints
        ``(n_tx, 3, nx, ny)`` taken at steps 0, K, 2K, ..."""
        if self.full_history():
            return self._forward_full(ca, cb)
        specs = self._specs(ca, cb)
        # IEEE division: the adjoint kernels transpose the step that divides by dx, dy
        br = BatchedRuns(specs, dx=self.dx, dy=self.dx, dt=self.dt, device=self.device, depth=self.depth,
                         tile_rows=self.tile_rows, n_check=10 ** 9, division="ieee")
        ckpts = []
        done = 0
        K = self.checkpoint
        while done < self.n_steps:
            ckpts.append(torch.stack([torch.stack(br.fields(k)) for k in range(len(specs))]))
            seg = min(K, self.n_steps - done)
            if seg >= 2:
                br.advance(seg)
            else:  # a single trailing step: the per-step kernel path (in place)
                from .engine import per_step_advance

                for r in br.runs:
                    r.step = done
                    if r.pstrip is None:
                        r.pstrip = torch.empty(4 * (self.nx + self.ny), dtype=torch.float32, device=self.device)
                per_step_advance(br.runs, 1, _native.stream_handle(self.device))
                br.steps_completed += 1
            done += seg
        traces = torch.stack([br.traces(k) for k in range(len(specs))])
        self._tables = [r.sources for r in br.runs]
        return traces, ckpts

    def _forward_full(self, ca: torch.Tensor, cb: torch.Tensor):
        """The forward as one pass of state-history launches from the zero
        state over all n_steps: slot k of the history holds u_k, and the
        launches record the receivers (the stepper's arithmetic, so traces and
        states equal the checkpointed forward's bit for bit)."""
        specs = self._specs(ca, cb)
        # the runs' source tables, receiver cells and trace buffer (their state is not stepped)
        br = BatchedRuns(specs, dx=self.dx, dy=self.dx, dt=self.dt, device=self.device, depth=1, tile_rows=48,
                         n_check=10 ** 9, division="ieee")
        self._tables = [r.sources for r in br.runs]
        hist = self._history(ca, cb, self.n_steps, recorders=br.runs)
        self._recompute(hist, None, 0, self.n_steps)
        traces = torch.stack([br.traces(k) for k in range(len(specs))])
        return traces, FullHistory(hist[0])

    def forward_fp64(self, ca64: torch.Tensor, cb64: torch.Tensor) -> torch.Tensor:
        """fp64 traces ``(n_tx, n_rx, n_steps)`` for fp64 coefficients (the
        double-precision step, no checkpoints): the forward of central-difference
        gradient checks, whose quotients fp32 rounding would swamp."""
        from .fp64 import Fp64Runs

        specs = [RunSpec(self.nx, self.ny, ca64, cb64, sources=[("Hz", i, j, self.wave, 1.0)],
                         receivers=list(self.rx), n_record=self.n_steps) for (i, j) in self.tx]
        runs = Fp64Runs(specs, dx=self.dx, dy=self.dx, dt=self.dt, device=self.device)
        runs.advance(self.n_steps)
        return torch.stack([runs.traces(k) for k in range(len(specs))])

    def _divisions(self):
        """The verified fast divisions by dx and dy (``engine.verified_division``),
        or None when either divisor has none."""
        from .engine import verified_division

        vx = verified_division(self.coef.dx, self.device)
        vy = verified_division(self.coef.dy, self.device)
        return None if vx is None or vy is None else (vx, vy)

    # ------------------------------------------------------------------
    def _history(self, ca, cb, K: Optional[int] = None, recorders=None):
        """State history buffer ``H[k] = u_{s0+k}`` (k = 0..K, K the checkpoint
        interval by default) and the run table whose entry k makes a launch
        read H[k] and write H[k+1], H[k+2], ... one slot per step (descriptor
        step = k; the launch's step_offset adds s0). With ``recorders`` (the
        forward's RunStates) the descriptors also record their receivers.
        Descriptors differ only in their state pointers and step."""
        n_tx, nx, ny = len(self.tx), self.nx, self.ny
        K = self.checkpoint if K is None else K
        H = torch.empty((K + 1, n_tx, 3, nx, ny), dtype=torch.float32, device=self.device)
        base = []
        for r in range(n_tx):
            rec = recorders[r] if recorders is not None else None
            rs = RunState(H[0, r, 0], H[0, r, 1], H[0, r, 2], ca, cb, nx, ny, self.coef,
                          sources=self._tables[r], step=0,
                          det_cell=None if rec is None else rec.det_cell,
                          traces=None if rec is None else rec.traces, trace_n0=0)
            rs.second = (H[1, r, 0], H[1, r, 1], H[1, r, 2])
            base.append(rs.struct())
        sz = ctypes.sizeof(WfRun2d)
        tbl = np.tile(np.frombuffer(bytes((WfRun2d * n_tx)(*base)), dtype=np.uint8).reshape(1, n_tx, sz), (K, 1, 1))
        ptr = tbl.view(np.uint64)  # (K, n_tx, sz/8)
        cell_bytes = nx * ny * 4
        k = np.arange(K, dtype=np.uint64)[:, None]
        rr = np.arange(n_tx, dtype=np.uint64)[None, :]
        h0 = np.uint64(H.data_ptr())
        for c, (f_lo, f_hi) in enumerate((("ex", "ex2"), ("ey", "ey2"), ("hz", "hz2"))):
            src = h0 + ((k * np.uint64(n_tx) + rr) * np.uint64(3) + np.uint64(c)) * np.uint64(cell_bytes)
            ptr[:, :, getattr(WfRun2d, f_lo).offset // 8] = src
            ptr[:, :, getattr(WfRun2d, f_hi).offset // 8] = src + np.uint64(3 * n_tx * cell_bytes)
        tbl.view(np.int32)[:, :, WfRun2d.step.offset // 4] = np.arange(K, dtype=np.int32)[:, None]
        table = torch.from_numpy(np.ascontiguousarray(tbl).reshape(-1)).to(self.device)
        return H, table

    # the state-history launches exist for 32- and 48-row tiles; 32 rows measured
    # 49 us a launch against 57 us for 48 on 16 runs of 256^2 (768 workgroups,
    # three a CU, against 576), the history stores bounding both
    recompute_tile_rows = 32
    recompute_depth = 8       # steps per recompute launch (each step's state is written)

    def _tile_list(self, depth: int) -> torch.Tensor:
        """(run, tile) pairs of every run for ``depth``-step launches (cached)."""
        cache = self.__dict__.setdefault("_tiles_by_depth", {})
        if depth not in cache:
            n_tx = len(self.tx)
            nt = tile_count(self.nx, self.ny, self.recompute_tile_rows, depth)
            tiles = np.stack([np.repeat(np.arange(n_tx, dtype=np.int32), nt),
                              np.tile(np.arange(nt, dtype=np.int32), n_tx)], axis=1)
            cache[depth] = torch.from_numpy(np.ascontiguousarray(tiles)).to(self.device)
        return cache[depth]

    def _recompute(self, hist, start: torch.Tensor, s0: int, L: int) -> torch.Tensor:
        """Fill ``H[0..L]`` with ``u_{s0}..u_{s0+L}`` from checkpoint ``start``:
        temporally blocked launches of up to ``recompute_depth`` steps, each
        writing the state after every one of its steps into its slot."""
        H, table = hist
        n_tx, nx, ny = len(self.tx), self.nx, self.ny
        if start is None:
            H[0].zero_()
        else:
            H[0].copy_(start)
        lib = _native.load()
        stream = _native.stream_handle(self.device)
        sz = ctypes.sizeof(WfRun2d)
        a = WfTbArgs()
        a.n_runs, a.tile_i = n_tx, self.recompute_tile_rows
        a.parity, a.step_offset, a.has_materials = 0, s0, 1
        a.hist_stride = n_tx * 3 * nx * ny  # floats from one slot to the next
        k = 0
        while k < L:
            d = min(self.recompute_depth, L - k)
            tiles_t = self._tile_list(d)
            a.tiles, a.n_tiles = tiles_t.data_ptr(), int(tiles_t.shape[0])
            a.depth, a.tile_j = d, _native.tile_cols(d)
            a.runs = table.data_ptr() + k * n_tx * sz
            _native.check(lib.wf_fdtd2d_tb_advance(ctypes.byref(a), ctypes.c_void_p(stream)), "recompute")
            k += d
        return H

    def backward(self, ca: torch.Tensor, cb: torch.Tensor, ckpts, dl_dd: torch.Tensor):
        """fp64 ``(dL/dCa, dL/dCb)`` summed over TX, given dL/d traces."""
        n_tx, nx, ny = len(self.tx), self.nx, self.ny
        dev = self.device
        dl_dd = dl_dd.to(device=dev, dtype=torch.float32).contiguous()
        z = lambda dt=torch.float32: torch.zeros((n_tx, nx, ny), dtype=dt, device=dev)  # noqa: E731
        lex, ley, lhz, lex2, ley2, lhz2, lp = (z() for _ in range(7))
        gca, gcb = z(torch.float64), z(torch.float64)
        lib = _native.load()
        stream = _native.stream_handle(dev)
        K = self.checkpoint
        a = WfAdjArgs()
        a.ca, a.cb = ca.data_ptr(), cb.data_ptr()
        a.lex, a.ley, a.lhz = lex.data_ptr(), ley.data_ptr(), lhz.data_ptr()
        a.lex2, a.ley2, a.lhz2, a.lp = lex2.data_ptr(), ley2.data_ptr(), lhz2.data_ptr(), lp.data_ptr()
        a.gca, a.gcb = gca.data_ptr(), gcb.data_ptr()
        a.dl_dd, a.rx_cell = dl_dd.data_ptr(), self.rx_cells.data_ptr()
        if len(self.rx) and self.receiver_map:
            a.rx_first, a.rx_order = self.rx_first.data_ptr(), self.rx_order.data_ptr()
        a.flags = (_native.WF_ADJ_TWO_PASS if self.two_pass else 0) | (_native.WF_ADJ_ONE_STEP if self.one_step else 0)
        div = self._divisions() if self.fast_div else None
        if div is not None:
            (a.div_rx, _, a.div_span_x), (a.div_ry, _, a.div_span_y) = div
            a.flags |= _native.WF_ADJ_FAST_DIV
        a.run_stride = nx * ny
        a.state_stride = 3 * nx * ny
        a.dd_run_stride, a.dd_rx_stride = dl_dd.stride(0), dl_dd.stride(1)
        a.nx, a.ny, a.n_runs, a.n_rx = nx, ny, n_tx, len(self.rx)
        c = self.coef
        a.dh, a.dx, a.dy, a.cmx, a.cmy = c.dh, c.dx, c.dy, c.cmx, c.cmy
        step_bytes = n_tx * 3 * nx * ny * 4
        a.hist_stride = step_bytes // 4

        def adjoint_segment(h0: int, s0: int, L: int) -> None:
            # L adjoint steps in one call, from step s0+L-1 back to s0, over a
            # history whose slot 0 (at address h0) holds u_{s0}
            k = L - 1
            a.step, a.n_steps = s0 + k, L
            a.ex_prev = h0 + k * step_bytes
            a.ey_prev = h0 + k * step_bytes + nx * ny * 4
            a.hz_next = h0 + (k + 1) * step_bytes + 2 * nx * ny * 4
            _native.check(lib.wf_adjoint_step(ctypes.byref(a), ctypes.c_void_p(stream)), "wf_adjoint_step")

        if isinstance(ckpts, FullHistory):
            adjoint_segment(ckpts.states.data_ptr(), 0, self.n_steps)
            return gca.sum(0), gcb.sum(0)
        hist = self._history(ca, cb)
        H = hist[0]
        for seg in range(len(ckpts) - 1, -1, -1):
            s0 = seg * K
            L = min(K, self.n_steps - s0)
            self._recompute(hist, ckpts[seg], s0, L)
            adjoint_segment(H.data_ptr(), s0, L)
        del hist, H
        return gca.sum(0), gcb.sum(0)


class _Traces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, problem, fp64):
        with torch.no_grad():
            ca64, cb64 = coefficients(eps.double(), sigma.double(), problem.dt)
            ca, cb = ca64.float().contiguous(), cb64.float().contiguous()
            traces, ckpts = problem.forward(ca, cb)
            if fp64:
                traces = problem.forward_fp64(ca64, cb64)
        ctx.problem, ctx.ca, ctx.cb, ctx.ckpts = problem, ca, cb, ckpts
        ctx.save_for_backward(eps, sigma)
        return traces

    @staticmethod
    def backward(ctx, grad):
        eps, sigma = ctx.saved_tensors
        gca, gcb = ctx.problem.backward(ctx.ca, ctx.cb, ctx.ckpts, grad)
        with torch.enable_grad():
            e = eps.detach().double().requires_grad_(True)
            s = sigma.detach().double().requires_grad_(True)
            ca64, cb64 = coefficients(e, s, ctx.problem.dt)
            ge, gs = torch.autograd.grad([ca64, cb64], [e, s], [gca, gcb])
        return ge.to(eps.dtype), gs.to(sigma.dtype), None, None


def traces(eps: torch.Tensor, sigma: torch.Tensor, problem: TraceProblem, fp64: bool = False) -> torch.Tensor:
    """Differentiable ``(n_tx, n_rx, n_steps)`` Hz traces of ``problem`` for
    per-cell ``eps`` and ``sigma`` maps ``(nx, ny)`` on the device.

    With ``fp64`` the returned traces come from the double-precision GPU step
    (fp64 coefficients), so a misfit whose residual is much smaller than the
    traces themselves is not dominated by fp32 rounding; checkpoints and the
    adjoint still run in fp32 (their rounding enters the gradient only
    relative to the states, not to the residual)."""
    return _Traces.apply(eps, sigma, problem, fp64)


def misfit_gradient(eps, sigma, problem: TraceProblem, d_obs: torch.Tensor, fp64_residual: bool = False):
    """``L = 0.5 ||d_sim - d_obs||^2`` (fp64) and its gradient w.r.t. eps, sigma."""
    e = eps.detach().clone().requires_grad_(True)
    s = sigma.detach().clone().requires_grad_(True)
    d = traces(e, s, problem, fp64_residual)
    loss = 0.5 * ((d.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    return float(loss.detach()), e.grad, s.grad
