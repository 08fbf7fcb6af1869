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
    depth: int = 6

    def __post_init__(self):
        self.device = self.wave.device
        if self.device.type != "cuda":
            raise RuntimeError("waveforge_amd: the adjoint runs on the ROCm GPU; there is no CPU path")
        _native.load()
        self.coef = StepCoefficients.from_grid(self.dt, self.dx, self.dx)
        self.rx_cells = torch.tensor([i * self.ny + j for i, j in self.rx], dtype=torch.int32, device=self.device)

    # ------------------------------------------------------------------
    def forward(self, ca: torch.Tensor, cb: torch.Tensor):
        """Traces ``(n_tx, n_rx, n_steps)`` and the list of checkpoints
        ``(n_tx, 3, nx, ny)`` taken at steps 0, K, 2K, ..."""
        specs = [RunSpec(self.nx, self.ny, ca, cb, sources=[("Hz", i, j, self.wave, 1.0)], receivers=list(self.rx),
                         n_record=self.n_steps) for (i, j) in self.tx]
        br = BatchedRuns(specs, dx=self.dx, dy=self.dx, dt=self.dt, device=self.device, depth=self.depth,
                         n_check=10 ** 9)
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

    # ------------------------------------------------------------------
    def _recompute(self, ca, cb, start: torch.Tensor, s0: int, L: int) -> torch.Tensor:
        """States ``H[k] = u_{s0+k}``, k = 0..L, from checkpoint ``start``: one
        depth-1 launch per step, reading H[k] and writing H[k+1]."""
        n_tx, nx, ny = len(self.tx), self.nx, self.ny
        H = torch.empty((L + 1, n_tx, 3, nx, ny), dtype=torch.float32, device=self.device)
        H[0].copy_(start)
        structs = []
        for k in range(L):
            for r in range(n_tx):
                rs = RunState(H[k, r, 0], H[k, r, 1], H[k, r, 2], ca, cb, nx, ny, self.coef,
                              sources=self._tables[r], step=s0 + k)
                rs.second = (H[k + 1, r, 0], H[k + 1, r, 1], H[k + 1, r, 2])
                structs.append(rs.struct())
        arr = (WfRun2d * len(structs))(*structs)
        table = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(self.device)
        tile_i = 32
        tiles = np.concatenate([np.stack([np.full(tile_count(nx, ny, tile_i, 1), r, dtype=np.int32),
                                          np.arange(tile_count(nx, ny, tile_i, 1), dtype=np.int32)], axis=1)
                                for r in range(n_tx)])
        tiles_t = torch.from_numpy(np.ascontiguousarray(tiles)).to(self.device)
        lib = _native.load()
        stream = _native.stream_handle(self.device)
        sz = ctypes.sizeof(WfRun2d)
        for k in range(L):
            a = WfTbArgs()
            a.runs = table.data_ptr() + k * n_tx * sz
            a.tiles = tiles_t.data_ptr()
            a.n_runs, a.n_tiles = n_tx, int(tiles.shape[0])
            a.tile_i, a.tile_j = tile_i, _native.tile_cols(1)
            a.depth, a.parity, a.step_offset, a.has_materials = 1, 0, 0, 1
            _native.check(lib.wf_fdtd2d_tb_advance(ctypes.byref(a), ctypes.c_void_p(stream)), "recompute")
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
        a.run_stride = nx * ny
        a.state_stride = 3 * nx * ny
        a.dd_run_stride, a.dd_rx_stride = dl_dd.stride(0), dl_dd.stride(1)
        a.nx, a.ny, a.n_runs, a.n_rx = nx, ny, n_tx, len(self.rx)
        c = self.coef
        a.dh, a.dx, a.dy, a.cmx, a.cmy = c.dh, c.dx, c.dy, c.cmx, c.cmy
        for seg in range(len(ckpts) - 1, -1, -1):
            s0 = seg * K
            L = min(K, self.n_steps - s0)
            H = self._recompute(ca, cb, ckpts[seg], s0, L)
            for k in range(L - 1, -1, -1):
                a.step = s0 + k
                a.ex_prev = H[k, :, 0].data_ptr()
                a.ey_prev = H[k, :, 1].data_ptr()
                a.hz_next = H[k + 1, :, 2].data_ptr()
                _native.check(lib.wf_adjoint_step(ctypes.byref(a), ctypes.c_void_p(stream)), "wf_adjoint_step")
            del H
        return gca.sum(0), gcb.sum(0)


class _Traces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, problem):
        with torch.no_grad():
            ca64, cb64 = coefficients(eps.double(), sigma.double(), problem.dt)
            ca, cb = ca64.float().contiguous(), cb64.float().contiguous()
            traces, ckpts = problem.forward(ca, cb)
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
        return ge.to(eps.dtype), gs.to(sigma.dtype), None


def traces(eps: torch.Tensor, sigma: torch.Tensor, problem: TraceProblem) -> torch.Tensor:
    """Differentiable ``(n_tx, n_rx, n_steps)`` Hz traces of ``problem`` for
    per-cell ``eps`` and ``sigma`` maps ``(nx, ny)`` on the device."""
    return _Traces.apply(eps, sigma, problem)


def misfit_gradient(eps, sigma, problem: TraceProblem, d_obs: torch.Tensor):
    """``L = 0.5 ||d_sim - d_obs||^2`` (fp64) and its gradient w.r.t. eps, sigma."""
    e = eps.detach().clone().requires_grad_(True)
    s = sigma.detach().clone().requires_grad_(True)
    d = traces(e, s, problem)
    loss = 0.5 * ((d.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    return float(loss.detach()), e.grad, s.grad
