"""Double-precision runs on the GPU: the same leapfrog step in fp64.

The reference runs in float64 when its grid dtype is float64
(``src/core/fields.py:156-168`` allocates the fields in the grid dtype; the
update coefficients are Python doubles, ``src/core/fdtd2d.py:158-159``, and
so are the Mur coefficients, ``src/core/boundaries.py:177-178``). This
module drives ``wf_fdtd2d_step_f64`` for a set of runs that may differ in
grid size: three launches per step cover every run. It is the FP64
reference run against which fp32 results state their error at full size,
and the forward of central-difference gradient checks, where fp32 rounding
would swamp the difference quotient.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np
import torch

from . import _native
from ._native import WfF64Args, WfRun2dF64
from .engine import C0, EPS0, MU0

INJ64_DTYPE = np.dtype([("wave", np.uint64), ("wave_len", np.int32), ("comp", np.int32), ("cell", np.int32),
                        ("pad", np.int32), ("amp", np.float64)])
assert INJ64_DTYPE.itemsize == ctypes.sizeof(_native.WfInj2dF64)


class Fp64Runs:
    """fp64 state, sources and receivers of independent runs.

    ``specs`` are ``mimo.RunSpec``-like objects: ``nx, ny, ca, cb`` (any
    float tensors or None), ``sources`` as ``(comp, i, j, wave, amp)`` tuples
    (wave: a device tensor, used in fp64), ``receivers`` as ``(i, j)`` Hz
    cells and ``n_record``. Contributions to one (component, cell) are applied
    in list order, as ``SourceCollection`` does.
    """

    def __init__(self, specs: Sequence, *, dx: float, dy: float, dt: float, device=None) -> None:
        if not specs:
            raise ValueError("no runs")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("waveforge_amd: fp64 runs need the ROCm GPU; there is no CPU path")
        _native.load()
        self.dx, self.dy, self.dt = float(dx), float(dy), float(dt)
        cmx = (C0 * dt - dx) / (C0 * dt + dx)
        cmy = (C0 * dt - dy) / (C0 * dt + dy)
        dev = self.device
        self._keep = []
        self.state, self.trace_bufs = [], []
        structs = []
        self.min_wave_len = math.inf
        for k, sp in enumerate(specs):
            nx, ny = int(sp.nx), int(sp.ny)
            if nx < 4 or ny < 4:
                raise ValueError(f"run {k}: grid must be at least 4x4")
            ex, ey, hz = (torch.zeros((nx, ny), dtype=torch.float64, device=dev) for _ in range(3))
            pstrip = torch.zeros(4 * (nx + ny), dtype=torch.float64, device=dev)
            r = WfRun2dF64()
            r.ex, r.ey, r.hz, r.pstrip = ex.data_ptr(), ey.data_ptr(), hz.data_ptr(), pstrip.data_ptr()
            self._keep += [pstrip]
            if (sp.ca is None) != (sp.cb is None):
                raise ValueError(f"run {k}: ca and cb must both be given or both None")
            if sp.ca is not None:
                ca = sp.ca.to(device=dev, dtype=torch.float64).reshape(nx, ny).contiguous()
                cb = sp.cb.to(device=dev, dtype=torch.float64).reshape(nx, ny).contiguous()
                self._keep += [ca, cb]
                r.ca, r.cb = ca.data_ptr(), cb.data_ptr()
            self._source_table(r, sp.sources, nx, ny, k)
            r.trace_n0 = int(getattr(sp, "trace_n0", 0))
            if sp.receivers:
                if any(not (0 <= i < nx and 0 <= j < ny) for i, j in sp.receivers):
                    raise ValueError(f"run {k}: receiver outside the grid")
                cells = torch.tensor([i * ny + j for i, j in sp.receivers], dtype=torch.int32, device=dev)
                tr = torch.zeros((len(sp.receivers), max(1, int(sp.n_record))), dtype=torch.float64, device=dev)
                self._keep.append(cells)
                r.det_cell, r.traces = cells.data_ptr(), tr.data_ptr()
                r.trace_stride, r.n_det, r.trace_len = tr.stride(0), len(sp.receivers), tr.shape[1]
                self.trace_bufs.append(tr)
            else:
                self.trace_bufs.append(None)
            r.nx, r.ny, r.step = nx, ny, 0
            r.dh, r.de, r.dx, r.dy, r.cmx, r.cmy = dt / MU0, dt / EPS0, self.dx, self.dy, cmx, cmy
            structs.append(r)
            self.state.append((ex, ey, hz))
        arr = (WfRun2dF64 * len(structs))(*structs)
        self._table = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(dev)
        self.max_nx = max(int(s.nx) for s in specs)
        self.max_ny = max(int(s.ny) for s in specs)
        self.steps_completed = 0

    def _source_table(self, r: WfRun2dF64, sources, nx: int, ny: int, k: int) -> None:
        if not sources:
            return
        rec = np.zeros(len(sources), dtype=INJ64_DTYPE)
        key = np.zeros(len(sources), dtype=np.int64)
        for m, (comp, i, j, wave, amp) in enumerate(sources):
            if comp not in _native.WF_COMP:
                raise ValueError(f"run {k}: component {comp!r} is not a 2D TM field")
            if not (0 <= i < nx and 0 <= j < ny):
                raise ValueError(f"run {k}: source cell ({i}, {j}) outside the grid")
            w = wave.to(device=self.device, dtype=torch.float64).contiguous()
            self._keep.append(w)
            self.min_wave_len = min(self.min_wave_len, w.numel())
            rec[m] = (w.data_ptr(), w.numel(), _native.WF_COMP[comp], i * ny + j, 0, float(amp))
            key[m] = _native.WF_COMP[comp] * nx * ny + i * ny + j
        perm = np.argsort(key, kind="stable")
        rec, key = rec[perm], key[perm]
        starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
        grp = torch.from_numpy(np.r_[starts, len(rec)].astype(np.int32)).to(self.device)
        inj = torch.from_numpy(rec.view(np.uint8).copy()).to(self.device)
        self._keep += [grp, inj]
        r.inj, r.grp_start, r.n_grp, r.n_inj = inj.data_ptr(), grp.data_ptr(), len(starts), len(rec)

    def advance(self, n_steps: int) -> None:
        """Enqueue ``n_steps`` fp64 steps of every run (asynchronous)."""
        if n_steps <= 0:
            return
        if self.steps_completed + n_steps > self.min_wave_len:
            raise RuntimeError(f"a source waveform has {self.min_wave_len} samples, "
                               f"cannot reach step {self.steps_completed + n_steps}")
        lib = _native.load()
        a = WfF64Args()
        a.runs, a.n_runs = self._table.data_ptr(), len(self.state)
        a.max_nx, a.max_ny = self.max_nx, self.max_ny
        a.n_steps, a.step_offset = int(n_steps), self.steps_completed
        _native.check(lib.wf_fdtd2d_step_f64(ctypes.byref(a), ctypes.c_void_p(_native.stream_handle(self.device))),
                      "wf_fdtd2d_step_f64")
        self.steps_completed += n_steps

    def fields(self, k: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        return self.state[k]

    def traces(self, k: int) -> torch.Tensor:
        return self.trace_bufs[k]


class SourceTable64:
    """fp64 device records (``wf_inj2d_f64``) of a list of reference-style
    source objects, grouped by (component, cell) with collection order kept
    inside a group, as ``engine.SourceTable`` does for fp32.

    Each contribution adds ``wave[n] * amp`` in double, which is the
    reference's arithmetic on a float64 grid (tables and amplitudes in the
    grid dtype, ``src/core/sources.py:826-833, 971-981``; ``values = w_now *
    amp`` then ``target[idx] += values``, :882-887). A source whose table or
    amplitudes are in another dtype gets a per-cell table of its own products
    ``w[n] * amp`` formed in that dtype and widened, with amp 1, so the
    rounding is the reference's too. Plane sources add ``w[n]`` (amp 1,
    :715-736). Components outside Ex/Ey/Hz are kept in ``inert``."""

    def __init__(self, sources, Nx: int, Ny: int, device) -> None:
        from .core.sources import PlaneSource

        self.inert, self._keep = [], []
        self.min_wave_len = None
        recs, keys = [], []
        for src in sources:
            comp = src.component
            wave = src._waveform_tensor
            self.min_wave_len = wave.numel() if self.min_wave_len is None else min(self.min_wave_len, wave.numel())
            if comp not in _native.WF_COMP:
                self.inert.append(src)
                continue
            if isinstance(src, PlaneSource):
                i, j, _, _ = src.cells3d(Nx, Ny, 1)
                amp = None
            else:
                i, j = src._i_idx.cpu(), src._j_idx.cpu()
                amp = src._amp.detach()
            w = wave.detach().to(device)
            if amp is None or (w.dtype == torch.float64 and amp.dtype == torch.float64):
                w64 = w.to(torch.float64).contiguous()
                self._keep.append(w64)
                a64 = [1.0] * i.numel() if amp is None else amp.cpu().tolist()
                tables = [(w64, a) for a in a64]
            else:  # products in the source's own dtype, then widened
                amp_d = amp.to(device)
                tables = []
                for c in range(i.numel()):
                    t = (w * amp_d[c:c + 1]).to(torch.float64).contiguous()
                    self._keep.append(t)
                    tables.append((t, 1.0))
            for c in range(i.numel()):
                cell = int(i[c]) * Ny + int(j[c])
                t, a = tables[c]
                recs.append((t.data_ptr(), t.numel(), _native.WF_COMP[comp], cell, 0, float(a)))
                keys.append(_native.WF_COMP[comp] * Nx * Ny + cell)
        self.n_grp = self.n_inj = 0
        self.inj = self.grp_start = None
        if recs:
            rec = np.array(recs, dtype=INJ64_DTYPE)
            key = np.array(keys, dtype=np.int64)
            perm = np.argsort(key, kind="stable")
            rec, key = rec[perm], key[perm]
            starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
            self.grp_start = torch.from_numpy(np.r_[starts, len(rec)].astype(np.int32)).to(device)
            self.inj = torch.from_numpy(rec.view(np.uint8).copy()).to(device)
            self.n_grp, self.n_inj = int(starts.size), int(len(rec))


class _TwinSpec:
    def __init__(self, nx, ny, ca, cb, sources, receivers, n_record, trace_n0):
        self.nx, self.ny, self.ca, self.cb = nx, ny, ca, cb
        self.sources, self.receivers, self.n_record, self.trace_n0 = sources, receivers, n_record, trace_n0


def fp64_twin(sim) -> Fp64Runs:
    """The same simulation as an fp32 ``core.FDTD2D`` (one batch member),
    set up as the reference sets it up on a float64 grid: waveform tables
    rebuilt in fp64 (``Waveform.build`` with the grid dtype,
    ``src/core/sources.py:112-138``), Ca/Cb as stored (the reference computes
    them in fp32 and converts, ``src/core/materials.py:218-250``), fp64 update
    and Mur coefficients, the same sources, cells and detectors, zero initial
    fields and step 0. Advance it with ``advance`` and compare."""
    from .core.sources import PlaneSource

    g = sim.grid
    nx, ny = g.Nx, g.Ny
    if sim.fields.Hz.dim() == 4:
        raise ValueError("fp64_twin: batched FieldSets are not supported")
    srcs = []
    for src in sim._source_list():
        comp = src.component
        if comp not in _native.WF_COMP:
            continue  # Ez, Hx, Hy never enter the 2D TM update
        wf = getattr(src, "_waveform", None)
        n = src._waveform_tensor.numel()
        wave = wf.build(n, g.dt, g.device, torch.float64) if wf is not None else src._waveform_tensor.double()
        if isinstance(src, PlaneSource):
            ii, jj, aa = src.cells(nx, ny)
        else:
            ii, jj, aa = src._i_idx.cpu(), src._j_idx.cpu(), src._amp.detach().double().cpu()
        for i, j, a in zip(ii.tolist(), jj.tolist(), aa.tolist()):
            srcs.append((comp, int(i), int(j), wave, float(a)))
    det = sim._detectors
    receivers, n_record, n0 = [], 0, 0
    if det is not None and len(det):
        receivers = [(int(c) // ny, int(c) % ny) for c in det.cell_index.cpu().tolist()]
        n_record, n0 = det.traces.shape[-1], det.start_step
    ca = None if sim._Ca is None else sim._Ca.reshape(nx, ny)
    cb = None if sim._Cb is None else sim._Cb.reshape(nx, ny)
    spec = _TwinSpec(nx, ny, ca, cb, srcs, receivers, n_record, n0)
    return Fp64Runs([spec], dx=g.dx, dy=g.dy, dt=g.dt, device=g.device)


def rel_l2(a: torch.Tensor, ref: torch.Tensor) -> float:
    """``||a - ref|| / ||ref||`` in fp64 (0 if both are zero)."""
    a, ref = a.double(), ref.double()
    den = float(torch.linalg.norm(ref))
    num = float(torch.linalg.norm(a - ref))
    return num / den if den > 0 else num
