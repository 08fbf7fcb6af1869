"""The allocation-free step (``wf_fdtd2d_step_noalloc``) against the step that
takes a caller workspace (``wf_fdtd2d_step``), through the C ABI, bit for bit.

Two calls in a row on one GPU share the library's scratch array, the second
with a smaller grid; a batch of two runs of different sizes in one call
reuses the array run after run. Grids whose Mur strips exceed the array are
refused before any launch.
"""

import ctypes

import pytest
import torch

from oracle.fdtd2d_ref import stable_dt

DX = 1e-3


def _run(device, nx, ny, seed, steps, with_strips):
    """A lossy run with random fields, an Hz source inside, an Ey source on the
    i = 0 face, an Hz source on the j = 0 face, and three Hz detectors."""
    from waveforge_amd.engine import CellSource, RunState, SourceTable, StepCoefficients

    dt = stable_dt(DX, DX)
    g = torch.Generator().manual_seed(seed)
    eps = 1.0 + 30.0 * torch.rand((nx, ny), generator=g, dtype=torch.float64)
    sig = torch.rand((nx, ny), generator=g, dtype=torch.float64)
    alpha = sig * dt / (2.0 * 8.8541878128e-12 * eps)
    ca = ((1.0 - alpha) / (1.0 + alpha)).float().to(device)
    cb = ((dt / (8.8541878128e-12 * eps)) / (1.0 + alpha)).float().to(device)
    f = [((torch.rand((nx, ny), generator=g) - 0.5) * s).to(device) for s in (1e-2, 1e-2, 1e-4)]
    t = torch.arange(steps + 4, dtype=torch.float32) * dt
    wave = torch.exp(-((t - 12 * dt) ** 2) / (2.0 * (4 * dt) ** 2)).to(device)
    srcs = [CellSource("Hz", wave, nx // 2, ny // 3), CellSource("Ey", wave, 0, ny // 2, 0.5),
            CellSource("Hz", wave, nx // 3, 0, -0.25)]
    rs = RunState(f[0], f[1], f[2], ca, cb, nx, ny, StepCoefficients.from_grid(dt, DX, DX),
                  sources=SourceTable(srcs, nx, ny, device))
    cells = [(1, 1), (nx // 2, ny // 3), (nx - 1, ny - 2)]
    rs.det_cell = torch.tensor([i * ny + j for i, j in cells], dtype=torch.int32, device=device)
    rs.traces = torch.zeros((len(cells), steps), dtype=torch.float32, device=device)
    if with_strips:
        rs.pstrip = torch.empty(4 * (nx + ny), dtype=torch.float32, device=device)
    return rs


def _step(entry, runs, steps, device):
    from waveforge_amd import _native

    lib = _native.load()
    arr = (_native.WfRun2d * len(runs))(*[r.struct() for r in runs])
    fn = getattr(lib, entry)
    _native.check(fn(arr, len(runs), steps, ctypes.c_void_p(_native.stream_handle(device))), entry)


def _same(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


@pytest.mark.gpu
def test_noalloc_step_equals_step_twice_then_a_batch(device):
    calls = [((150, 133), 37), ((41, 29), 23)]  # the second call smaller than the first
    for k, ((nx, ny), steps) in enumerate(calls):
        print(f"[no-alloc] call {k}: {nx}x{ny}, {steps} steps", flush=True)
        ref = _run(device, nx, ny, 10 + k, steps, with_strips=True)
        na = _run(device, nx, ny, 10 + k, steps, with_strips=False)
        assert na.pstrip is None
        _step("wf_fdtd2d_step", [ref], steps, device)
        _step("wf_fdtd2d_step_noalloc", [na], steps, device)
        torch.cuda.synchronize()
        for name in ("ex", "ey", "hz", "traces"):
            assert _same(getattr(ref, name), getattr(na, name)), f"call {k}: {name} differs"
        assert float(na.hz.abs().max()) > 0.0
    print("[no-alloc] batch of two runs in one call", flush=True)
    refs = [_run(device, 70, 90, 21, 19, True), _run(device, 33, 47, 22, 19, True)]
    nas = [_run(device, 70, 90, 21, 19, False), _run(device, 33, 47, 22, 19, False)]
    _step("wf_fdtd2d_step", refs, 19, device)
    _step("wf_fdtd2d_step_noalloc", nas, 19, device)
    torch.cuda.synchronize()
    for b in range(2):
        for name in ("ex", "ey", "hz", "traces"):
            assert _same(getattr(refs[b], name), getattr(nas[b], name)), f"batch run {b}: {name} differs"


def test_noalloc_rejects_grids_beyond_its_scratch():
    """4 * (nx + ny) strip floats above 2^18 are refused before any launch."""
    from waveforge_amd import _native

    lib = _native.load()
    r = _native.WfRun2d()
    r.ex = r.ey = r.hz = 16  # never dereferenced: the size check fails first
    r.nx, r.ny = 4, 65533
    rc = lib.wf_fdtd2d_step_noalloc(ctypes.byref(r), 1, 1, None)
    assert rc == -1 and b"scratch" in lib.wf_last_error()
    assert lib.wf_fdtd2d_step_noalloc(None, 1, 1, None) == -1
