"""The device status word through the C ABI.

The reference raises when a step runs past a source's waveform table
(``src/core/sources.py:882-887``); the Python layer checks this before it
launches, but a direct ABI caller relies on the kernels: a contribution past
the table is not applied and raises WF_STATUS_WAVE_END, and a tile region
with more source groups (or a tile core with more detectors) than the
blocked kernel stages raises WF_STATUS_TILE_GROUPS (WF_STATUS_TILE_DETS).
``wf_status_check`` turns a raised word into WF_ERR_ARG with a message.
"""

import ctypes

import numpy as np
import pytest
import torch

from cases import assert_bitwise, waveform_table
from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource, stable_dt

pytestmark = pytest.mark.gpu

DX = 1e-3
WF_STATUS_WAVE_END, WF_STATUS_TILE_GROUPS, WF_STATUS_TILE_DETS = 1, 2, 4


def _run(device, n, sources=(), dets=(), n_record=0):
    """A free-space n x n run with point sources (comp, i, j, wave) and Hz
    detectors, both state buffers and a zeroed status word."""
    from waveforge_amd.engine import CellSource, RunState, SourceTable, StepCoefficients

    dt = stable_dt(DX, DX)
    f = torch.zeros((6, n, n), dtype=torch.float32, device=device)
    table = SourceTable([CellSource(c, w, i, j, 1.0) for c, i, j, w in sources], n, n, device)
    rs = RunState(f[0], f[1], f[2], None, None, n, n, StepCoefficients.from_grid(dt, DX, DX), sources=table,
                  status=torch.zeros(1, dtype=torch.int32, device=device))
    rs.second = (f[3], f[4], f[5])
    rs.pstrip = torch.empty(4 * (n + n), dtype=torch.float32, device=device)
    if dets:
        rs.det_cell = torch.tensor([i * n + j for i, j in dets], dtype=torch.int32, device=device)
        rs.traces = torch.zeros((len(dets), max(1, n_record)), dtype=torch.float32, device=device)
    return rs, dt


def _check(rs, device):
    from waveforge_amd import _native

    lib = _native.load()
    rc = lib.wf_status_check(rs.status.data_ptr(), 1, ctypes.c_void_p(_native.stream_handle(device)))
    return rc, lib.wf_last_error().decode(), int(rs.status.item())


def test_clean_run_reports_ok(device):
    from waveforge_amd import _native
    from waveforge_amd.engine import TBPlan

    wave = torch.ones(40, device=device)
    rs, _ = _run(device, 96, sources=[("Hz", 40, 40, wave)], dets=[(10, 10)], n_record=16)
    plan = TBPlan([rs], 48, device, fast_div=False)
    plan.advance(16, 8, _native.stream_handle(device))
    rc, msg, word = _check(rs, device)
    print(f"[status] clean run: rc={rc} word={word}", flush=True)
    assert rc == 0 and word == 0


def test_tile_source_overflow_is_an_error(device):
    """65 source cells inside one tile region of a blocked launch."""
    from waveforge_amd import _native
    from waveforge_amd.engine import TBPlan

    wave = torch.ones(40, device=device)
    cells = [(10 + k // 9, 10 + k % 9) for k in range(65)]
    rs, _ = _run(device, 96, sources=[("Hz", i, j, wave) for i, j in cells])
    plan = TBPlan([rs], 48, device, fast_div=False)
    plan.advance(16, 8, _native.stream_handle(device))
    rc, msg, word = _check(rs, device)
    print(f"[status] 65 groups: rc={rc} word={word} msg={msg!r}", flush=True)
    assert rc == -1 and word & WF_STATUS_TILE_GROUPS and "source groups" in msg


def test_tile_detector_overflow_is_an_error(device):
    from waveforge_amd import _native
    from waveforge_amd.engine import TBPlan

    dets = [(20 + k // 9, 20 + k % 9) for k in range(65)]
    rs, _ = _run(device, 96, dets=dets, n_record=16)
    plan = TBPlan([rs], 48, device, fast_div=False)
    plan.advance(16, 8, _native.stream_handle(device))
    rc, msg, word = _check(rs, device)
    print(f"[status] 65 detectors: rc={rc} word={word}", flush=True)
    assert rc == -1 and word == WF_STATUS_TILE_DETS and "detectors" in msg


@pytest.mark.parametrize("path", ["per_step", "blocked"])
def test_step_past_waveform_end_is_an_error_and_injects_nothing(device, path):
    """A 10-sample table stepped 24 times: the status word says so, and the
    fields equal the oracle's with the table padded by zeros (nothing was
    read past the table)."""
    from waveforge_amd import _native
    from waveforge_amd.engine import TBPlan, per_step_advance

    n, steps, n_wave = 40, 24, 10
    dt = stable_dt(DX, DX)
    w = waveform_table("gauss", n_wave, dt)
    w[:] = torch.linspace(0.5, 1.0, n_wave)  # nonzero to the last sample
    rs, _ = _run(device, n, sources=[("Hz", 20, 17, w.to(device))], dets=[(5, 5), (20, 17)], n_record=steps)
    stream = _native.stream_handle(device)
    if path == "per_step":
        per_step_advance([rs], steps, stream)
    else:
        TBPlan([rs], 16, device, fast_div=False).advance(steps, 8, stream)
    rc, msg, word = _check(rs, device)
    print(f"[status] {path} past the table: rc={rc} word={word}", flush=True)
    assert rc == -1 and word == WF_STATUS_WAVE_END and "waveform table" in msg
    padded = torch.cat([w, torch.zeros(steps - n_wave)])
    o = OracleFDTD2D(n, n, DX, DX, dt, sources=[OracleSource("Hz", padded, torch.tensor([20]), torch.tensor([17]),
                                                            torch.tensor([1.0]))], detectors=[(5, 5), (20, 17)],
                     n_check=10 ** 9)
    o.run(steps)
    assert_bitwise(rs.hz, o.Hz.reshape(n, n), f"{path} Hz")
    assert_bitwise(rs.ex, o.Ex.reshape(n, n), f"{path} Ex")
    assert_bitwise(rs.traces, o.trace_array(), f"{path} traces")


def test_status_check_rejects_bad_arguments():
    from waveforge_amd import _native

    lib = _native.load()
    assert lib.wf_status_check(None, 1, None) == -1
    assert b"bad arguments" in lib.wf_last_error()


def test_odd_launch_count_uses_the_third_buffer(device):
    """20 steps at depth 8 run as three launches (7, 7, 6) through the third
    state buffer and end in the FieldSet tensors, equal to the oracle."""
    from cases import build_oracle, build_product, case

    c = case(60, 70, steps=20, material=True, init=True, sources=[("point", "Hz", 30, 35, "dgauss")],
             detectors=[(2, 3), (30, 35)])
    sim, det = build_product(c, device, depth=8, tile=16)
    o = build_oracle(c)
    sim.run(20)
    o.run(20)
    torch.cuda.synchronize()
    assert sim._third is not None
    for k in ("Ex", "Ey", "Hz"):
        assert_bitwise(getattr(sim.fields, k), o.fields[k], k)
    assert_bitwise(det.traces, o.trace_array(), "traces")
    assert np.isfinite(sim.fields.Hz.cpu().numpy()).all()
