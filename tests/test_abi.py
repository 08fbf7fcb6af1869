"""The C-ABI library: loads without a GPU, exports every declared symbol, and
its argument structs have the layout the header gives them (no compute)."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from waveforge_amd import _native

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "waveforge_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(wf_\w+)\s*\(", text, flags=re.M)))


def test_header_declarations_match_binding():
    assert declared_functions() == sorted(_native.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_symbol():
    lib = _native.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.wf_abi_version() == _native.ABI_VERSION
    assert isinstance(lib.wf_last_error(), bytes)


def test_header_constants_match_binding():
    text = HEADER.read_text()
    defs = dict(re.findall(r"^#define (WF_\w+) \(?(-?\d+)\)?", text, flags=re.M))
    assert int(defs["WF_RUN3D_SNAP_FROM_STATE"]) == _native.WF_RUN3D_SNAP_FROM_STATE
    assert int(defs["WF_ADJ_TWO_PASS"]) == _native.WF_ADJ_TWO_PASS
    for name, code in _native.WF_COMP3.items():
        assert int(defs[f"WF_COMP_{name.upper()}"]) == code


def test_tile_shapes_listed():
    shapes = _native.tile_shapes()
    assert (32, 1) in shapes and (48, 2) in shapes and (8, 1) in shapes and all(r > 0 and c in (1, 2) for r, c in shapes)
    assert _native.tile_rows() == [r for r, c in shapes if c == 1]
    assert [_native.tile_cols(d) for d in (1, 4, 8)] == [62, 56, 48]
    assert [_native.tile_cols(d, 2) for d in (1, 4, 8)] == [126, 120, 112]


def test_argument_validation_without_gpu():
    """Bad arguments are rejected on the host before any launch."""
    lib = _native.load()
    rc = lib.wf_fdtd2d_step(None, 0, 1, None)
    assert rc == -1 and b"bad arguments" in lib.wf_last_error()
    a = _native.WfTbArgs()
    assert lib.wf_fdtd2d_tb_advance(ctypes.byref(a), None) == -1
    d = _native.WfDas2dArgs()
    assert lib.wf_das2d(ctypes.byref(d), None) == -1
    # a grid beyond the int32 cell range is refused before any launch
    r = _native.WfRun2d()
    r.ex = r.ey = r.hz = r.pstrip = 16  # never dereferenced: validation fails first
    r.nx = r.ny = 46341
    rc = lib.wf_fdtd2d_step(ctypes.byref(r), 1, 1, None)
    assert rc == -3 and b"2^31-cell index range" in lib.wf_last_error()


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "waveforge_hip.h"
#define O(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("wf_inj2d size %zu\nwf_run2d size %zu\nwf_tb_args size %zu\nwf_das2d_args size %zu\n",
         sizeof(wf_inj2d), sizeof(wf_run2d), sizeof(wf_tb_args), sizeof(wf_das2d_args));
  O(wf_inj2d, wave) O(wf_inj2d, wave_len) O(wf_inj2d, comp) O(wf_inj2d, cell) O(wf_inj2d, amp)
  O(wf_run2d, ex) O(wf_run2d, ey) O(wf_run2d, hz) O(wf_run2d, ex2) O(wf_run2d, ey2) O(wf_run2d, hz2)
  O(wf_run2d, ex3) O(wf_run2d, ey3) O(wf_run2d, hz3) O(wf_run2d, ca) O(wf_run2d, cb) O(wf_run2d, pstrip) O(wf_run2d, inj) O(wf_run2d, grp_start)
  O(wf_run2d, det_cell) O(wf_run2d, traces) O(wf_run2d, trace_stride) O(wf_run2d, n_grp) O(wf_run2d, n_inj)
  O(wf_run2d, n_det) O(wf_run2d, trace_n0) O(wf_run2d, trace_len) O(wf_run2d, nx) O(wf_run2d, ny)
  O(wf_run2d, dh) O(wf_run2d, de) O(wf_run2d, dx) O(wf_run2d, dy) O(wf_run2d, cmx) O(wf_run2d, cmy)
  O(wf_run2d, step) O(wf_run2d, flags) O(wf_run2d, status) O(wf_run2d, step_end) O(wf_run2d, rdx) O(wf_run2d, rdy) O(wf_run2d, pad)
  O(wf_tb_args, runs) O(wf_tb_args, tiles) O(wf_tb_args, n_runs) O(wf_tb_args, n_tiles) O(wf_tb_args, tile_i)
  O(wf_tb_args, tile_j) O(wf_tb_args, depth) O(wf_tb_args, parity) O(wf_tb_args, step_offset)
  O(wf_tb_args, has_materials) O(wf_tb_args, maxabs_out) O(wf_tb_args, rdx) O(wf_tb_args, rdy)
  O(wf_tb_args, div_lo_x) O(wf_tb_args, div_span_x) O(wf_tb_args, div_lo_y) O(wf_tb_args, div_span_y)
  O(wf_tb_args, fast_div) O(wf_tb_args, pad) O(wf_tb_args, stop) O(wf_tb_args, hist_stride)
  O(wf_das2d_args, traces) O(wf_das2d_args, ant_x) O(wf_das2d_args, ant_y) O(wf_das2d_args, image)
  O(wf_das2d_args, n_tx) O(wf_das2d_args, n_rx) O(wf_das2d_args, n_t) O(wf_das2d_args, nx_img)
  O(wf_das2d_args, ny_img) O(wf_das2d_args, gate_start) O(wf_das2d_args, clip) O(wf_das2d_args, pad)
  O(wf_das2d_args, pitch_x) O(wf_das2d_args, pitch_y) O(wf_das2d_args, v_medium) O(wf_das2d_args, dt)
  printf("wf_adj_args size %zu\n", sizeof(wf_adj_args));
  O(wf_adj_args, ca) O(wf_adj_args, hz_next) O(wf_adj_args, lp) O(wf_adj_args, gcb) O(wf_adj_args, dl_dd)
  O(wf_adj_args, rx_cell) O(wf_adj_args, run_stride) O(wf_adj_args, dd_rx_stride) O(wf_adj_args, nx)
  O(wf_adj_args, step) O(wf_adj_args, dh) O(wf_adj_args, cmy) O(wf_adj_args, n_steps) O(wf_adj_args, hist_stride) O(wf_adj_args, rx_first) O(wf_adj_args, rx_order) O(wf_adj_args, flags) O(wf_adj_args, div_rx) O(wf_adj_args, div_span_y)
  printf("wf_inj2d_f64 size %zu\nwf_run2d_f64 size %zu\nwf_f64_args size %zu\n",
         sizeof(wf_inj2d_f64), sizeof(wf_run2d_f64), sizeof(wf_f64_args));
  O(wf_inj2d_f64, wave) O(wf_inj2d_f64, cell) O(wf_inj2d_f64, amp)
  O(wf_run2d_f64, ex) O(wf_run2d_f64, pstrip) O(wf_run2d_f64, traces) O(wf_run2d_f64, trace_stride)
  O(wf_run2d_f64, n_grp) O(wf_run2d_f64, step) O(wf_run2d_f64, dh) O(wf_run2d_f64, cmy) O(wf_run2d_f64, status)
  O(wf_f64_args, runs) O(wf_f64_args, n_runs) O(wf_f64_args, step_offset) O(wf_f64_args, pad)
  printf("wf_das_env_args size %zu\n", sizeof(wf_das_env_args));
  O(wf_das_env_args, traces) O(wf_das_env_args, layer_eps) O(wf_das_env_args, cx) O(wf_das_env_args, dt)
  O(wf_das_env_args, n_ant) O(wf_das_env_args, stages)
  printf("wf_run3d size %zu\n", sizeof(wf_run3d));
  O(wf_run3d, ex) O(wf_run3d, hz) O(wf_run3d, ex2) O(wf_run3d, hz2) O(wf_run3d, ca) O(wf_run3d, pstrip) O(wf_run3d, det_comp) O(wf_run3d, traces)
  O(wf_run3d, trace_stride) O(wf_run3d, n_grp) O(wf_run3d, nz) O(wf_run3d, dh) O(wf_run3d, cmz)
  O(wf_run3d, step) O(wf_run3d, flags) O(wf_run3d, status)
  return 0;
}
"""


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    try:
        res = subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", "-I/opt/rocm/include",
                              str(src), "-o", str(exe)], capture_output=True, text=True)
    except FileNotFoundError as exc:
        pytest.skip(f"no C compiler for the layout probe: {exc}")
    assert res.returncode == 0, f"layout probe does not compile against the header:\n{res.stderr}"
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    types = {"wf_inj2d": _native.WfInj2d, "wf_run2d": _native.WfRun2d, "wf_tb_args": _native.WfTbArgs,
             "wf_das2d_args": _native.WfDas2dArgs, "wf_adj_args": _native.WfAdjArgs,
             "wf_inj2d_f64": _native.WfInj2dF64, "wf_run2d_f64": _native.WfRun2dF64,
             "wf_f64_args": _native.WfF64Args, "wf_das_env_args": _native.WfDasEnvArgs,
             "wf_run3d": _native.WfRun3d}
    checked = 0
    for line in out:
        if not line.strip():
            continue
        t, f, v = line.split()
        if f == "size":
            assert ctypes.sizeof(types[t]) == int(v), line
        else:
            assert getattr(types[t], f).offset == int(v), line
        checked += 1
    assert checked > 50
