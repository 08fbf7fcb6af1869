"""ctypes binding of libwaveforge_hip.so (declared in include/waveforge_hip.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded, every entry point raises ``RuntimeError`` naming the
problem. There is no CPU or eager-torch fallback.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libwaveforge_hip.so"
ABI_VERSION = 3

WF_COMP = {"Ex": 0, "Ey": 1, "Hz": 2}
WF_COMP3 = {"Ex": 0, "Ey": 1, "Hz": 2, "Ez": 3, "Hx": 4, "Hy": 5}
WF_RUN3D_SNAP_FROM_STATE = 1
WF_ADJ_TWO_PASS = 1  # wf_adj_args.flags: two launches per adjoint step
WF_ADJ_ONE_STEP = 2  # wf_adj_args.flags: one launch per step (no two-step launches)
WF_ADJ_FAST_DIV = 4  # wf_adj_args.flags: verified fast division in two-step launches
TB_MAX_DEPTH = 8
WF_RUN2D_RECIPROCAL = 1  # wf_run2d.flags: quotients as products by rdx, rdy (torch's GPU form)
WF_DIV_IEEE, WF_DIV_FAST, WF_DIV_RECIPROCAL = 0, 1, 2  # wf_tb_args.fast_div

# Every symbol the header declares; tests check the built library exports them.
EXPORTED_SYMBOLS = (
    "wf_abi_version",
    "wf_last_error",
    "wf_status_check",
    "wf_fdtd2d_step",
    "wf_fdtd2d_step_noalloc",
    "wf_fdtd2d_maxabs",
    "wf_fdtd2d_maxabs_f64",
    "wf_fdtd2d_tb_advance",
    "wf_fdtd2d_judge",
    "wf_runs_add_step",
    "wf_tb_tile_shapes",
    "wf_div_verify",
    "wf_das2d",
    "wf_das2d_batch",
    "wf_adjoint_step",
    "wf_fdtd2d_step_f64",
    "wf_das_envelope",
    "wf_fdtd3d_step",
    "wf_fdtd3d_strip_floats",
    "wf_fdtd3d_fused_step",
)

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class WfInj2d(ctypes.Structure):
    """``wf_inj2d``."""

    _fields_ = [
        ("wave", _vp),
        ("wave_len", _i32),
        ("comp", _i32),
        ("cell", _i32),
        ("amp", _f32),
    ]


class WfRun2d(ctypes.Structure):
    """``wf_run2d``."""

    _fields_ = [
        ("ex", _vp), ("ey", _vp), ("hz", _vp),
        ("ex2", _vp), ("ey2", _vp), ("hz2", _vp),
        ("ex3", _vp), ("ey3", _vp), ("hz3", _vp),
        ("ca", _vp), ("cb", _vp),
        ("pstrip", _vp),
        ("inj", _vp), ("grp_start", _vp),
        ("det_cell", _vp), ("traces", _vp),
        ("trace_stride", _i64),
        ("n_grp", _i32), ("n_inj", _i32),
        ("n_det", _i32), ("trace_n0", _i32), ("trace_len", _i32),
        ("nx", _i32), ("ny", _i32),
        ("dh", _f32), ("de", _f32), ("dx", _f32), ("dy", _f32),
        ("cmx", _f32), ("cmy", _f32),
        ("step", _i32), ("flags", _i32),
        ("status", _vp),
        ("step_end", _i32),
        ("rdx", _f32), ("rdy", _f32),
        ("pad", _i32),
    ]


class WfTbArgs(ctypes.Structure):
    """``wf_tb_args``."""

    _fields_ = [
        ("runs", _vp), ("tiles", _vp),
        ("n_runs", _i32), ("n_tiles", _i32),
        ("tile_i", _i32), ("tile_j", _i32),
        ("depth", _i32), ("parity", _i32),
        ("step_offset", _i32), ("has_materials", _i32),
        ("maxabs_out", _vp),
        ("rdx", _f32), ("rdy", _f32),
        ("div_lo_x", ctypes.c_uint32), ("div_span_x", ctypes.c_uint32),
        ("div_lo_y", ctypes.c_uint32), ("div_span_y", ctypes.c_uint32),
        ("fast_div", _i32), ("pad", _i32),
        ("stop", _vp),
        ("hist_stride", ctypes.c_int64),
        ("tile_lists", _vp), ("grp_list", _vp), ("det_list", _vp),
    ]


class WfDas2dArgs(ctypes.Structure):
    """``wf_das2d_args``."""

    _fields_ = [
        ("traces", _vp), ("ant_x", _vp), ("ant_y", _vp), ("image", _vp),
        ("n_tx", _i32), ("n_rx", _i32), ("n_t", _i32),
        ("nx_img", _i32), ("ny_img", _i32),
        ("gate_start", _i32), ("clip", _i32), ("pad", _i32),
        ("pitch_x", _f64), ("pitch_y", _f64),
        ("v_medium", _f64), ("dt", _f64),
    ]


class WfAdjArgs(ctypes.Structure):
    """``wf_adj_args``."""

    _fields_ = [(n, _vp) for n in ("ca", "cb", "ex_prev", "ey_prev", "hz_next", "lex", "ley", "lhz", "lex2",
                                   "ley2", "lhz2", "lp", "gca", "gcb", "dl_dd", "rx_cell")] + \
        [(n, _i64) for n in ("run_stride", "state_stride", "dd_run_stride", "dd_rx_stride")] + \
        [(n, _i32) for n in ("nx", "ny", "n_runs", "n_rx", "step")] + \
        [(n, _f32) for n in ("dh", "dx", "dy", "cmx", "cmy")] + [("n_steps", _i32), ("hist_stride", _i64)] + \
        [("rx_first", ctypes.c_void_p), ("rx_order", ctypes.c_void_p), ("flags", _i32), ("pad2", _i32)] + \
        [("div_rx", _f32), ("div_ry", _f32), ("div_span_x", ctypes.c_uint32), ("div_span_y", ctypes.c_uint32)]


class WfInj2dF64(ctypes.Structure):
    """``wf_inj2d_f64``."""

    _fields_ = [("wave", _vp), ("wave_len", _i32), ("comp", _i32), ("cell", _i32), ("pad", _i32), ("amp", _f64)]


class WfRun2dF64(ctypes.Structure):
    """``wf_run2d_f64``."""

    _fields_ = [(n, _vp) for n in ("ex", "ey", "hz", "ca", "cb", "pstrip", "inj", "grp_start", "det_cell",
                                   "traces")] + \
        [("trace_stride", _i64)] + \
        [(n, _i32) for n in ("n_grp", "n_inj", "n_det", "trace_n0", "trace_len", "nx", "ny", "step")] + \
        [(n, _f64) for n in ("dh", "de", "dx", "dy", "cmx", "cmy")] + [("status", _vp)]


class WfF64Args(ctypes.Structure):
    """``wf_f64_args``."""

    _fields_ = [("runs", _vp)] + [(n, _i32) for n in ("n_runs", "max_nx", "max_ny", "n_steps", "step_offset",
                                                       "pad")]


class WfDasEnvArgs(ctypes.Structure):
    """``wf_das_env_args``."""

    _fields_ = [(n, _vp) for n in ("traces", "hilbert_k", "env", "tt", "image", "ant_x", "ant_y", "px", "py",
                                   "layer_r", "layer_eps")] + \
        [(n, _f64) for n in ("cx", "cy", "dt")] + \
        [(n, _i32) for n in ("n_ant", "n_t", "img_w", "img_h", "n_layers", "stages")]


class WfRun3d(ctypes.Structure):
    """``wf_run3d``."""

    _fields_ = [(n, _vp) for n in ("ex", "ey", "ez", "hx", "hy", "hz", "ex2", "ey2", "ez2", "hx2", "hy2", "hz2",
                                   "ca", "cb", "pstrip", "inj", "grp_start", "det_cell", "det_comp", "traces")] + \
        [("trace_stride", _i64)] + \
        [(n, _i32) for n in ("n_grp", "n_inj", "n_det", "trace_n0", "trace_len", "nx", "ny", "nz")] + \
        [(n, _f32) for n in ("dh", "de", "dx", "dy", "dz", "cmx", "cmy", "cmz")] + \
        [("step", _i32), ("flags", _i32), ("status", _vp)]


_lock = threading.Lock()
_lib = None


def load(path: Path | None = None) -> ctypes.CDLL:
    """Load (once) and return the library, with argument types declared."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"waveforge_amd: HIP library {p} is not built. Build it with "
                "`python -m waveforge_amd.build` (hipcc, gfx950); there is no "
                "CPU fallback."
            )
        try:
            lib = ctypes.CDLL(str(p))
        except OSError as exc:
            raise RuntimeError(f"waveforge_amd: cannot load {p}: {exc}") from exc
        lib.wf_abi_version.restype = ctypes.c_int
        lib.wf_last_error.restype = ctypes.c_char_p
        lib.wf_status_check.argtypes = [_vp, _i32, _vp]
        lib.wf_status_check.restype = ctypes.c_int
        lib.wf_fdtd2d_step.argtypes = [ctypes.POINTER(WfRun2d), _i32, _i32, _vp]
        lib.wf_fdtd2d_step.restype = ctypes.c_int
        lib.wf_fdtd2d_step_noalloc.argtypes = [ctypes.POINTER(WfRun2d), _i32, _i32, _vp]
        lib.wf_fdtd2d_step_noalloc.restype = ctypes.c_int
        lib.wf_fdtd2d_maxabs.argtypes = [_vp, _vp, _vp, _i64, _vp, _i32, _vp, _vp]
        lib.wf_fdtd2d_maxabs.restype = ctypes.c_int
        lib.wf_fdtd2d_maxabs_f64.argtypes = [_vp, _vp, _vp, _i64, _vp, _i32, _vp, _vp]
        lib.wf_fdtd2d_maxabs_f64.restype = ctypes.c_int
        lib.wf_fdtd2d_tb_advance.argtypes = [ctypes.POINTER(WfTbArgs), _vp]
        lib.wf_fdtd2d_tb_advance.restype = ctypes.c_int
        lib.wf_tb_tile_shapes.argtypes = [ctypes.POINTER(_i32), _i32]
        lib.wf_tb_tile_shapes.restype = ctypes.c_int
        lib.wf_fdtd2d_judge.argtypes = [_vp, _i32, _f64, _i32, _vp, _vp]
        lib.wf_fdtd2d_judge.restype = ctypes.c_int
        lib.wf_runs_add_step.argtypes = [_vp, _i32, _i32, _vp]
        lib.wf_runs_add_step.restype = ctypes.c_int
        lib.wf_div_verify.argtypes = [_f32, _f32, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp]
        lib.wf_div_verify.restype = ctypes.c_int
        lib.wf_adjoint_step.argtypes = [ctypes.POINTER(WfAdjArgs), _vp]
        lib.wf_adjoint_step.restype = ctypes.c_int
        lib.wf_fdtd2d_step_f64.argtypes = [ctypes.POINTER(WfF64Args), _vp]
        lib.wf_fdtd2d_step_f64.restype = ctypes.c_int
        lib.wf_fdtd3d_step.argtypes = [ctypes.POINTER(WfRun3d), _i32, _i32, _vp]
        lib.wf_fdtd3d_step.restype = ctypes.c_int
        lib.wf_fdtd3d_fused_step.argtypes = [ctypes.POINTER(WfRun3d), _i32, _i32, _vp]
        lib.wf_fdtd3d_fused_step.restype = ctypes.c_int
        lib.wf_fdtd3d_strip_floats.argtypes = [_i32, _i32, _i32]
        lib.wf_fdtd3d_strip_floats.restype = ctypes.c_int64
        lib.wf_das_envelope.argtypes = [ctypes.POINTER(WfDasEnvArgs), _vp]
        lib.wf_das_envelope.restype = ctypes.c_int
        lib.wf_das2d.argtypes = [ctypes.POINTER(WfDas2dArgs), _vp]
        lib.wf_das2d.restype = ctypes.c_int
        lib.wf_das2d_batch.argtypes = [_vp, _i32, _vp]
        lib.wf_das2d_batch.restype = ctypes.c_int
        ver = lib.wf_abi_version()
        if ver != ABI_VERSION:
            raise RuntimeError(f"waveforge_amd: ABI version {ver} != expected {ABI_VERSION}")
        _lib = lib
        return lib


def check(rc: int, what: str) -> None:
    """Raise ``RuntimeError`` with the library's message on a non-zero code."""
    if rc != 0:
        msg = load().wf_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def tile_shapes() -> list[tuple[int, int]]:
    """(tile rows, columns per lane) of every shape compiled into the
    temporally blocked kernel; the tile width is ``64*cpl - 2*depth``."""
    lib = load()
    buf = (_i32 * 64)()
    n = lib.wf_tb_tile_shapes(buf, 32)
    return [(int(buf[2 * k]), int(buf[2 * k + 1])) for k in range(n)]


def tile_rows() -> list[int]:
    """Tile heights (tile_i) compiled for one column per lane."""
    return [ti for ti, cpl in tile_shapes() if cpl == 1]


def tile_cols(depth: int, cpl: int = 1) -> int:
    """Tile width for a launch of ``depth`` steps: one 64-lane wavefront per
    row with ``cpl`` columns per lane, minus the halo."""
    return 64 * cpl - 2 * depth


def stream_handle(device) -> int:
    """Raw ``hipStream_t`` of torch's current stream on ``device``."""
    import torch

    return int(torch.cuda.current_stream(device).cuda_stream)
