"""Pin the torch oracle with the independent per-cell C stepper.

``oracle/fdtd2d_ref.py`` restates the reference's eager slice operations;
``oracle/fdtd2d_brute.c`` is written from the update equations as scalar
loops. They share no code, so an off-by-one slice, a swapped sign or operand,
or a Mur edge in the wrong order makes them disagree. In fp32 they must agree
bit for bit (both round every operation once, in the same order); the fp64
variants must agree too. Both division forms are pinned: IEEE ``x / fp32(d)``
(the reference's ``x / d`` on the CPU) and ``x * fp32(1/d)`` (the same
expression on a GPU; ``division="reciprocal"`` in the oracle, the ``*_rcp``
entry points in C).
"""

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from cases import build_oracle, case

ORACLE = Path(__file__).resolve().parent.parent / "oracle"
LIB = ORACLE / "build" / "libwf_brute.so"


@pytest.fixture(scope="module")
def brute():
    subprocess.run(["make", "-s", "-C", str(ORACLE)], check=True)
    lib = ctypes.CDLL(str(LIB))
    for name, ct in (("wf_brute_step_f32", ctypes.c_float), ("wf_brute_step_f64", ctypes.c_double),
                     ("wf_brute_step_f32_rcp", ctypes.c_float)):
        fn = getattr(lib, name)
        P = ctypes.c_void_p
        fn.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, ct, ct, ct, ct, ct, ct, ctypes.c_int, P, P, P]
    return lib


def brute_run(lib, o, steps, dtype):
    """Run the C stepper from a fresh copy of the oracle's initial state."""
    npdt = np.float32 if dtype == torch.float32 else np.float64
    nx, ny = o.Nx, o.Ny
    ex = o.Ex.reshape(nx, ny).numpy().astype(npdt).copy()
    ey = o.Ey.reshape(nx, ny).numpy().astype(npdt).copy()
    hz = o.Hz.reshape(nx, ny).numpy().astype(npdt).copy()
    ca = cb = None
    if o.Ca is not None:
        ca = o.Ca.reshape(nx, ny).numpy().astype(npdt).copy()
        cb = o.Cb.reshape(nx, ny).numpy().astype(npdt).copy()
    rcp = o.division == "reciprocal"
    fn = lib.wf_brute_step_f32 if dtype == torch.float32 else lib.wf_brute_step_f64
    qx, qy = o.dx, o.dy
    if rcp:  # the C entry takes the reciprocals in place of the cell sizes
        assert dtype == torch.float32
        fn, qx, qy = lib.wf_brute_step_f32_rcp, float(o._rdx), float(o._rdy)
    comp_code = {"Ex": 0, "Ey": 1, "Hz": 2}
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    traces = []
    for n in range(steps):
        comps, cells, vals = [], [], []
        for s in o.sources:
            if s.kind == "cells":
                v = (s.wave[n:n + 1] * s.amp).numpy().astype(npdt)
                for i, j, val in zip(s.i_idx.tolist(), s.j_idx.tolist(), v.tolist()):
                    comps.append(comp_code[s.component]); cells.append(i * ny + j); vals.append(val)
            else:
                w = float(s.wave[n])
                if s.kind == "plane_xz":
                    cl = [i * ny + s.position for i in range(nx)]
                else:
                    cl = [s.position * ny + j for j in range(ny)]
                for cc in cl:
                    comps.append(comp_code[s.component]); cells.append(cc); vals.append(w)
        comps_a = np.array(comps, dtype=np.int32)
        cells_a = np.array(cells, dtype=np.int32)
        vals_a = np.array(vals, dtype=npdt)
        fn(nx, ny, ptr(ex), ptr(ey), ptr(hz), ptr(ca), ptr(cb),
           o.Dh, o.De, qx, qy, o.cx, o.cy, len(comps), ptr(comps_a), ptr(cells_a), ptr(vals_a))
        if o.detectors:
            traces.append([hz[i, j] for i, j in o.detectors])
    return ex, ey, hz, np.array(traces, dtype=npdt).T if traces else None


BRUTE_CASES = [
    case(4, 4, steps=10, init=True),
    case(5, 9, steps=20, init=True, material=True, seed=3),
    case(23, 17, steps=40, material=True, sources=[
        ("point", "Hz", 11, 8, "dgauss"), ("point", "Hz", 1, 3, "gauss"), ("point", "Ex", 0, 5, "sin"),
        ("line", "Ey", "x", 12, 2, 15, "gauss", [0.1 * k for k in range(13)]), ("point", "Hz", 11, 8, "sin"),
        ("plane", "Hz", "yz", 21, "dgauss"), ("point", "Hz", 22, 16, "gauss")],
        detectors=[(0, 0), (11, 8), (22, 16), (1, 3)]),
    case(31, 30, steps=60, init=True, sources=[("plane", "Ey", "xz", 4, "sin"), ("point", "Hz", 15, 15, "gauss")]),
]


@pytest.mark.parametrize("k", range(len(BRUTE_CASES)))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_oracle_matches_brute(brute, k, dtype):
    c = BRUTE_CASES[k]
    o = build_oracle(c, dtype=dtype)
    ex, ey, hz, tr = brute_run(brute, o, c["steps"], dtype)
    o.run(c["steps"])
    np.testing.assert_array_equal(o.Ex.reshape(c["nx"], c["ny"]).numpy(), ex)
    np.testing.assert_array_equal(o.Ey.reshape(c["nx"], c["ny"]).numpy(), ey)
    np.testing.assert_array_equal(o.Hz.reshape(c["nx"], c["ny"]).numpy(), hz)
    if tr is not None:
        np.testing.assert_array_equal(o.trace_array().numpy(), tr)
    # a non-trivial trajectory: fields moved away from the initial state
    assert float(o.Hz.abs().max()) > 0.0


@pytest.mark.parametrize("k", range(len(BRUTE_CASES)))
def test_reciprocal_oracle_matches_brute(brute, k):
    """The oracle's reciprocal form (the reference's quotients as torch forms
    them on a GPU) equals the C loops with x * fp32(1/d), bit for bit."""
    c = BRUTE_CASES[k]
    o = build_oracle(c, division="reciprocal")
    ex, ey, hz, tr = brute_run(brute, o, c["steps"], torch.float32)
    o.run(c["steps"])
    for got, want in ((o.Ex, ex), (o.Ey, ey), (o.Hz, hz)):
        assert np.array_equal(got.reshape(c["nx"], c["ny"]).numpy().view(np.int32), want.view(np.int32))
    if tr is not None:
        assert np.array_equal(o.trace_array().numpy().view(np.int32), tr.view(np.int32))


def test_reciprocal_and_ieee_forms_differ_but_agree_closely():
    """The two division forms are different arithmetic (the reciprocal of
    dx = 1 mm rounds to 1000 exactly, not to 1/fp32(0.001)), so their fields
    part after a few steps, but only by fp32 rounding: well inside the 1e-5
    relative L2 parity bar after 300 steps of a lossy medium."""
    c = case(40, 44, steps=300, material=True, init=True, seed=4,
             sources=[("point", "Hz", 20, 22, "dgauss"), ("point", "Ey", 7, 30, "sin")])
    a = build_oracle(c)
    b = build_oracle(c, division="reciprocal")
    assert float(b._rdx) == 1000.0 and np.float32(1.0) / np.float32(1e-3) != np.float32(1000.0)
    a.run(c["steps"])
    b.run(c["steps"])
    for name in ("Ex", "Ey", "Hz"):
        x, y = a.fields[name].double(), b.fields[name].double()
        assert not torch.equal(x, y)
        assert float(torch.linalg.norm(x - y) / torch.linalg.norm(x)) < 1e-5


def test_brute_detects_planted_reciprocal_error(brute):
    """A reciprocal taken of the fp32 cell size (1/fp32(dx), 999.99994 at
    1 mm) instead of the double's must be caught by the comparison."""
    c = BRUTE_CASES[2]
    o = build_oracle(c, division="reciprocal")
    ex, ey, hz, _ = brute_run(brute, o, 12, torch.float32)
    o._rdx = torch.tensor(np.float32(1.0) / np.float32(o.dx))  # plant the mistake in the oracle (CPU only)
    o._rdy = torch.tensor(np.float32(1.0) / np.float32(o.dy))
    o.run(12)
    assert not np.array_equal(o.Hz.reshape(c["nx"], c["ny"]).numpy(), hz)


def test_brute_detects_planted_sign_error(brute):
    """A flipped Faraday sign in the oracle must be caught by the comparison."""
    c = BRUTE_CASES[2]
    o = build_oracle(c)
    ex, ey, hz, _ = brute_run(brute, o, 5, torch.float32)
    o.Dh = -o.Dh  # plant the mistake in the oracle (CPU only)
    o.run(5)
    assert not np.array_equal(o.Hz.reshape(c["nx"], c["ny"]).numpy(), hz)


def brute_run_many(lib, o, steps):
    """The same run through wf_brute_run_f32 (all steps in one call, rows in
    parallel): point/line sources only, as the full-size workloads use."""
    lib.wf_brute_run_f32.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5 + \
        [ctypes.c_float] * 6 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int] + \
        [ctypes.c_void_p] * 2
    nx, ny = o.Nx, o.Ny
    ex, ey, hz = (getattr(o, k).reshape(nx, ny).numpy().astype(np.float32).copy() for k in ("Ex", "Ey", "Hz"))
    ca = cb = None
    if o.Ca is not None:
        ca = o.Ca.reshape(nx, ny).numpy().astype(np.float32).copy()
        cb = o.Cb.reshape(nx, ny).numpy().astype(np.float32).copy()
    comp_code = {"Ex": 0, "Ey": 1, "Hz": 2}
    comps, cells, vals = [], [], []
    for s in o.sources:
        assert s.kind == "cells"
        for c, (i, j) in enumerate(zip(s.i_idx.tolist(), s.j_idx.tolist())):
            comps.append(comp_code[s.component])
            cells.append(i * ny + j)
            vals.append((s.wave[:steps] * s.amp[c]).numpy().astype(np.float32))
    vals_a = np.ascontiguousarray(np.stack(vals)) if vals else np.zeros((0, steps), np.float32)
    det = np.array([i * ny + j for i, j in o.detectors], dtype=np.int32)
    tr = np.zeros((len(det), steps), dtype=np.float32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    lib.wf_brute_run_f32(nx, ny, ptr(ex), ptr(ey), ptr(hz), ptr(ca), ptr(cb), o.Dh, o.De, o.dx, o.dy, o.cx, o.cy,
                         steps, len(comps), ptr(np.array(comps, np.int32)), ptr(np.array(cells, np.int32)),
                         ptr(vals_a), len(det), ptr(det), ptr(tr))
    return ex, ey, hz, tr


def test_multi_step_entry_equals_step_by_step(brute):
    c = case(61, 47, steps=80, material=True, init=True, seed=9,
             sources=[("point", "Hz", 30, 20, "dgauss"), ("point", "Ey", 0, 11, "gauss"),
                      ("line", "Ex", "x", 0, 3, 40, "sin", [0.2 + 0.01 * k for k in range(37)]),
                      ("point", "Hz", 30, 20, "sin")],
             detectors=[(0, 0), (30, 20), (60, 46), (5, 40)])
    o = build_oracle(c)
    ex, ey, hz, tr = brute_run_many(brute, o, c["steps"])
    ex1, ey1, hz1, tr1 = brute_run(brute, build_oracle(c), c["steps"], torch.float32)
    for a, b in ((ex, ex1), (ey, ey1), (hz, hz1), (tr, tr1)):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    o.run(c["steps"])
    assert np.array_equal(hz.view(np.int32), o.Hz.reshape(61, 47).numpy().view(np.int32))
    assert np.array_equal(tr.view(np.int32), o.trace_array().numpy().view(np.int32))
