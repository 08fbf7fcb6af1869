"""HIP stepping kernels vs the CPU oracle: bit-exact fp32 fields and traces.

Two division forms are checked. ``division="ieee"`` (IEEE ``x / fp32(d)``)
against the oracle as the reference runs on the CPU; ``division="reciprocal"``
(``x * fp32(1/d)``, the default) against the oracle's reciprocal form and
against the reference's own expressions evaluated by torch on this GPU,
which form the quotient that way.

Both the per-step path (``wf_fdtd2d_step``) and the temporally blocked path
(``wf_fdtd2d_tb_advance``) must reproduce the oracle bit for bit: the kernels
use the reference's fp32 operation order and no fused multiply-adds. Cases
cover tiny grids (4x4 and odd sizes), grids that are not multiples of the
tile, free space and lossy media, random initial fields (also ones spread
from subnormals to 2^80 with both zeros, for the division), every source kind
(point, line with amplitude profile, plane) on Ex, Ey and Hz, sources on Mur
rows/columns and stacked on one cell, detectors, and batched FieldSets.
"""

import pytest
import torch

from cases import assert_bitwise, build_oracle, build_product, case

pytestmark = pytest.mark.gpu

CASES = {
    "tiny4_free": case(4, 4, steps=12, init=True),
    "tiny5x7_mat": case(5, 7, steps=15, init=True, material=True),
    "free_point_hz": case(40, 33, steps=60, sources=[("point", "Hz", 20, 16, "gauss")], detectors=[(5, 5), (30, 20)]),
    "mat_all_sources": case(
        70, 90, steps=50, material=True, init=True,
        sources=[
            ("point", "Hz", 35, 40, "dgauss"),
            ("point", "Ex", 10, 12, "sin"),
            ("line", "Ey", "x", 60, 3, 50, "gauss", None),
            ("line", "Hz", "y", 5, 0, 90, "sin", [0.5 + 0.01 * k for k in range(90)]),
            ("plane", "Hz", "xz", 7, "dgauss"),
            ("point", "Hz", 35, 40, "sin"),
        ],
        detectors=[(0, 0), (69, 89), (35, 40), (1, 1), (68, 2)],
    ),
    "edge_sources": case(
        20, 24, steps=30, material=True,
        sources=[
            ("point", "Hz", 1, 5, "gauss"), ("point", "Hz", 0, 6, "gauss"), ("point", "Hz", 18, 1, "sin"),
            ("point", "Hz", 5, 22, "dgauss"), ("point", "Ex", 0, 0, "sin"), ("point", "Ey", 19, 23, "gauss"),
            ("point", "Hz", 19, 23, "gauss"), ("point", "Hz", 1, 1, "sin"),
        ],
        detectors=[(0, 6), (1, 5), (19, 23)],
    ),
    # the cells Ampere never writes (Ex of column 0, Ey of row 0) carry values
    # and take soft sources, on a grid wider than one two-column region
    "face_cells": case(
        23, 150, steps=20, material=True, init=True, seed=3,
        sources=[("point", "Ex", 9, 0, "sin"), ("point", "Ey", 0, 7, "gauss"), ("point", "Ey", 0, 140, "sin"),
                 ("point", "Ex", 0, 0, "gauss"), ("line", "Ex", "x", 0, 2, 20, "dgauss", None)],
        detectors=[(0, 0), (11, 1), (22, 149)],
    ),
    "odd_tiles_mat": case(133, 71, steps=41, material=True, sources=[("point", "Hz", 66, 35, "dgauss")],
                          detectors=[(10, 10), (120, 60), (66, 35)]),
    "batch3": case(17, 19, steps=25, init=True, batch=3, sources=[("point", "Hz", 8, 9, "gauss")],
                   detectors=[(3, 4)]),
    "extreme_mat": case(45, 70, steps=5, init="extreme", material=True, seed=5),
    "extreme_free": case(38, 66, steps=5, init="extreme", seed=6),
}


def _compare(sim, det, o, name):
    f = sim.fields
    assert_bitwise(f.Ex, o.Ex, f"{name} Ex")
    assert_bitwise(f.Ey, o.Ey, f"{name} Ey")
    assert_bitwise(f.Hz, o.Hz, f"{name} Hz")
    if det is not None:
        assert_bitwise(det.traces, o.trace_array(), f"{name} traces")


@pytest.mark.parametrize("name", sorted(CASES))
def test_per_step_path_bitwise(device, name):
    c = CASES[name]
    print(f"[per-step] case={name} grid={c['nx']}x{c['ny']} steps={c['steps']}", flush=True)
    sim, det = build_product(c, device)
    o = build_oracle(c)
    for _ in range(c["steps"]):
        sim.step()
        o.step()
    torch.cuda.synchronize()
    _compare(sim, det, o, name)


@pytest.mark.parametrize("fast_div", [True, False])
@pytest.mark.parametrize("tile", [32, 16, 48, 64, 8, (16, 2), (32, 2), (48, 2)])
@pytest.mark.parametrize("depth", [1, 2, 3, 4, 5, 8])
@pytest.mark.parametrize("name", ["tiny5x7_mat", "mat_all_sources", "edge_sources", "face_cells", "odd_tiles_mat",
                                  "batch3", "free_point_hz", "extreme_mat", "extreme_free"])
def test_tb_path_bitwise(device, name, tile, depth, fast_div):
    c = CASES[name]
    print(f"[tb] case={name} tile={tile} depth={depth} fast_div={fast_div}", flush=True)
    sim, det = build_product(c, device, depth=depth, tile=tile, fast_div=fast_div)
    o = build_oracle(c)
    sim.run(c["steps"])
    o.run(c["steps"])
    torch.cuda.synchronize()
    _compare(sim, det, o, name)


def test_mixed_step_and_run(device):
    """step() and run() interleave on the same live tensors; user writes between calls are honoured."""
    c = CASES["mat_all_sources"]
    sim, det = build_product(c, device, depth=4)
    o = build_oracle(c)
    sim.step(); o.step()
    sim.run(7); o.run(7)
    sim.fields.Hz[30:32, 10:40] *= 0.5
    o.Hz[30:32, 10:40] *= 0.5
    sim.run(20); o.run(20)
    sim.step(); o.step()
    sim.run(3); o.run(3)
    torch.cuda.synchronize()
    f = sim.fields
    assert_bitwise(f.Ex, o.Ex, "Ex")
    assert_bitwise(f.Ey, o.Ey, "Ey")
    assert_bitwise(f.Hz, o.Hz, "Hz")
    assert sim.steps_completed == o.steps_completed == 32


def test_fast_division_is_licensed_on_the_device(device):
    """The exhaustive device check passes for the divisors the tests use, so
    the fast_div=True cases above really run the fast quotient; stretched to
    every finite float, the same check counts the mismatches above the range
    (where x * 2^(31-e) overflows)."""
    import ctypes

    from waveforge_amd import _native
    from waveforge_amd.engine import division_range, f32, verified_division

    for d in (1e-3, 5e-4, 2.5e-3):
        v = verified_division(d, device)
        assert v is not None, d
        r, lo, span = v
        assert (lo, span) == division_range(f32(d))
    d = f32(1e-3)
    r, lo, span = verified_division(d, device)
    bad = torch.zeros(1, dtype=torch.int32, device=device)
    lib = _native.load()
    _native.check(lib.wf_div_verify(d, r, 0, 0x7F800000, bad.data_ptr(),
                                    ctypes.c_void_p(_native.stream_handle(device))), "wf_div_verify")
    print(f"[fast-div] mismatches over every finite float: {int(bad.item())}", flush=True)
    assert int(bad.item()) > 0


RCP_CASES = ["tiny4_free", "tiny5x7_mat", "mat_all_sources", "edge_sources", "face_cells", "odd_tiles_mat",
             "batch3", "free_point_hz", "extreme_mat", "extreme_free"]


def _gpu_reference(c, device):
    """The oracle's literal form of the reference step (``x / dx``),
    evaluated by torch on the GPU."""
    return build_oracle(c).to(device)


@pytest.mark.parametrize("name", RCP_CASES)
def test_per_step_path_reciprocal_bitwise(device, name):
    c = CASES[name]
    print(f"[per-step rcp] case={name} grid={c['nx']}x{c['ny']} steps={c['steps']}", flush=True)
    sim, det = build_product(c, device, division="reciprocal")
    o = build_oracle(c, division="reciprocal")
    g = _gpu_reference(c, device)
    for _ in range(c["steps"]):
        sim.step()
        o.step()
        g.step()
    torch.cuda.synchronize()
    _compare(sim, det, o, name + " vs CPU reciprocal form")
    for k in ("Ex", "Ey", "Hz"):
        assert_bitwise(g.fields[k], o.fields[k], f"{name} {k}: torch-GPU reference vs CPU reciprocal form")
    if det is not None:
        assert_bitwise(g.trace_array().cpu(), o.trace_array(), f"{name} traces: torch-GPU reference")


@pytest.mark.parametrize("tile", [32, 16, 48, 64, 8, (16, 2), (32, 2), (48, 2)])
@pytest.mark.parametrize("depth", [1, 2, 3, 5, 7, 8])
@pytest.mark.parametrize("name", RCP_CASES)
def test_tb_path_reciprocal_bitwise(device, name, tile, depth):
    c = CASES[name]
    print(f"[tb rcp] case={name} tile={tile} depth={depth}", flush=True)
    sim, det = build_product(c, device, depth=depth, tile=tile, division="reciprocal")
    o = build_oracle(c, division="reciprocal")
    sim.run(c["steps"])
    o.run(c["steps"])
    torch.cuda.synchronize()
    _compare(sim, det, o, name)


def test_default_division_is_the_reference_on_the_gpu(device):
    """An FDTD2D built without a division argument steps as the reference's
    expressions do under torch on this GPU (blocked and per-step paths)."""
    from waveforge_amd import core

    c = CASES["mat_all_sources"]
    sim, det = build_product(c, device, division="reciprocal", depth=8, tile=48)
    assert sim._division == "reciprocal"
    grid = core.YeeGrid(8, 8, 1e-3, 1e-3, device=device)
    f = core.FieldSet(grid)
    assert core.FDTD2D(grid, f, core.MurABC(grid, f.Hz))._division == "reciprocal"
    g = _gpu_reference(c, device)
    sim.run(c["steps"])
    g.run(c["steps"])
    torch.cuda.synchronize()
    _compare(sim, det, g, "blocked path vs torch-GPU reference")


@pytest.mark.parametrize("d", [1e-3, 5e-4, 2e-3, 7e-4, 1.1e-3, 2.5e-4, 0.37, 3.7])
def test_torch_gpu_divides_by_a_python_float_as_a_reciprocal_product(device, d):
    """The fact the reciprocal division form rests on: torch on this GPU
    evaluates ``float32 tensor / python_float`` as ``x * fp32(1/d)`` (the
    reciprocal of the double rounded once), not as IEEE ``x / fp32(d)``,
    which torch computes on the CPU. Values span subnormals to 2^100, both
    signs and zeros."""
    import numpy as np

    rng = np.random.default_rng(11)
    n = 1 << 20
    x = (rng.choice([-1.0, 1.0], n) * rng.uniform(1, 2, n) * np.exp2(rng.uniform(-140, 100, n))).astype(np.float32)
    x[:8] = 0.0
    x[8:16] = -0.0
    q_gpu = (torch.from_numpy(x).to(device) / d).cpu().numpy()
    rcp = (x * np.float32(1.0 / d)).astype(np.float32)
    with np.errstate(all="ignore"):
        ieee = (x / np.float32(d)).astype(np.float32)
    assert np.array_equal(q_gpu.view(np.uint32), rcp.view(np.uint32))
    q_cpu = (torch.from_numpy(x) / d).numpy()
    assert np.array_equal(q_cpu.view(np.uint32), ieee.view(np.uint32))
