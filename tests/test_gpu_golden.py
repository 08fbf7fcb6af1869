"""The HIP product against the reference's own outputs, bit for bit: both
stepping paths on the fixture trajectories, and the DAS kernel on the
fixture images."""

import numpy as np
import pytest
import torch

from golden_cases import CASES, fixtures

pytestmark = pytest.mark.gpu


def _product(F, name, device, depth):
    from waveforge_amd import core
    from waveforge_amd.engine import CellSource

    c = CASES[name]
    g = core.YeeGrid(c["nx"], c["ny"], 1e-3, 1e-3, courant=c["courant"], device=device)
    f = core.FieldSet(g)
    srcs = []
    for comp, wave, ii, jj, amp, kind, pos in c["sources"](F):
        w = wave.to(device)
        if kind == "cells":
            srcs += [CellSource(comp, w, i, j, a) for i, j, a in zip(ii, jj, amp)]
        else:
            class _Plane(core.PlaneSource):
                def __init__(self, wv):
                    self._component, self._plane, self._position = comp, "xz", pos
                    self._N_steps, self._n_cells, self._waveform_tensor = len(wv), g.Nx, wv
            srcs.append(_Plane(w))
    det = core.DetectorSet(c["detectors"], c["steps"], grid=g) if c["detectors"] else None
    sim = core.FDTD2D(g, f, core.MurABC(g, f.Hz), core.SourceCollection(srcs),
                      Ca=torch.from_numpy(F[f"{name}_Ca"]).to(device), Cb=torch.from_numpy(F[f"{name}_Cb"]).to(device),
                      n_check=c["n_check"], detectors=det, depth=depth, division="ieee")
    return sim, f, det


@pytest.mark.parametrize("depth", [None, 6])
@pytest.mark.parametrize("name", sorted(CASES))
def test_product_trajectory_equals_reference(device, name, depth):
    F = fixtures()
    c = CASES[name]
    sim, f, det = _product(F, name, device, depth)
    if depth is None:
        for _ in range(c["steps"]):
            sim.step()
    else:
        sim.run(c["steps"])
    torch.cuda.synchronize()
    for comp in ("Ex", "Ey", "Hz"):
        np.testing.assert_array_equal(getattr(f, comp).cpu().numpy(), F[f"{name}_{comp}"], err_msg=f"{name} {comp}")
    if det is not None:
        np.testing.assert_array_equal(det.traces.cpu().numpy(), F[f"{name}_traces"])
        assert sim.last_field_max == float(F["slab_last_field_max"])


def test_gpu_das_equals_reference(device):
    from waveforge_amd import core

    F = fixtures()
    g = core.YeeGrid(30, 30, 1e-3, 1e-3, device=device)
    ants = [tuple(int(v) for v in a) for a in F["das_antennas"]]
    np.testing.assert_array_equal(core.das_backprojection(F["das_traces"], ants, g, 200), F["das_image"])
    np.testing.assert_array_equal(core.das_backprojection(F["das_traces"], ants, g, 200, (20, 25)),
                                  F["das_image_20x25"])
    g9 = core.YeeGrid(40, 40, 1e-3, 1e-3, device=device)
    ants9 = [tuple(int(v) for v in a) for a in F["das09_antennas"]]
    gate = max(0, int(0.9e-9 / g9.dt))
    np.testing.assert_array_equal(core.das_backprojection(F["das09_traces"], ants9, g9, 500, gate_start=gate),
                                  F["das09_image"])
