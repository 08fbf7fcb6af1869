"""The oracle (and the product's host-side tables) against the reference's
own outputs (``tests/golden/reference_fixtures.npz``), bit for bit."""

import numpy as np
import pytest
import torch

from golden_cases import CASES, fixtures
from oracle import das_ref
from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource, stable_dt
from waveforge_amd import core


@pytest.fixture(scope="module")
def F():
    return fixtures()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_trajectory_equals_reference(F, name):
    c = CASES[name]
    dt = stable_dt(1e-3, 1e-3, c["courant"])
    srcs = []
    for comp, wave, ii, jj, amp, kind, pos in c["sources"](F):
        if kind == "cells":
            srcs.append(OracleSource(comp, wave, torch.tensor(ii), torch.tensor(jj), torch.tensor(amp, dtype=torch.float32)))
        else:
            srcs.append(OracleSource(comp, wave, kind=kind, position=pos))
    o = OracleFDTD2D(c["nx"], c["ny"], 1e-3, 1e-3, dt, Ca=torch.from_numpy(F[f"{name}_Ca"]),
                     Cb=torch.from_numpy(F[f"{name}_Cb"]), sources=srcs, detectors=c["detectors"],
                     n_check=c["n_check"])
    o.run(c["steps"])
    for comp in ("Ex", "Ey", "Hz"):
        np.testing.assert_array_equal(o.fields[comp].numpy(), F[f"{name}_{comp}"])
    if c["detectors"]:
        np.testing.assert_array_equal(o.trace_array().numpy(), F[f"{name}_traces"])
    if name == "slab":
        assert o.last_field_max == float(F["slab_last_field_max"])


def test_product_material_map_equals_reference(F):
    g = core.YeeGrid(48, 48, 1e-3, 1e-3, courant=0.99, device="cpu")
    mm = core.MaterialMap(g)
    mm.add_rectangle((28, 33), (0, 47), core.Material("slab", 4.0, 0.0))
    Ca, Cb = mm.build()
    np.testing.assert_array_equal(Ca.numpy(), F["slab_Ca"])
    np.testing.assert_array_equal(Cb.numpy(), F["slab_Cb"])
    g = core.YeeGrid(40, 36, 1e-3, 1e-3, device="cpu")
    mm = core.MaterialMap(g, default=core.Material("coupling", 20.0, 0.1))
    mm.add_circle((20, 18), 12, core.Material("brain", 45.0, 0.8))
    mm.add_circle((24, 15), 4, core.Material("clot", 60.0, 1.2))
    Ca, Cb = mm.build()
    np.testing.assert_array_equal(Ca.numpy(), F["lossy_Ca"])
    np.testing.assert_array_equal(Cb.numpy(), F["lossy_Cb"])


def test_product_waveforms_equal_reference(F):
    import math

    dt = float(F["wave_dt"])
    waves = {
        "gaussian": core.GaussianPulse(1.3, sigma=7e-12),
        "gaussian_freq": core.GaussianPulse(0.5, freq=30e9),
        "sinusoidal": core.SinusoidalSource(0.5, 3e9, 0.0),
        "ricker": core.RickerWavelet(1.0, 1e9),
        "modulated": core.ModulatedGaussian(1.0, 1e9, sigma=1 / (2 * math.pi * 0.5e9)),
        "chirp": core.Chirp(1.0, 1e8, 1e9, 1e-10, 1.2e-9),
        "uwb": core.UWBPulse(1.0, 1e9, 3e9),
    }
    for name, w in waves.items():
        np.testing.assert_array_equal(w.build(600, dt, torch.device("cpu")).numpy(), F[f"wave_{name}"], err_msg=name)


def test_oracle_das_equals_reference(F):
    S, ants = F["das_traces"], [tuple(a) for a in F["das_antennas"]]
    dt = float(F["das_dt"])
    np.testing.assert_array_equal(das_ref.das_backprojection(S, ants, 30, 30, 1e-3, dt, 200), F["das_image"])
    np.testing.assert_array_equal(das_ref.das_backprojection(S, ants, 30, 30, 1e-3, dt, 200, (20, 25)),
                                  F["das_image_20x25"])
    S9, ants9 = F["das09_traces"], [tuple(a) for a in F["das09_antennas"]]
    dt9 = float(F["das09_dt"])
    gate = max(0, int(0.9e-9 / dt9))
    np.testing.assert_array_equal(das_ref.das_backprojection(S9, ants9, 40, 40, 1e-3, dt9, 500, gate_start=gate),
                                  F["das09_image"])
    # the per-pixel brute force agrees too (pins the NumPy restatement independently)
    np.testing.assert_allclose(das_ref.das_brute(S, ants, 30, 30, 1e-3, dt, 200), F["das_image"], rtol=1e-12)


def test_antenna_positions_equal_reference(F):
    g = core.YeeGrid(30, 30, 1e-3, 1e-3, device="cpu")
    assert core.get_antenna_positions(6, 12, (15, 15), g) == [tuple(int(v) for v in a) for a in F["das_antennas"]]
