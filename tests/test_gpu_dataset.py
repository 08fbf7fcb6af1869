"""A generated imaging sample: batched runs equal standalone runs, the image is
the envelope DAS of the scattered signals, and the file round-trips."""

import numpy as np
import pytest
import torch

from oracle import das_env_ref as E
from test_dataset_format import REFERENCE_KEYS

pytestmark = pytest.mark.gpu


def test_sample_contents_and_file(device, tmp_path):
    from waveforge_amd import workloads
    from waveforge_amd.core import YeeGrid
    from waveforge_amd.core.materials import coefficients_from_maps
    from waveforge_amd.dataset import head_layers, load_sample, make_sample, save_sample
    from waveforge_amd.mimo import run_mimo

    n, E_, steps = 96, 8, 600
    s = make_sample(n, E_, steps, seed=11, device=device, image_size=32)
    assert set(s) == REFERENCE_KEYS
    dt = workloads.imaging_dt()
    ph = workloads.head_phantom(n, E_, seed=11, clot=True)
    g = YeeGrid(n, n, 1e-3, 1e-3, courant=0.99, device=device)
    ca, cb = coefficients_from_maps(ph.eps, ph.sigma, dt, device=device)
    alone = run_mimo(g, ca, cb, ph.antennas, steps)
    np.testing.assert_array_equal(s["signals_total"], alone)
    np.testing.assert_array_equal(s["signals_scattered"], s["signals_total"] - s["signals_reference"])
    c = (n - 1) / 2.0
    geom = {"cx_cells": c, "cy_cells": c, "layers": head_layers(n)}
    ref = E.compute_das_image(s["signals_scattered"], ph.antennas, n, n, 1e-3, dt, steps, image_size=32,
                              head_geometry=geom)
    rel = float(np.linalg.norm(s["das_image"].astype(np.float64) - ref) / np.linalg.norm(ref))
    print(f"[dataset] das_image vs oracle rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-5 and np.argmax(s["das_image"]) == np.argmax(ref)
    p = save_sample(s, tmp_path, 0)
    d = load_sample(p)
    np.testing.assert_array_equal(d["signals_scattered"], s["signals_scattered"])
    np.testing.assert_array_equal(d["das_image"], s["das_image"])
    assert d["bleed_type"] == "clot" and int(d["label"]) == 1
