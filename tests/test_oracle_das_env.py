"""The envelope / layered-medium DAS oracle, pinned.

* bit for bit against the reference's own ``AntennaRing.compute_das_image``
  (fixtures written by ``tests/golden/make_golden.py``), free space with an
  even record length and a layered head with an odd one;
* the analytic-signal kernel: its closed form equals the FFT form the product
  uses, and convolving with it reproduces SciPy's Hilbert transform;
* the envelope of a cosine at a bin frequency is its amplitude;
* travel times: free space is distance / c; a ray through the centre of one
  layer spends exactly its diameter there;
* the vectorised accumulation equals a per-pixel loop.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import das_env_ref as E

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures.npz"


def _head(v):
    return {"cx_cells": v[0], "cy_cells": v[1], "layers": [(v[2 + 2 * k], v[3 + 2 * k]) for k in range((len(v) - 2) // 2)]}


@pytest.mark.parametrize("tag", ["env_free", "env_head"])
def test_oracle_equals_reference_fixture(tag):
    f = np.load(FIX)
    S, pos, dt, img = f[f"{tag}_traces"], [tuple(p) for p in f[f"{tag}_positions"]], float(f[f"{tag}_dt"]), \
        f[f"{tag}_image"]
    geom = _head(f["env_head_geometry"]) if tag == "env_head" else None
    got = E.compute_das_image(S, pos, 30, 30, 1e-3, dt, S.shape[-1], image_size=img.shape[0], head_geometry=geom)
    np.testing.assert_array_equal(got, img)


@pytest.mark.parametrize("n", [7, 8, 300, 301])
def test_hilbert_kernel_closed_form_and_fft_form(n):
    from scipy.signal import hilbert
    from waveforge_amd.antenna import _hilbert_kernel_np

    k = E.hilbert_kernel(n)
    np.testing.assert_allclose(_hilbert_kernel_np(n), k, atol=1e-13)
    x = np.random.default_rng(n).standard_normal(n)
    y = np.array([sum(x[m] * k[(t - m) % n] for m in range(n)) for t in range(n)])
    np.testing.assert_allclose(y, hilbert(x).imag, atol=1e-11)


def test_envelope_of_bin_cosine_is_its_amplitude():
    n, f = 256, 12
    t = np.arange(n)
    s = (2.5 * np.cos(2 * np.pi * f * t / n)).astype(np.float32)[None, None, :]
    env = E.envelope(s)
    np.testing.assert_allclose(env[0, 0], 2.5, rtol=1e-5)


def test_travel_time_closed_forms():
    dt = 2e-12
    # free space: 30 mm at 3e8 m/s
    assert E.travel_time_samples(0.0, 0.0, 30.0, 0.0, 100.0, 100.0, [], dt) == pytest.approx(30e-3 / 3e8 / dt, rel=1e-14)
    # one layer of radius 10 mm, eps 9, centred on the ray: 20 mm at c/3, 20 mm at c
    t = E.travel_time_samples(-20.0, 0.0, 20.0, 0.0, 0.0, 0.0, [(10.0, 9.0)], dt)
    assert t == pytest.approx((20e-3 / (3e8 / 3) + 20e-3 / 3e8) / dt, rel=1e-12)
    # nested layers: inner layer's chord is not counted twice
    t2 = E.travel_time_samples(-20.0, 0.0, 20.0, 0.0, 0.0, 0.0, [(10.0, 4.0), (5.0, 16.0)], dt)
    assert t2 == pytest.approx((10e-3 / (3e8 / 2) + 10e-3 / (3e8 / 4) + 20e-3 / 3e8) / dt, rel=1e-12)
    # a ray that misses the circle
    assert E.ray_circle_chord(-20.0, 15.0, 20.0, 15.0, 0.0, 0.0, 10.0) == 0.0


def test_vectorised_accumulation_equals_per_pixel_loop():
    rng = np.random.default_rng(2)
    env = np.abs(rng.standard_normal((4, 4, 60))).astype(np.float32)
    tt = (rng.uniform(-3, 33, (4, 9, 7))).astype(np.float32)
    np.testing.assert_array_equal(E.das_from_envelope(env, tt, 60), E.das_brute(env, tt, 60))
