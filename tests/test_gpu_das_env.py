"""Envelope / layered-medium DAS on the GPU (``wf_das_envelope``).

Stage by stage against the oracle: the travel-time table bit for bit, the
accumulation bit for bit when given the oracle's envelope, the envelope to
fp32 rounding of SciPy's single-precision FFT; end to end against the
reference's own images (relative L2 < 1e-5, same peak pixel).
"""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import das_env_ref as E

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures.npz"


def _head(v):
    return {"cx_cells": v[0], "cy_cells": v[1], "layers": [(v[2 + 2 * k], v[3 + 2 * k]) for k in range((len(v) - 2) // 2)]}


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _peak(a):
    return np.unravel_index(int(np.argmax(a)), a.shape)


@pytest.mark.parametrize("tag", ["env_free", "env_head"])
def test_antenna_ring_matches_reference_fixture(device, tag):
    from waveforge_amd import core
    from waveforge_amd.antenna import AntennaRing, EnvelopeDAS

    f = np.load(FIX)
    S, img_ref = f[f"{tag}_traces"], f[f"{tag}_image"]
    geom = _head(f["env_head_geometry"]) if tag == "env_head" else None
    g = core.YeeGrid(30, 30, 1e-3, 1e-3, device=device)
    assert g.dt == float(f[f"{tag}_dt"])
    ring = AntennaRing(6, 13, 0, g, component="Hz")
    assert [(i, j) for i, j, _ in ring.positions] == [tuple(p) for p in f[f"{tag}_positions"]]
    ring.build_sources(core.GaussianPulse(1.0, sigma=5 * g.dt), 0, S.shape[-1])
    got = ring.compute_das_image(S, image_size=img_ref.shape[0], head_geometry=geom)
    assert got.dtype == np.float32 and got.shape == img_ref.shape
    e = _rel(got, img_ref)
    print(f"[das_env] {tag}: rel L2 vs reference {e:.2e}", flush=True)
    assert e < 1e-5 and _peak(got) == _peak(img_ref)
    # stage by stage
    pos = [(i, j) for i, j, _ in ring.positions]
    das = EnvelopeDAS(pos, (30, 30), 1e-3, g.dt, img_ref.shape[0], geom, device=device)
    tt = E.travel_time_table(pos, 30, 30, 1e-3, g.dt, img_ref.shape[0], geom)
    np.testing.assert_array_equal(das.tt.cpu().numpy(), tt)
    env_o = E.envelope(S)
    np.testing.assert_array_equal(das.image_from_envelope(torch.from_numpy(env_o)).cpu().numpy(),
                                  E.das_from_envelope(env_o, tt, S.shape[-1]))
    env = das.envelope(S).cpu().numpy()
    assert float(np.abs(env - env_o).max()) <= 2e-6 * float(env_o.max())


@pytest.mark.parametrize("n_t", [2000, 2001, 4999])
def test_envelope_long_records(device, n_t):
    from waveforge_amd.antenna import EnvelopeDAS

    rng = np.random.default_rng(n_t)
    t = np.arange(n_t)
    S = np.zeros((3, 3, n_t), np.float32)
    for a in range(3):
        for b in range(3):
            c = rng.uniform(200, n_t - 200)
            S[a, b] = (np.sin(2 * np.pi * rng.uniform(0.01, 0.2) * t) * np.exp(-((t - c) / 80.0) ** 2)
                       + 0.01 * rng.standard_normal(n_t)).astype(np.float32)
    das = EnvelopeDAS([(5, 5), (20, 30), (40, 8)], (48, 48), 1e-3, 2e-12, 16, device=device)
    env = das.envelope(S).cpu().numpy()
    ref = E.envelope(S)
    err = float(np.abs(env - ref).max() / ref.max())
    print(f"[das_env] n_t={n_t}: max |env - scipy| / max = {err:.2e}", flush=True)
    assert err < 2e-6


def test_head_phantom_scattered_image(device):
    """Scattered signals of a simulated head (clot minus clot-free twin), layered
    geometry from the phantom, GPU image vs the oracle's."""
    from waveforge_amd import workloads
    from waveforge_amd.antenna import EnvelopeDAS
    from waveforge_amd.core import YeeGrid
    from waveforge_amd.core.materials import coefficients_from_maps
    from waveforge_amd.mimo import run_mimo

    n, steps = 96, 600
    dt = workloads.imaging_dt()
    g = YeeGrid(n, n, 1e-3, 1e-3, courant=0.99, device=device)
    ph = workloads.head_phantom(n, 8, seed=3, clot=True)
    free = workloads.head_phantom(n, 8, seed=3, clot=False)
    sig = []
    for p in (ph, free):
        ca, cb = coefficients_from_maps(p.eps, p.sigma, dt, device=device)
        sig.append(run_mimo(g, ca, cb, p.antennas, steps))
    scat = sig[0] - sig[1]
    c = (n - 1) / 2.0
    geom = {"cx_cells": c, "cy_cells": c, "layers": [(0.33 * n, 12.0), (0.33 * n - 7, 45.0)]}
    das = EnvelopeDAS(ph.antennas, (n, n), 1e-3, dt, 48, geom, device=device)
    got = das(scat).cpu().numpy()
    ref = E.compute_das_image(scat, ph.antennas, n, n, 1e-3, dt, steps, image_size=48, head_geometry=geom)
    e = _rel(got, ref)
    print(f"[das_env] head phantom scattered image: rel L2 {e:.2e}, peak {_peak(got)} vs {_peak(ref)}", flush=True)
    assert e < 1e-5 and _peak(got) == _peak(ref)


def test_bad_shapes_raise(device):
    from waveforge_amd.antenna import EnvelopeDAS

    das = EnvelopeDAS([(5, 5), (20, 30)], (48, 48), 1e-3, 2e-12, 8, device=device)
    with pytest.raises(ValueError):
        das(np.zeros((3, 3, 10), np.float32))
