"""Imaging samples in the dataset generator's ``.npz`` format (2D).

``BrainDatasetGenerator`` writes one ``sample_NNNNNN.npz`` per phantom
(``src/datasets/generator.py:356-428``): the MIMO signal matrices of the
phantom with a bleed and of its bleed-free twin, their difference, the
envelope DAS image of the difference (layered-medium delays from the head
geometry), labels, geometry, simulation parameters and a throughput figure;
string fields are stored as 0-d unicode arrays, so the file loads with
``numpy.load(..., allow_pickle=False)``.

``make_sample`` produces the same keys and dtypes for a 2D head phantom
(``workloads.head_phantom``, whose clot plays the bleed): both phantoms' TX
runs advance together in one batched launch sequence, and the image comes
from ``antenna.EnvelopeDAS``. The phantom's elliptical layers enter the
ray-circle delay model as circles of equal area. Shapes that are 3D in the
reference are 2D here with a unit third axis (``grid_shape`` ``[N, N, 1]``,
centres ``(i, j, 0)``).
"""

from __future__ import annotations

import math
import time
from pathlib import Path

import numpy as np
import torch

from .antenna import EnvelopeDAS
from .core.materials import coefficients_from_maps
from .core.sources import RickerWavelet
from .mimo import BatchedRuns, mimo_specs
from .workloads import HEAD_MEDIA, head_phantom, imaging_dt

STRING_FIELDS = ("bleed_type", "bleed_age", "phantom_id")


def head_layers(n: int, dx: float = 1e-3):
    """Equal-area circle radii (cells) of the head phantom's outer boundary and
    skull inner boundary, with the media inside each (outermost first)."""
    a_out, b_out = 0.36 * n, 0.30 * n
    t = 7e-3 / dx
    r_out = math.sqrt(a_out * b_out)
    r_in = math.sqrt((a_out - t) * (b_out - t))
    return [(r_out, HEAD_MEDIA["skull"][0]), (r_in, HEAD_MEDIA["brain"][0])]


def make_sample(n: int, n_elements: int, n_steps: int, *, seed: int, device="cuda", image_size: int = 40,
                phantom_id: str = "train", dx: float = 1e-3, depth: int = 6) -> dict:
    """One sample: a head phantom with a clot (label 1) and its clot-free twin."""
    dev = torch.device(device)
    dt = imaging_dt(dx)
    ph = head_phantom(n, n_elements, seed=seed, clot=True, dx=dx)
    ref = head_phantom(n, n_elements, seed=seed, clot=False, dx=dx)
    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9).build(n_steps, dt, dev, torch.float32)
    specs = []
    for p, tag in ((ph, "total"), (ref, "reference")):
        ca, cb = coefficients_from_maps(p.eps, p.sigma, dt, device=dev)
        specs += mimo_specs(ca, cb, p.antennas, wave, n_steps, tag=tag)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    br = BatchedRuns(specs, dx=dx, dy=dx, dt=dt, device=dev, depth=depth)
    br.advance(n_steps)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    E = len(ph.antennas)
    total = torch.stack([br.traces(k) for k in range(E)])
    refs = torch.stack([br.traces(E + k) for k in range(E)])
    scat = total - refs
    c = (n - 1) / 2.0
    geom = {"cx_cells": c, "cy_cells": c, "layers": head_layers(n, dx)}
    image = EnvelopeDAS(ph.antennas, (n, n), dx, dt, image_size, geom, device=dev)(scat)
    ci, cj, rad = ph.clot
    r_mm = rad * dx * 1e3
    a_out, b_out = 0.36 * n, 0.30 * n
    t_skull = 7e-3 / dx
    return {
        "signals_total": total.cpu().numpy(),
        "signals_reference": refs.cpu().numpy(),
        "signals_scattered": scat.cpu().numpy(),
        "das_image": image.cpu().numpy(),
        "label": np.int32(1),
        "bleed_type": "clot",
        "bleed_age": "none",
        "bleed_center_cells": np.array([int(round(ci)), int(round(cj)), 0], dtype=np.int32),
        "bleed_center_mm": np.array([int(round(ci)), int(round(cj)), 0], dtype=np.float32) * float(dx * 1e3),
        "bleed_radius_cells": np.int32(int(round(rad))),
        "bleed_radius_mm": np.float32(r_mm),
        "bleed_volume_ml": np.float32((4 / 3) * math.pi * r_mm ** 3 / 1e3),
        "freq_hz": np.float32(1e9),
        "freq_low_hz": np.float32(1e9),
        "freq_high_hz": np.float32(1e9),
        "uwb_mode": np.bool_(False),
        "dx_mm": np.float32(dx * 1e3),
        "grid_shape": np.array([n, n, 1], dtype=np.int32),
        "n_tx": np.int32(E),
        "n_rx": np.int32(E),
        "n_steps": np.int32(n_steps),
        "dt_s": np.float64(dt),
        "phantom_id": phantom_id,
        "phantom_seed": np.int32(seed),
        "phantom_skull_inner_r": np.int32(int(round(math.sqrt((a_out - t_skull) * (b_out - t_skull))))),
        "phantom_gray_r": np.int32(int(round(math.sqrt((a_out - t_skull) * (b_out - t_skull))))),
        "phantom_scalp_outer_r": np.int32(int(round(math.sqrt(a_out * b_out)))),
        "_mcells_per_second": np.float32(br.n_cells * n_steps / elapsed / 1e6),
    }


def save_sample(sample: dict, out_dir, sample_idx: int) -> Path:
    """``sample_{idx:06d}.npz`` with the generator's layout (``generator.py:418-428``)."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    path = out / f"sample_{sample_idx:06d}.npz"
    arrays = {k: v for k, v in sample.items() if isinstance(v, (np.ndarray, np.generic))}
    for key in STRING_FIELDS:
        arrays[key] = np.array(sample[key])
    np.savez_compressed(path, **arrays)
    return path


def load_sample(path) -> dict:
    """Read a sample back (no pickles: ``allow_pickle=False``); string fields as ``str``."""
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    for key in STRING_FIELDS:
        if key in d:
            d[key] = str(d[key])
    return d
