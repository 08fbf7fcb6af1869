"""Seeded synthetic workloads (the benchmark configurations), built on ``core``.

* ``slab_256``            — 256x256 slab run: vacuum, lossless eps_r=4 slab in
                            columns i=150..169, Gaussian-derivative Hz pulse at
                            (64, 128), detectors (100,128), (200,128), 1000 steps,
                            dt = 0.99 dx/(c sqrt 2). The reference's example 02
                            scenario (``examples/02_dielectric_slab.py``).
* ``blob_medium``         — N x N (4096 default) medium: eps_r = 1 + sum of 64
                            seeded Gaussian blobs (peak eps_r U[1,60], radius
                            U[20,200] cells), clipped to 60, sigma = 0.02 (eps_r-1);
                            1 GHz modulated-Gaussian Hz source at the centre, 64
                            detectors on a circle of radius 1800 cells.
* ``head_phantom``        — 2D head: coupling medium (20, 0.1 S/m), elliptical
                            7 mm skull ring (12, 0.2), brain (45, 0.8), optional
                            clot radius 5-15 mm (60, 1.2) at a random interior
                            spot, with a ring array around it.

Blob combination rule (stated here because the workload definition leaves it
open): each blob adds ``(peak - 1) * exp(-r^2 / (2 R^2))`` over the cells
within ``4 R`` of its centre (beyond that the term is < 3.4e-4 of the peak and
is dropped), in fp64; the sum is clipped to ``[1, 60]``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .core import (
    DetectorSet,
    FDTD2D,
    FieldSet,
    GaussianDerivative,
    ModulatedGaussian,
    MurABC,
    PointSource,
    RickerWavelet,
    SourceCollection,
    YeeGrid,
    coefficients_from_maps,
    get_antenna_positions,
)
from .core.materials import MaterialMap, Material

C0 = 299_792_458.0


def blob_eps_map(n: int, n_blobs: int = 64, seed: int = 0) -> np.ndarray:
    """eps_r map (n x n, fp64) of the blob medium."""
    rng = np.random.default_rng(seed)
    eps = np.ones((n, n), dtype=np.float64)
    cx = rng.uniform(0, n, n_blobs)
    cy = rng.uniform(0, n, n_blobs)
    peak = rng.uniform(1.0, 60.0, n_blobs)
    rad = rng.uniform(20.0, 200.0, n_blobs)
    for x0, y0, p, R in zip(cx, cy, peak, rad):
        r4 = 4.0 * R
        i0, i1 = max(0, int(math.floor(x0 - r4))), min(n, int(math.ceil(x0 + r4)) + 1)
        j0, j1 = max(0, int(math.floor(y0 - r4))), min(n, int(math.ceil(y0 + r4)) + 1)
        if i0 >= i1 or j0 >= j1:
            continue
        ii = np.arange(i0, i1, dtype=np.float64)[:, None]
        jj = np.arange(j0, j1, dtype=np.float64)[None, :]
        r2 = (ii - x0) ** 2 + (jj - y0) ** 2
        eps[i0:i1, j0:j1] += np.where(r2 <= r4 * r4, (p - 1.0) * np.exp(-r2 / (2.0 * R * R)), 0.0)
    return np.clip(eps, 1.0, 60.0)


@dataclass
class Simulation:
    """A ready-to-run FDTD2D with its detectors and bookkeeping."""

    sim: FDTD2D
    detectors: DetectorSet | None
    n_steps: int
    grid: YeeGrid
    Ca: torch.Tensor | None
    Cb: torch.Tensor | None
    source_cell: tuple[int, int]
    description: str


def slab_256(device="cuda", n: int = 256, n_steps: int = 1000, **sim_kw) -> Simulation:
    dx = 1e-3
    grid = YeeGrid(n, n, dx, dx, courant=0.99, device=device)  # dt = 0.99 dx / (c sqrt 2)
    dt = grid.dt
    fields = FieldSet(grid)
    mm = MaterialMap(grid)
    mm.add_rectangle((150, 169), (0, n - 1), Material("slab", eps_r=4.0, sigma=0.0))
    Ca, Cb = mm.build()
    src = PointSource(GaussianDerivative.in_steps(dt, 60, 15), 64, 128, "Hz", grid=grid, N_steps=n_steps)
    det = DetectorSet([(100, 128), (200, 128)], n_steps, grid=grid)
    sim = FDTD2D(grid, fields, MurABC(grid, fields.Hz), SourceCollection([src]), Ca=Ca, Cb=Cb, detectors=det,
                 **sim_kw)
    return Simulation(sim, det, n_steps, grid, Ca, Cb, (64, 128), f"slab {n}x{n}, {n_steps} steps")


def blob_medium(device="cuda", n: int = 4096, n_steps: int = 4000, seed: int = 0, n_det: int = 64,
                det_radius: float | None = None, inner_rings=(), **sim_kw) -> Simulation:
    dx = 1e-3
    grid = YeeGrid(n, n, dx, dx, device=device)
    eps = blob_eps_map(n, seed=seed)
    sigma = 0.02 * (eps - 1.0)
    Ca, Cb = coefficients_from_maps(eps, sigma, grid.dt, device=grid.device)
    fields = FieldSet(grid)
    wave = ModulatedGaussian(fc=1e9, sigma=1.0 / (2.0 * math.pi * 0.5e9))
    c = n // 2
    src = PointSource(wave, c, c, "Hz", grid=grid, N_steps=n_steps)
    r = det_radius if det_radius is not None else 1800.0 * n / 4096.0
    cells = [(int(round(c + r * math.cos(2 * math.pi * k / n_det))), int(round(c + r * math.sin(2 * math.pi * k / n_det))))
             for k in range(n_det)]
    for m, rr in inner_rings:  # further (count, radius) rings after the main one
        cells += [(int(round(c + rr * math.cos(2 * math.pi * k / m))), int(round(c + rr * math.sin(2 * math.pi * k / m))))
                  for k in range(m)]
    det = DetectorSet(cells, n_steps, grid=grid)
    sim = FDTD2D(grid, fields, MurABC(grid, fields.Hz), SourceCollection([src]), Ca=Ca, Cb=Cb, detectors=det,
                 **sim_kw)
    return Simulation(sim, det, n_steps, grid, Ca, Cb, (c, c), f"blob medium {n}x{n}, {n_steps} steps")


# ---------------------------------------------------------------------------
# Head phantoms (batched imaging / serving / adjoint workloads)
# ---------------------------------------------------------------------------

HEAD_MEDIA = {
    "coupling": (20.0, 0.1),
    "skull": (12.0, 0.2),
    "brain": (45.0, 0.8),
    "clot": (60.0, 1.2),
}


@dataclass
class HeadPhantom:
    n: int
    eps: np.ndarray
    sigma: np.ndarray
    clot: tuple[float, float, float] | None  # (ci, cj, radius) in cells
    antennas: list[tuple[int, int]]


def head_phantom(n: int, n_elements: int = 16, *, seed: int = 0, clot: bool = True, dx: float = 1e-3) -> HeadPhantom:
    """Elliptical head in a coupling medium with a ``n_elements`` ring array.

    Head semi-axes 0.36 n and 0.30 n cells, a 7 mm skull ring, brain inside;
    the clot (radius U[5, 15] mm) sits at a random spot fully inside the brain.
    The array ring has radius 0.45 n, clamped to ``[1, n-2]``.
    """
    rng = np.random.default_rng(seed)
    I = np.arange(n, dtype=np.float64)[:, None]
    J = np.arange(n, dtype=np.float64)[None, :]
    c = (n - 1) / 2.0
    a_out, b_out = 0.36 * n, 0.30 * n
    t_skull = 7e-3 / dx
    a_in, b_in = a_out - t_skull, b_out - t_skull
    r_out = ((I - c) / a_out) ** 2 + ((J - c) / b_out) ** 2
    r_in = ((I - c) / a_in) ** 2 + ((J - c) / b_in) ** 2
    eps = np.full((n, n), HEAD_MEDIA["coupling"][0])
    sig = np.full((n, n), HEAD_MEDIA["coupling"][1])
    skull = r_out <= 1.0
    eps[skull], sig[skull] = HEAD_MEDIA["skull"]
    brain = r_in <= 1.0
    eps[brain], sig[brain] = HEAD_MEDIA["brain"]
    clot_geom = None
    rad = rng.uniform(5e-3, 15e-3) / dx
    ang = rng.uniform(0, 2 * math.pi)
    frac = rng.uniform(0.0, 1.0)
    max_off_a = max(0.0, a_in - rad - 1)
    max_off_b = max(0.0, b_in - rad - 1)
    ci = c + frac * max_off_a * math.cos(ang)
    cj = c + frac * max_off_b * math.sin(ang)
    if clot:
        m = ((I - ci) ** 2 + (J - cj) ** 2) <= rad * rad
        eps[m], sig[m] = HEAD_MEDIA["clot"]
        clot_geom = (ci, cj, rad)
    g = type("G", (), {"Nx": n, "Ny": n})
    ants = get_antenna_positions(n_elements, 0.45 * n, (c, c), g, clamp="inner")
    return HeadPhantom(n, eps, sig, clot_geom, ants)


def ricker_1ghz() -> RickerWavelet:
    """1 GHz Ricker pulse with t0 = 1.5/f (``examples/brain_mimo_imaging.py:184-185``)."""
    return RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9)


def phantom_set(n_phantoms: int = 32, sizes=(150, 200, 256, 384), n_elements: int = 16, seed: int = 0):
    """Seeded head phantoms cycling through ``sizes`` (8 of each for 32): the
    batched-imaging workload. Returns ``[(eps_r, sigma, antennas), ...]``."""
    out = []
    for p in range(n_phantoms):
        ph = head_phantom(sizes[p % len(sizes)], n_elements, seed=seed * 1000 + p, clot=True)
        out.append((ph.eps, ph.sigma, ph.antennas))
    return out


def imaging_dt(dx: float = 1e-3) -> float:
    """dt = 0.99 dx / (c sqrt 2), the time step of the imaging workloads, computed
    as ``YeeGrid(..., courant=0.99)`` computes it."""
    from .core.grid import compute_stable_dt

    return compute_stable_dt(dx, dx, courant=0.99)


# ---------------------------------------------------------------------------
# Adjoint workload: trace-misfit gradient on a head phantom
# ---------------------------------------------------------------------------

@dataclass
class AdjointCase:
    problem: object             # adjoint.TraceProblem
    eps: torch.Tensor           # clot-free eps_r, fp64 on the device (the model being differentiated)
    sigma: torch.Tensor
    d_obs: torch.Tensor         # fp32 traces of the clot phantom (the "observed" data)
    phantom: HeadPhantom        # the clot phantom


def adjoint_case(device="cuda", n: int = 256, n_steps: int = 3000, seed: int = 0, checkpoint: int = 100,
                 depth: int = 6, n_elements: int = 16) -> AdjointCase:
    """256^2 head phantom, 16 TX x 16 RX, 3000 steps: d_obs from the clot
    phantom, the model is its clot-free twin (same seed, same head), loss
    0.5 ||d_sim - d_obs||^2; the gradient keeps the full state history when
    it fits (``TraceProblem.history``), else checkpoints every ``checkpoint``
    steps."""
    from .adjoint import TraceProblem, coefficients

    dev = torch.device(device)
    dt = imaging_dt()
    ph = head_phantom(n, n_elements, seed=seed, clot=True)
    free = head_phantom(n, n_elements, seed=seed, clot=False)
    wave = ricker_1ghz().build(n_steps + 1, dt, dev)
    prob = TraceProblem(n, n, 1e-3, dt, list(ph.antennas), list(ph.antennas), wave, n_steps,
                        checkpoint=checkpoint, depth=depth)
    e_c = torch.from_numpy(ph.eps).to(dev)
    s_c = torch.from_numpy(ph.sigma).to(dev)
    ca, cb = coefficients(e_c, s_c, dt)
    prob.history = "checkpoint"  # the data need only the traces
    d_obs, _ = prob.forward(ca.float().contiguous(), cb.float().contiguous())
    prob.history = "auto"
    return AdjointCase(prob, torch.from_numpy(free.eps).to(dev), torch.from_numpy(free.sigma).to(dev), d_obs, ph)
