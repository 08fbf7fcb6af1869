"""NumPy restatement of the reference DAS backprojection (the oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``das_backprojection`` follows ``examples/brain_mimo_imaging.py:211-266``:
pixel pitch ``Nx*dx/Nx_img``, fp64 distances, ``np.round`` (half-even) of
``((d_tx + d_rx) / v) / dt`` cast to int32, ``clip`` to ``[0, T-1]``, fp64
accumulation in tx-major, rx-minor order, squared at the end. The optional
``gate_start`` zeroes ``S[:, :, :gate_start]`` first
(``examples/09_brain_clot_dataset_sample.py:176-178``) and
``out_of_range="zero"`` drops samples whose index falls outside
``[0, T)`` instead of clipping (``src/datasets/brain/antenna.py:333-336``).
"""

from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0


def das_backprojection(
    S: np.ndarray,
    antenna_positions,
    Nx: int,
    Ny: int,
    dx: float,
    dt: float,
    n_steps: int,
    image_shape=None,
    v_medium: float | None = None,
    gate_start: int = 0,
    out_of_range: str = "clip",
    dy: float | None = None,
) -> np.ndarray:
    if image_shape is None:
        Nx_img, Ny_img = Nx, Ny
    else:
        Nx_img, Ny_img = image_shape
    if v_medium is None:
        v_medium = C0 / math.sqrt(40.0)
    n_tx, n_rx = S.shape[0], S.shape[1]
    if gate_start > 0:
        S = S.copy()
        S[:, :, :gate_start] = 0.0
    ant_x = np.array([p[0] * dx for p in antenna_positions])
    ant_y = np.array([p[1] * dx for p in antenna_positions])
    px = np.arange(Nx_img) * (Nx * dx / Nx_img)
    py = np.arange(Ny_img) * (Ny * dx / Ny_img)
    PX, PY = np.meshgrid(px, py, indexing="ij")
    image = np.zeros((Nx_img, Ny_img), dtype=np.float64)
    for tx in range(n_tx):
        for rx in range(n_rx):
            d_tx = np.sqrt((PX - ant_x[tx]) ** 2 + (PY - ant_y[tx]) ** 2)
            d_rx = np.sqrt((PX - ant_x[rx]) ** 2 + (PY - ant_y[rx]) ** 2)
            tau = (d_tx + d_rx) / v_medium
            n_tau = np.round(tau / dt).astype(np.int32)
            if out_of_range == "clip":
                n_tau = np.clip(n_tau, 0, n_steps - 1)
                image += S[tx, rx, n_tau]
            else:
                valid = (n_tau >= 0) & (n_tau < n_steps)
                n_c = np.clip(n_tau, 0, n_steps - 1)
                image += np.where(valid, S[tx, rx, n_c], 0.0)
    return image ** 2


def das_brute(S, antenna_positions, Nx, Ny, dx, dt, n_steps, image_shape=None,
              v_medium=None, gate_start=0, out_of_range="clip"):
    """Per-pixel Python-loop restatement used to pin ``das_backprojection`` on
    tiny images (independent of NumPy broadcasting)."""
    if image_shape is None:
        Nx_img, Ny_img = Nx, Ny
    else:
        Nx_img, Ny_img = image_shape
    if v_medium is None:
        v_medium = C0 / math.sqrt(40.0)
    pitch_x = Nx * dx / Nx_img
    pitch_y = Ny * dx / Ny_img
    img = np.zeros((Nx_img, Ny_img))
    for u in range(Nx_img):
        for v in range(Ny_img):
            x, y = u * pitch_x, v * pitch_y
            acc = 0.0
            for tx, (ti, tj) in enumerate(antenna_positions):
                dtx = math.sqrt((x - ti * dx) ** 2 + (y - tj * dx) ** 2)
                for rx, (ri, rj) in enumerate(antenna_positions):
                    drx = math.sqrt((x - ri * dx) ** 2 + (y - rj * dx) ** 2)
                    n = int(round(((dtx + drx) / v_medium) / dt))
                    if out_of_range == "clip":
                        n = min(max(n, 0), n_steps - 1)
                    elif not (0 <= n < n_steps):
                        continue
                    if n < gate_start:
                        continue
                    acc += float(S[tx, rx, n])
            img[u, v] = acc * acc
    return img
