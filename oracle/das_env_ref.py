"""CPU restatement of the envelope / layered-medium DAS image (test infrastructure).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``AntennaRing.compute_das_image`` (``src/datasets/brain/antenna.py:262-341``):

* envelope ``|hilbert(signals, axis=-1)|`` cast to fp32 (``:294``; SciPy's
  FFT analytic signal, computed in single precision for fp32 input);
* pixel coordinates ``linspace(0, (N-1)*dx*1e3, image_size)`` in mm (``:288-289``);
* a fp32 table ``tt[ant, iy, ix]`` of one-way travel times in samples: the
  layered ray-circle model ``_travel_time_samples`` (``:239-260``, chords
  ``_ray_sphere_chord`` ``:218-237``) when a head geometry is given, else the
  free-space time ``(d*1e-3)/C0/dt`` with ``C0 = 3e8`` (``:314-321``);
* accumulation in fp64 over tx-major, rx-minor pairs of
  ``env[tx, rx][rint(tt[tx] + tt[rx])]``, samples outside ``[0, N_steps)``
  contributing 0 (``:324-331``), and ``(img**2).astype(float32)`` (``:333``).

The travel-time functions below are written in plain Python floats with the
same expression order, so their fp64 values (hence the fp32 table) are the
reference's.
"""

from __future__ import annotations

import math

import numpy as np

C0_RAY = 3e8  # the travel-time model's speed of light (antenna.py:249, :318)


def ray_circle_chord(ax, ay, bx, by, cx, cy, r) -> float:
    """Length of the part of segment a->b inside the circle (centre c, radius r) (``antenna.py:218-237``)."""
    dx_ = bx - ax
    dy_ = by - ay
    seg_len = math.sqrt(dx_ ** 2 + dy_ ** 2) + 1e-30
    ux, uy = dx_ / seg_len, dy_ / seg_len
    ox, oy = ax - cx, ay - cy
    b_c = ox * ux + oy * uy
    c_c = ox ** 2 + oy ** 2 - r ** 2
    disc = b_c ** 2 - c_c
    if disc <= 0:
        return 0.0
    sq = math.sqrt(disc)
    t0 = max(0.0, min(seg_len, -b_c - sq))
    t1 = max(0.0, min(seg_len, -b_c + sq))
    return max(0.0, t1 - t0)


def travel_time_samples(ax, ay, bx, by, cx, cy, layers, dt) -> float:
    """One-way travel time in samples through concentric layers, outermost first (``antenna.py:239-260``)."""
    total_d = math.sqrt((bx - ax) ** 2 + (by - ay) ** 2) + 1e-30
    prev = 0.0
    t = 0.0
    for idx, (r, eps) in enumerate(layers):
        co = ray_circle_chord(ax, ay, bx, by, cx, cy, r)
        if idx + 1 < len(layers):
            ci = ray_circle_chord(ax, ay, bx, by, cx, cy, layers[idx + 1][0])
        else:
            ci = 0.0
        sc = max(0.0, co - ci)
        t += (sc * 1e-3) / (C0_RAY / math.sqrt(eps))
        prev += sc
    t += (max(0.0, total_d - prev) * 1e-3) / C0_RAY
    return t / dt


def pixel_axes_mm(Nx, Ny, dx, image_size):
    return (np.linspace(0, (Nx - 1) * dx * 1e3, image_size), np.linspace(0, (Ny - 1) * dx * 1e3, image_size))


def travel_time_table(positions, Nx, Ny, dx, dt, image_size, head_geometry=None) -> np.ndarray:
    """``tt[ant, iy, ix]`` fp32 (``antenna.py:296-321``)."""
    x_px, y_px = pixel_axes_mm(Nx, Ny, dx, image_size)
    ants = [(i * dx * 1e3, j * dx * 1e3) for i, j in positions]
    tt = np.zeros((len(ants), image_size, image_size), dtype=np.float32)
    if head_geometry is not None:
        cx = head_geometry["cx_cells"] * dx * 1e3
        cy = head_geometry["cy_cells"] * dx * 1e3
        layers = [(r * dx * 1e3, eps) for r, eps in head_geometry["layers"]]
        for a, (ax, ay) in enumerate(ants):
            for iy, py in enumerate(y_px):
                for ix, px in enumerate(x_px):
                    tt[a, iy, ix] = travel_time_samples(ax, ay, px, py, cx, cy, layers, dt)
    else:
        for a, (ax, ay) in enumerate(ants):
            for iy, py in enumerate(y_px):
                for ix, px in enumerate(x_px):
                    d = math.sqrt((px - ax) ** 2 + (py - ay) ** 2)
                    tt[a, iy, ix] = (d * 1e-3) / C0_RAY / dt
    return tt


def envelope(signals: np.ndarray) -> np.ndarray:
    """``|hilbert(signals, axis=-1)|`` as fp32 (``antenna.py:294``)."""
    from scipy.signal import hilbert

    return np.abs(hilbert(signals, axis=-1)).astype(np.float32)


def das_from_envelope(env: np.ndarray, tt: np.ndarray, n_steps: int) -> np.ndarray:
    """The accumulation and power step (``antenna.py:324-333``)."""
    n = tt.shape[0]
    img = np.zeros(tt.shape[1:], dtype=np.float64)
    for tx in range(n):
        for rx in range(n):
            delay = np.rint(tt[tx] + tt[rx]).astype(np.int32)
            valid = (delay >= 0) & (delay < n_steps)
            delay = np.clip(delay, 0, n_steps - 1)
            img += np.where(valid, env[tx, rx][delay], 0.0)
    return (img ** 2).astype(np.float32)


def compute_das_image(signals, positions, Nx, Ny, dx, dt, n_steps, image_size=40, head_geometry=None):
    """The whole of ``AntennaRing.compute_das_image`` for 2D positions ``(i, j)``."""
    env = envelope(signals)
    tt = travel_time_table(positions, Nx, Ny, dx, dt, image_size, head_geometry)
    return das_from_envelope(env, tt, n_steps)


def das_brute(env, tt, n_steps) -> np.ndarray:
    """Per-pixel loop form of ``das_from_envelope`` (independent check)."""
    n, h, w = tt.shape
    out = np.zeros((h, w), dtype=np.float64)
    for iy in range(h):
        for ix in range(w):
            acc = 0.0
            for tx in range(n):
                for rx in range(n):
                    k = int(np.rint(np.float32(tt[tx, iy, ix]) + np.float32(tt[rx, iy, ix])))
                    if 0 <= k < n_steps:
                        acc += float(env[tx, rx, k])
            out[iy, ix] = acc
    return (out ** 2).astype(np.float32)


def hilbert_kernel(n: int) -> np.ndarray:
    """Imaginary part of the impulse response of SciPy's analytic-signal
    filter for length ``n`` (the circular kernel whose convolution with x is
    Im(hilbert(x))), from its closed form: ``k[d] = (2/n) sum_{m=1}^{K} sin(2 pi m d / n)``
    with ``K = n/2 - 1`` (even n) or ``(n-1)/2`` (odd n)."""
    K = n // 2 - 1 if n % 2 == 0 else (n - 1) // 2
    d = np.arange(n, dtype=np.float64)
    m = np.arange(1, K + 1, dtype=np.float64)
    if K <= 0:
        return np.zeros(n)
    return (2.0 / n) * np.sin(2.0 * np.pi * np.outer(d, m) / n).sum(axis=1)
