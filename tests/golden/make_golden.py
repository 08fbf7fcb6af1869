"""Write the reference fixtures (run in the build container only).

The reference ships no golden vectors. This script imports the reference's
own ``core`` package and imaging example functions from /root/reference (pure
Python) and records small input/output vectors for what the oracle's
mathematical checks cannot pin at the bit level: the exact eager-op
trajectory of ``FDTD2D.step`` (fields and detector traces), the ``Waveform.build``
tables, and the DAS images of ``examples/brain_mimo_imaging.py`` and
``examples/09_brain_clot_dataset_sample.py``, and the envelope / layered-medium
DAS of ``AntennaRing.compute_das_image``, and ``FDTD3D`` trajectories (lossy
``MaterialMap3D`` media, free space, and the Nz = 1 mode). Everything runs on the CPU in
fp32, as the reference's tests do. Output: ``tests/golden/reference_fixtures.npz``.
"""

from __future__ import annotations

import importlib.util
import math
import sys
from pathlib import Path

import numpy as np
import torch

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_fixtures.npz"


def load_module(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main() -> None:
    sys.path.insert(0, str(REF / "src"))
    from core.boundaries import MurABC
    from core.fdtd2d import FDTD2D
    from core.fields import FieldSet
    from core.grid import YeeGrid
    from core.materials import Material, MaterialMap
    from core import sources as S

    out = {}
    torch.manual_seed(0)

    # 1. slab trajectory: 48x48, courant 0.99, eps_r=4 slab, Gaussian Hz point source, 2 detectors, 150 steps
    g = YeeGrid(48, 48, 1e-3, 1e-3, courant=0.99, device="cpu")
    f = FieldSet(g)
    mm = MaterialMap(g)
    mm.add_rectangle((28, 33), (0, 47), Material("slab", 4.0, 0.0))
    Ca, Cb = mm.build()
    src = S.PointSource(S.GaussianPulse(1.0, sigma=8 * g.dt), 12, 24, "Hz", grid=g, N_steps=200)
    sim = FDTD2D(g, f, MurABC(g, f.Hz), S.SourceCollection([src]), Ca=Ca, Cb=Cb, n_check=50)
    tr = []
    for _ in range(150):
        sim.step()
        tr.append([float(f.Hz[20, 24, 0]), float(f.Hz[40, 24, 0])])
    out["slab_dt"] = np.array(g.dt)
    out["slab_Ca"], out["slab_Cb"] = Ca.numpy(), Cb.numpy()
    out["slab_wave"] = src._waveform_tensor.numpy()
    out["slab_Ex"], out["slab_Ey"], out["slab_Hz"] = f.Ex.numpy(), f.Ey.numpy(), f.Hz.numpy()
    out["slab_traces"] = np.array(tr, dtype=np.float32).T
    out["slab_last_field_max"] = np.array(sim.last_field_max)

    # 2. lossy media + mixed sources, free-space boundary values, 40x36, default courant, 80 steps
    g = YeeGrid(40, 36, 1e-3, 1e-3, device="cpu")
    f = FieldSet(g)
    mm = MaterialMap(g, default=Material("coupling", 20.0, 0.1))
    mm.add_circle((20, 18), 12, Material("brain", 45.0, 0.8))
    mm.add_circle((24, 15), 4, Material("clot", 60.0, 1.2))
    Ca, Cb = mm.build()
    srcs = [
        S.PointSource(S.RickerWavelet(1.0, 20e9), 20, 18, "Hz", grid=g, N_steps=100),
        S.PointSource(S.SinusoidalSource(0.3, 30e9), 3, 4, "Ex", grid=g, N_steps=100),
        S.LineSource(S.ModulatedGaussian(0.5, 25e9, sigma=2e-11), "x", 30, 5, 35, "Ey", grid=g, N_steps=100,
                     amplitude_profile=[0.5 + 0.02 * k for k in range(30)]),
        S.PlaneSource(S.GaussianPulse(0.2, sigma=5 * g.dt), "xz", 2, "Hz", grid=g, N_steps=100),
        S.PointSource(S.UWBPulse(1.0, 10e9, 40e9), 20, 18, "Hz", grid=g, N_steps=100),
    ]
    sim = FDTD2D(g, f, MurABC(g, f.Hz), S.SourceCollection(srcs), Ca=Ca, Cb=Cb, n_check=1000)
    for _ in range(80):
        sim.step()
    out["lossy_Ca"], out["lossy_Cb"] = Ca.numpy(), Cb.numpy()
    for k, s_ in enumerate(srcs):
        out[f"lossy_wave{k}"] = s_._waveform_tensor.numpy()
    out["lossy_Ex"], out["lossy_Ey"], out["lossy_Hz"] = f.Ex.numpy(), f.Ey.numpy(), f.Hz.numpy()

    # 3. waveform tables at dt = default 1 mm grid dt, 600 samples
    dt = YeeGrid(8, 8, 1e-3, 1e-3, device="cpu").dt
    waves = {
        "gaussian": S.GaussianPulse(1.3, sigma=7e-12),
        "gaussian_freq": S.GaussianPulse(0.5, freq=30e9),
        "sinusoidal": S.SinusoidalSource(0.5, 3e9, 0.0),
        "ricker": S.RickerWavelet(1.0, 1e9),
        "modulated": S.ModulatedGaussian(1.0, 1e9, sigma=1 / (2 * math.pi * 0.5e9)),
        "chirp": S.Chirp(1.0, 1e8, 1e9, 1e-10, 1.2e-9),
        "uwb": S.UWBPulse(1.0, 1e9, 3e9),
    }
    out["wave_dt"] = np.array(dt)
    for name, w in waves.items():
        out[f"wave_{name}"] = w.build(600, dt, torch.device("cpu"), torch.float32).numpy()

    # 4. DAS images (brain example; example 09 time-gated)
    rng = np.random.default_rng(5)
    bm = load_module("ref_brain_mimo", REF / "examples" / "brain_mimo_imaging.py")
    g = YeeGrid(30, 30, 1e-3, 1e-3, device="cpu")
    ants = bm.get_antenna_positions(6, 12, (15, 15), g)
    Sx = rng.standard_normal((6, 6, 200)).astype(np.float32)
    out["das_traces"] = Sx
    out["das_antennas"] = np.array(ants, dtype=np.int64)
    out["das_dt"] = np.array(g.dt)
    out["das_image"] = bm.das_backprojection(Sx, ants, g, 200)
    out["das_image_20x25"] = bm.das_backprojection(Sx, ants, g, 200, image_shape=(20, 25))
    ex9 = load_module("ref_ex09", REF / "examples" / "09_brain_clot_dataset_sample.py")
    g9 = YeeGrid(40, 40, 1e-3, 1e-3, device="cpu")
    ants9 = ex9.get_antenna_positions(5, 16, (20, 20), g9) if hasattr(ex9, "get_antenna_positions") else \
        bm.get_antenna_positions(5, 16, (20, 20), g9)
    S9 = rng.standard_normal((5, 5, 500)).astype(np.float32)
    out["das09_traces"], out["das09_antennas"] = S9, np.array(ants9, dtype=np.int64)
    out["das09_dt"] = np.array(g9.dt)
    out["das09_image"] = ex9.das_backprojection(S9, ants9, g9, 500)

    # 5. envelope / layered-medium DAS of AntennaRing.compute_das_image (src/datasets/brain/antenna.py)
    ant = load_module("ref_antenna", REF / "src" / "datasets" / "brain" / "antenna.py")
    ge = YeeGrid(30, 30, 1e-3, 1e-3, device="cpu")
    head = {"cx_cells": 15, "cy_cells": 14, "layers": [(12, 40.0), (11, 13.1), (9, 44.0), (8, 68.0), (6, 52.7),
                                                       (4, 38.1)]}
    for tag, n_t, geom, size in (("env_free", 300, None, 24), ("env_head", 301, head, 20)):
        ring = ant.AntennaRing(6, 13, 0, ge, component="Hz")
        ring.build_sources(S.GaussianPulse(1.0, sigma=5 * ge.dt), 0, n_t)
        Se = rng.standard_normal((6, 6, n_t)).astype(np.float32)
        out[f"{tag}_traces"] = Se
        out[f"{tag}_positions"] = np.array([(i, j) for i, j, _ in ring.positions], dtype=np.int64)
        out[f"{tag}_dt"] = np.array(ge.dt)
        out[f"{tag}_image"] = ring.compute_das_image(Se, image_size=size, head_geometry=geom)
    # 6. 3D trajectories (src/core/fdtd3d.py, MurABC3D, MaterialMap3D)
    from core.boundaries import MurABC3D
    from core.fdtd3d import FDTD3D
    from core.materials import MaterialMap3D

    g3 = YeeGrid(12, 10, 1e-3, 1e-3, Nz=8, dz=1e-3, device="cpu")
    f3 = FieldSet(g3)
    m3 = MaterialMap3D(g3, default=Material("coupling", 20.0, 0.1))
    m3.add_sphere((6, 5, 4), 3.2, Material("brain", 45.0, 0.8))
    m3.add_box((1, 1, 5), (4, 3, 7), Material("bone", 12.0, 0.2))
    m3.add_ellipsoid((8, 6, 3), (2.5, 1.5, 2.0), Material("clot", 60.0, 1.2))
    Ca3, Cb3 = m3.build3d()
    src3 = [
        S.PointSource(S.RickerWavelet(1.0, 30e9), 6, 5, "Ez", k=4, grid=g3, N_steps=60),
        S.PointSource(S.GaussianPulse(0.7, sigma=6 * g3.dt), 0, 4, "Hx", k=3, grid=g3, N_steps=60),
        S.PointSource(S.SinusoidalSource(0.3, 40e9), 5, 9, "Hy", k=2, grid=g3, N_steps=60),
        S.PointSource(S.ModulatedGaussian(0.5, 30e9, sigma=2e-11), 3, 3, "Ex", k=7, grid=g3, N_steps=60),
        S.PlaneSource(S.GaussianPulse(0.05, sigma=5 * g3.dt), "xy", 2, "Ey", grid=g3, N_steps=60),
        S.PointSource(S.UWBPulse(1.0, 10e9, 40e9), 6, 5, "Ez", k=4, grid=g3, N_steps=60),
        S.PointSource(S.GaussianPulse(0.4, sigma=4 * g3.dt), 7, 0, "Hz", k=5, grid=g3, N_steps=60),
    ]
    sim3 = FDTD3D(g3, f3, MurABC3D(g3, f3.Hx, f3.Hy, f3.Hz), S.SourceCollection(src3), Ca=Ca3, Cb=Cb3,
                  n_check=10)
    for _ in range(45):
        sim3.step()
    out["fd3_Ca"], out["fd3_Cb"], out["fd3_dt"] = Ca3.numpy(), Cb3.numpy(), np.array(g3.dt)
    for k, s_ in enumerate(src3):
        out[f"fd3_wave{k}"] = s_._waveform_tensor.numpy()
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        out[f"fd3_{c}"] = getattr(f3, c).numpy()
    out["fd3_last_field_max"] = np.array(sim3.last_field_max)
    # free space, 30 steps
    g4 = YeeGrid(10, 9, 1e-3, 1e-3, Nz=7, dz=1e-3, device="cpu")
    f4 = FieldSet(g4)
    src4 = [S.PointSource(S.GaussianPulse(1.0, sigma=5 * g4.dt), 5, 4, "Ez", k=3, grid=g4, N_steps=40),
            S.PointSource(S.RickerWavelet(0.5, 40e9), 2, 6, "Hz", k=1, grid=g4, N_steps=40)]
    sim4 = FDTD3D(g4, f4, MurABC3D(g4, f4.Hx, f4.Hy, f4.Hz), S.SourceCollection(src4), n_check=10)
    for _ in range(30):
        sim4.step()
    out["fs3_dt"] = np.array(g4.dt)
    for k, s_ in enumerate(src4):
        out[f"fs3_wave{k}"] = s_._waveform_tensor.numpy()
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        out[f"fs3_{c}"] = getattr(f4, c).numpy()
    out["fs3_last_field_max"] = np.array(sim4.last_field_max)
    # Nz = 1 mode
    g5 = YeeGrid(16, 14, 1e-3, 1e-3, Nz=1, device="cpu")
    f5 = FieldSet(g5)
    src5 = [S.PointSource(S.GaussianPulse(1.0, sigma=5 * g5.dt), 8, 7, "Hz", grid=g5, N_steps=40),
            S.PointSource(S.GaussianPulse(0.5, sigma=4 * g5.dt), 5, 9, "Ez", grid=g5, N_steps=40),
            S.PointSource(S.GaussianPulse(0.5, sigma=4 * g5.dt), 4, 3, "Hx", grid=g5, N_steps=40)]
    sim5 = FDTD3D(g5, f5, MurABC3D(g5, f5.Hx, f5.Hy, f5.Hz), S.SourceCollection(src5))
    for _ in range(30):
        sim5.step()
    out["nz1_dt"] = np.array(g5.dt)
    for k, s_ in enumerate(src5):
        out[f"nz1_wave{k}"] = s_._waveform_tensor.numpy()
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        out[f"nz1_{c}"] = getattr(f5, c).numpy()

    out["env_head_geometry"] = np.array([head["cx_cells"], head["cy_cells"]] +
                                        [v for layer in head["layers"] for v in layer], dtype=np.float64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
