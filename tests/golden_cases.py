"""Load the reference fixtures and rebuild their set-ups for the oracle and the product."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

FIX = Path(__file__).resolve().parent / "golden" / "reference_fixtures.npz"


def fixtures():
    return np.load(FIX, allow_pickle=False)


def slab_sources(F):
    """(component, wave, i, j, amp, kind, position) of the slab case."""
    return [("Hz", torch.from_numpy(F["slab_wave"]), [12], [24], [1.0], "cells", 0)]


def lossy_sources(F):
    return [
        ("Hz", torch.from_numpy(F["lossy_wave0"]), [20], [18], [1.0], "cells", 0),
        ("Ex", torch.from_numpy(F["lossy_wave1"]), [3], [4], [1.0], "cells", 0),
        ("Ey", torch.from_numpy(F["lossy_wave2"]), list(range(5, 35)), [30] * 30,
         [0.5 + 0.02 * k for k in range(30)], "cells", 0),
        ("Hz", torch.from_numpy(F["lossy_wave3"]), None, None, None, "plane_xz", 2),
        ("Hz", torch.from_numpy(F["lossy_wave4"]), [20], [18], [1.0], "cells", 0),
    ]


CASES = {
    "slab": dict(nx=48, ny=48, courant=0.99, steps=150, sources=slab_sources, detectors=[(20, 24), (40, 24)],
                 n_check=50),
    "lossy": dict(nx=40, ny=36, courant=0.98, steps=80, sources=lossy_sources, detectors=[], n_check=1000),
}
