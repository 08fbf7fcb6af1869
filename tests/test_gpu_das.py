"""DAS HIP kernel vs the NumPy oracle: bit-identical fp64 images.

The kernel forms every fp64 value as the reference's NumPy code does
(``examples/brain_mimo_imaging.py:211-266``), so images must be equal bit for
bit, which also gives an identical peak pixel. Covers the default clip mode,
the time gate of example 09, another medium speed (example 10), the
zero-outside-window mode, non-square images with a pitch different from dx,
and a single TX/RX pair.
"""

import math

import numpy as np
import pytest
import torch

from oracle import das_ref

pytestmark = pytest.mark.gpu


def _grid(nx, ny, dev):
    from waveforge_amd import core

    return core.YeeGrid(nx, ny, 1e-3, 1e-3, device=dev)


@pytest.mark.parametrize(
    "nx,ny,n_ant,T,img,kw",
    [
        (40, 40, 8, 300, None, {}),
        (64, 48, 16, 500, None, dict(gate_start=120)),
        (50, 70, 6, 200, (25, 35), dict(v_medium=299_792_458.0 / math.sqrt(13.1))),
        (32, 32, 5, 80, (40, 20), dict(out_of_range="zero")),
        (20, 20, 1, 50, None, {}),
        (60, 60, 32, 150, (24, 24), {}),  # the most receive distances a thread keeps in registers
        (60, 60, 40, 150, (20, 28), dict(out_of_range="zero")),  # beyond that: recomputed per pair
    ],
)
def test_das_matches_oracle_bitwise(device, nx, ny, n_ant, T, img, kw):
    from waveforge_amd import core

    rng = np.random.default_rng(nx * 7 + T)
    S = rng.standard_normal((n_ant, n_ant, T)).astype(np.float32)
    g = _grid(nx, ny, device)
    ants = core.get_antenna_positions(n_ant, min(nx, ny) * 0.42, (nx // 2, ny // 2), g)
    out = core.das_backprojection(S, ants, g, T, img, **kw)
    ref = das_ref.das_backprojection(S, ants, nx, ny, 1e-3, g.dt, T, img, **kw)
    assert out.dtype == np.float64 and out.shape == ref.shape
    np.testing.assert_array_equal(out, ref)
    assert np.unravel_index(out.argmax(), out.shape) == np.unravel_index(ref.argmax(), ref.shape)


def test_das_tensor_io(device):
    from waveforge_amd import core

    g = _grid(24, 24, device)
    S = torch.randn(4, 4, 100, device=device)
    ants = core.get_antenna_positions(4, 10, (12, 12), g)
    out = core.das_backprojection(S, ants, g, 100, return_tensor=True)
    assert isinstance(out, torch.Tensor) and out.device.type == "cuda"
    ref = das_ref.das_backprojection(S.cpu().numpy(), ants, 24, 24, 1e-3, g.dt, 100)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
