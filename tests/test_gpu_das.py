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


def test_das_batch_equals_single_and_oracle(device):
    """wf_das2d_batch over 41 images of mixed sizes, element counts (1..40,
    both sides of the register limit), gates, speeds and both range modes
    (two launches: 32 + 9 images): each image equals wf_das2d's and the
    oracle's bit for bit."""
    from waveforge_amd.core.imaging import DasBatch, run_das

    rng = np.random.default_rng(5)
    batch = DasBatch(device)
    cases = []
    for k in range(41):
        n_ant = [1, 4, 8, 16, 32, 33, 40][k % 7]
        nx, ny = int(rng.integers(12, 70)), int(rng.integers(12, 70))
        T = int(rng.integers(50, 400))
        v = 299_792_458.0 / math.sqrt([40.0, 13.1, 1.0][k % 3])
        gate = [0, 0, 30][k % 3]
        clip = k % 4 != 3
        dt = 1.6e-12 * (1 + k % 2)
        S = torch.from_numpy(rng.standard_normal((n_ant, n_ant, T)).astype(np.float32)).to(device)
        ang = np.linspace(0, 2 * np.pi, n_ant, endpoint=False)
        ants = [(int(nx / 2 + 0.4 * nx * np.cos(a)), int(ny / 2 + 0.4 * ny * np.sin(a))) for a in ang]
        gx, gy = nx + 5 * (k % 3), ny + 3 * (k % 3)  # the grid the image spans (pitch = grid extent / pixels)
        ax = torch.tensor([a[0] * 1e-3 for a in ants], dtype=torch.float64, device=device)
        ay = torch.tensor([a[1] * 1e-3 for a in ants], dtype=torch.float64, device=device)
        kw = dict(pitch_x=gx * 1e-3 / nx, pitch_y=gy * 1e-3 / ny, v_medium=v, dt=dt, gate_start=gate, clip=clip)
        img_b = torch.full((nx, ny), float("nan"), dtype=torch.float64, device=device)
        img_s = torch.full((nx, ny), float("nan"), dtype=torch.float64, device=device)
        batch.add(S, ax, ay, img_b, **kw)
        run_das(S, ax, ay, img_s, **kw)
        cases.append((S, ants, img_b, img_s, kw, (gx, gy)))
    assert len(batch) == 41
    print("[das-batch] 41 images in one call", flush=True)
    batch.launch()
    torch.cuda.synchronize()
    for k, (S, ants, img_b, img_s, kw, (gx, gy)) in enumerate(cases):
        assert torch.equal(img_b, img_s), f"image {k}: batch != single"
        if k < 14:  # the oracle on a sample (pure NumPy, slower)
            nx, ny = img_b.shape
            ref = das_ref.das_backprojection(S.cpu().numpy(), ants, gx, gy, 1e-3, kw["dt"], S.shape[2],
                                             (nx, ny), v_medium=kw["v_medium"], gate_start=kw["gate_start"],
                                             out_of_range="clip" if kw["clip"] else "zero")
            np.testing.assert_array_equal(img_b.cpu().numpy(), ref, err_msg=f"image {k}")


def test_das_batch_rejects_a_bad_descriptor(device):
    """A bad descriptor anywhere in the array fails the whole call before any
    launch, with the reason, and no image is written."""
    from waveforge_amd import _native
    from waveforge_amd.core.imaging import das_args

    S = torch.zeros((2, 2, 10), device=device)
    ax = torch.zeros(2, dtype=torch.float64, device=device)
    imgs = [torch.full((8, 8), 7.0, dtype=torch.float64, device=device) for _ in range(3)]
    args = [das_args(S, ax, ax, im, pitch_x=1e-3, pitch_y=1e-3, v_medium=1e8, dt=1e-12) for im in imgs]
    args[2].dt = 0.0
    arr = (_native.WfDas2dArgs * 3)(*args)
    lib = _native.load()
    import ctypes

    rc = lib.wf_das2d_batch(arr, 3, ctypes.c_void_p(_native.stream_handle(device)))
    assert rc == -1 and b"v_medium and dt" in lib.wf_last_error()
    torch.cuda.synchronize()
    assert all(bool((im == 7.0).all()) for im in imgs)
