"""The largest grids the stepping kernels take. In 2D: 46340 x 46340 cells (2.15e9,
just below the int32 cell range of the source and detector tables; each field
is 8.6 GB, about 95 GB in all with the coefficients and the blocked path's
second state).

No oracle finishes at this size, so the check is a size-independent one: a
point source at the centre reaches only cells within n steps after n steps,
so as long as the pulse stays away from the boundary the window around the
source is independent of the grid around it. The window of the huge grid must
equal the whole of a 96 x 96 grid with the source at its centre, bit for bit,
on the temporally blocked path (the headline launch shape) and on the
per-step path; the small grid itself is checked against the oracle by the
other parity tests. This exercises the 64-bit addressing of every kernel on
the path at a size where 32-bit offsets would wrap.
"""

import pytest
import torch

pytestmark = pytest.mark.gpu

BIG = 46340
SMALL = 96
STEPS = 24


def _sim(n, device, depth, tile):
    from waveforge_amd import core

    grid = core.YeeGrid(n, n, 1e-3, 1e-3, courant=0.98, device=device)
    fields = core.FieldSet(grid)
    boundary = core.MurABC(grid, fields.Hz)
    c = n // 2
    src = core.PointSource(core.RickerWavelet(1.0, 30e9, t0=40e-12), c, c, "Hz", grid=grid, N_steps=STEPS + 4)
    ca = torch.full((n, n, 1), 0.9925, dtype=torch.float32, device=device)
    cb = torch.full((n, n, 1), 0.0625, dtype=torch.float32, device=device)
    sim = core.FDTD2D(grid, fields, boundary, core.SourceCollection([src]), Ca=ca, Cb=cb, depth=depth, tile=tile)
    return sim, c


@pytest.mark.parametrize("path", ["blocked", "per_step"])
def test_largest_grid_window_equals_small_grid(device, path):
    print(f"[large-grid] {BIG}x{BIG} path={path} steps={STEPS}", flush=True)
    depth, tile = (8, 48) if path == "blocked" else (None, None)
    small, cs = _sim(SMALL, device, depth, tile)
    big, cb = _sim(BIG, device, depth, tile)
    for sim in (small, big):
        if path == "blocked":
            sim.run(STEPS)
        else:
            for _ in range(STEPS):
                sim.step()
    torch.cuda.synchronize()
    lo = cb - cs
    for name in ("Ex", "Ey", "Hz"):
        w = getattr(big.fields, name)[lo:lo + SMALL, lo:lo + SMALL].contiguous().cpu()
        s = getattr(small.fields, name).contiguous().cpu()
        assert float(s.abs().max()) > 0.0, name
        assert torch.equal(w.view(torch.int32), s.view(torch.int32)), name
    del big
    torch.cuda.empty_cache()


BIG3 = 1290   # 1290^3 = 2.147e9 cells, just below the int32 cell range
SMALL3 = 48
STEPS3 = 12


def _sim3(n, device, fused):
    from waveforge_amd import core

    grid = core.YeeGrid(n, n, 1e-3, 1e-3, Nz=n, dz=1e-3, courant=0.98, device=device)
    fields = core.FieldSet(grid)
    c = n // 2
    src = core.PointSource(core.RickerWavelet(1.0, 30e9, t0=40e-12), c, c, "Ez", k=c, grid=grid,
                           N_steps=STEPS3 + 4)
    sim = core.FDTD3D(grid, fields, core.MurABC3D(grid, fields.Hx, fields.Hy, fields.Hz),
                      core.SourceCollection([src]), fused=fused)
    return sim, c


@pytest.mark.parametrize("fused", [True, False])
def test_largest_3d_grid_window_equals_small_grid(device, fused):
    """The same check in 3D: a 1290^3 free-space grid (8.6 GB per component,
    about 103 GB with the fused path's second state) against a 48^3 grid."""
    print(f"[large-grid] {BIG3}^3 fused={fused} steps={STEPS3}", flush=True)
    small, cs = _sim3(SMALL3, device, fused)
    small.run(STEPS3)
    big, cb = _sim3(BIG3, device, fused)
    big.run(STEPS3)
    torch.cuda.synchronize()
    lo = cb - cs
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        w = getattr(big.fields, name)[lo:lo + SMALL3, lo:lo + SMALL3, lo:lo + SMALL3].contiguous().cpu()
        s = getattr(small.fields, name).contiguous().cpu()
        assert float(s.abs().max()) > 0.0, name
        assert torch.equal(w.view(torch.int32), s.view(torch.int32)), name
    del big
    torch.cuda.empty_cache()
