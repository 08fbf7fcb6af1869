"""Batched multi-run driver vs the oracle, run by run, bit for bit.

Runs of different grid sizes, media, sources and receivers advance together
in one launch sequence; each run must equal its own standalone oracle run
(alone-vs-in-batch identity). Also: a diverging run is isolated, graph
replays equal eager stepping, and ``run_mimo`` equals the oracle per TX.
"""

import math

import numpy as np
import pytest
import torch

from cases import assert_bitwise, waveform_table
from oracle.fdtd2d_ref import OracleDivergedError, OracleFDTD2D, OracleSource, stable_dt

pytestmark = pytest.mark.gpu

DX = 1e-3


def _media(nx, ny, seed, dt):
    g = torch.Generator().manual_seed(seed)
    eps = 1.0 + 40.0 * torch.rand((nx, ny), generator=g, dtype=torch.float64)
    sig = 1.5 * torch.rand((nx, ny), generator=g, dtype=torch.float64)
    alpha = sig * dt / (2.0 * 8.8541878128e-12 * eps)
    ca = ((1.0 - alpha) / (1.0 + alpha)).float()
    cb = ((dt / (8.8541878128e-12 * eps)) / (1.0 + alpha)).float()
    return ca, cb


def _specs(device, steps, dt, sizes=((37, 41), (64, 64), (100, 53), (20, 90), (150, 150)), amps=None,
           division="reciprocal"):
    from waveforge_amd.mimo import RunSpec

    specs, oracles = [], []
    for k, (nx, ny) in enumerate(sizes):
        ca, cb = _media(nx, ny, k, dt)
        kinds = ["gauss", "dgauss", "sin"]
        a0 = 1.0 if amps is None else amps[k]
        src = [("Hz", nx // 2, ny // 3, kinds[k % 3], a0), ("Ey", 2, ny - 3, kinds[(k + 1) % 3], 0.5 * a0)]
        rx = [(1, 1), (nx - 2, ny // 2), (nx // 2, ny // 3)]
        waves = {kd: waveform_table(kd, steps + 4, dt) for kd in kinds}
        specs.append(RunSpec(nx, ny, ca.to(device), cb.to(device),
                             sources=[(c, i, j, waves[w].to(device), a) for c, i, j, w, a in src],
                             receivers=rx, n_record=steps))
        osrc = [OracleSource(c, waves[w], torch.tensor([i]), torch.tensor([j]), torch.tensor([a]))
                for c, i, j, w, a in src]
        oracles.append(OracleFDTD2D(nx, ny, DX, DX, dt, Ca=ca.unsqueeze(2), Cb=cb.unsqueeze(2), sources=osrc,
                                    detectors=rx, n_check=10 ** 9, division=division))
    return specs, oracles


@pytest.mark.parametrize("division", ["reciprocal", "ieee"])
@pytest.mark.parametrize("depth,tile", [(6, 32), (4, 16), (8, 48), (3, 32)])
def test_batched_runs_each_equal_standalone_oracle(device, depth, tile, division):
    from waveforge_amd.mimo import BatchedRuns

    steps = 70
    dt = stable_dt(DX, DX)
    specs, oracles = _specs(device, steps, dt, division=division)
    br = BatchedRuns(specs, dx=DX, dy=DX, dt=dt, device=device, depth=depth, tile_rows=tile, n_check=25,
                     division=division)
    br.advance(30)
    br.advance(steps - 30)
    st = br.status()
    assert not any(st["diverged"]) and st["checks"] == 2
    for k, o in enumerate(oracles):
        print(f"[batched] run={k} grid={o.Nx}x{o.Ny} depth={depth} tile={tile} division={division}", flush=True)
        o.run(steps)
        ex, ey, hz = br.fields(k)
        assert_bitwise(ex, o.Ex.reshape(o.Nx, o.Ny), f"run {k} Ex")
        assert_bitwise(ey, o.Ey.reshape(o.Nx, o.Ny), f"run {k} Ey")
        assert_bitwise(hz, o.Hz.reshape(o.Nx, o.Ny), f"run {k} Hz")
        assert_bitwise(br.traces(k), o.trace_array(), f"run {k} traces")


@pytest.mark.parametrize("tile", [32, 48])
def test_diverging_run_is_isolated(device, tile):
    from waveforge_amd.mimo import BatchedRuns

    steps = 60
    dt = stable_dt(DX, DX)
    specs, oracles = _specs(device, steps, dt, sizes=((30, 30), (40, 35), (25, 50)), amps=(1e-3, 1e6, 1e-3))
    br = BatchedRuns(specs, dx=DX, dy=DX, dt=dt, device=device, depth=5, tile_rows=tile, n_check=10,
                     stability_threshold=10.0)
    br.advance(steps)
    st = br.status()
    for k, o in enumerate(oracles):
        o.n_check, o.threshold = 10, 10.0
        diverged = False
        try:
            o.run(steps)
        except OracleDivergedError:
            diverged = True
        assert st["diverged"][k] == diverged, k
        ex, ey, hz = br.fields(k)
        # a diverged run keeps the fields of its failing check step
        assert_bitwise(hz, o.Hz.reshape(o.Nx, o.Ny), f"run {k} Hz")
        assert_bitwise(ex, o.Ex.reshape(o.Nx, o.Ny), f"run {k} Ex")
    assert st["diverged"] == [False, True, False]


def test_graph_replay_equals_eager(device):
    """A captured 24-step chunk replayed several times, with eager launches and
    synchronisations between replays, equals eager stepping after every replay."""
    from waveforge_amd.mimo import BatchedRuns

    dt = stable_dt(DX, DX)
    steps = 24 * 4 + 12
    specs_a, _ = _specs(device, steps, dt, sizes=((40, 33), (64, 70)))
    specs_b, _ = _specs(device, steps, dt, sizes=((40, 33), (64, 70)))
    a = BatchedRuns(specs_a, dx=DX, dy=DX, dt=dt, device=device, depth=6, n_check=10 ** 9)
    b = BatchedRuns(specs_b, dx=DX, dy=DX, dt=dt, device=device, depth=6, n_check=10 ** 9)
    a.capture(24)
    a.replay(1)
    b.advance(24)
    a.advance(12)  # eager work between replays
    b.advance(12)
    torch.cuda.synchronize()
    for r in range(3):
        a.replay(1)
        torch.cuda.synchronize()
        b.advance(24)
        for k in range(2):
            for fa, fb, nm in zip(a.fields(k), b.fields(k), ("Ex", "Ey", "Hz")):
                assert torch.equal(fa, fb), f"replay {r} run {k} {nm}"
            assert torch.equal(a.traces(k), b.traces(k)), f"replay {r} run {k} traces"
    assert a.steps_completed == b.steps_completed == steps


def test_graph_replay_judges_stability_at_chunk_ends(device):
    """Three replays of a 20-step chunk with threshold 10 flag the same run
    as the oracle checking every 20 steps, and that run keeps the fields of
    its failing check step; the others equal the oracle."""
    from waveforge_amd.mimo import BatchedRuns

    steps, chunk = 60, 20
    dt = stable_dt(DX, DX)
    specs, oracles = _specs(device, steps, dt, sizes=((30, 30), (40, 35), (25, 50)), amps=(1e-3, 1e6, 1e-3))
    br = BatchedRuns(specs, dx=DX, dy=DX, dt=dt, device=device, depth=5, tile_rows=32, n_check=10 ** 9,
                     stability_threshold=10.0)
    br.capture(chunk)
    br.replay(steps // chunk)
    st = br.status()
    for k, o in enumerate(oracles):
        o.n_check, o.threshold = chunk, 10.0
        diverged = False
        try:
            o.run(steps)
        except OracleDivergedError:
            diverged = True
        assert st["diverged"][k] == diverged, k
        ex, ey, hz = br.fields(k)
        assert_bitwise(hz, o.Hz.reshape(o.Nx, o.Ny), f"run {k} Hz")
        assert_bitwise(ex, o.Ex.reshape(o.Nx, o.Ny), f"run {k} Ex")
    assert st["diverged"] == [False, True, False] and st["checks"] == 3
    assert st["field_max"][1] > 10.0 and st["field_max"][0] < 10.0


@pytest.mark.parametrize("steps", [97, 49, 95])
def test_image_phantoms_graph_equals_eager(device, steps):
    """image_phantoms with graph replay of 48-step chunks equals the eager
    path bit for bit (traces and images), also when one step is left over
    (97 = 2 x 48 + 1, 49 = 48 + 1) or a remainder of 47."""
    from waveforge_amd import mimo, workloads

    dt = workloads.imaging_dt()
    phantoms = [workloads.head_phantom(n, 8, seed=40 + n, clot=True) for n in (48, 64)]
    ph = [(p.eps, p.sigma, p.antennas) for p in phantoms]
    print(f"[image-phantoms] steps={steps} chunk=48", flush=True)
    img_g, tr_g, br_g = mimo.image_phantoms(ph, steps, dt=dt, device=device, use_graph=True, chunk=48, depth=6)
    img_e, tr_e, br_e = mimo.image_phantoms(ph, steps, dt=dt, device=device, use_graph=False, depth=6)
    torch.cuda.synchronize()
    assert br_g.steps_completed == br_e.steps_completed == steps
    assert not any(br_g.status()["diverged"])
    for p in range(2):
        assert torch.equal(tr_g[p], tr_e[p]), f"phantom {p} traces"
        assert torch.equal(img_g[p], img_e[p]), f"phantom {p} image"
        assert float(img_g[p].abs().max()) > 0


def test_run_mimo_matches_oracle_per_tx(device):
    from waveforge_amd import core
    from waveforge_amd.mimo import run_mimo

    n, steps = 60, 150
    g = core.YeeGrid(n, n, DX, DX, device=device)
    mm = core.MaterialMap(g, default=core.Material("coupling", 20.0, 0.1))
    mm.add_circle((30, 30), 20, core.Material("brain", 45.0, 0.8))
    Ca, Cb = mm.build()
    ants = core.get_antenna_positions(6, 26, (30, 30), g)
    S = run_mimo(g, Ca, Cb, ants, steps)
    assert S.shape == (6, 6, steps) and S.dtype == np.float32
    wave = core.RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9).build(steps, g.dt, torch.device("cpu"))
    for tx, (i, j) in enumerate(ants):
        o = OracleFDTD2D(n, n, DX, DX, g.dt, Ca=Ca.cpu(), Cb=Cb.cpu(),
                         sources=[OracleSource("Hz", wave, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]))],
                         detectors=ants, n_check=500, division="reciprocal")
        o.run(steps)
        np.testing.assert_array_equal(S[tx], o.trace_array().numpy())


def test_default_launch_shape_follows_the_batch_size(device):
    """With no shape given, a batch runs 8 steps deep with the tile height of
    FDTD2D's rule over the whole set (32-row tiles below 1000 tiles of 48 x 48
    cells, 64-row tiles above; 8-row tiles while there are at most 288 of
    them, else 16-row tiles below 64 tiles of 48 rows); one of the two given
    takes the other from that default."""
    from waveforge_amd.mimo import BatchedRuns, RunSpec

    dt = stable_dt(DX, DX)

    def batch(n, count, **kw):
        ca = torch.full((n, n), 0.99, device=device)
        cb = torch.full((n, n), 0.05, device=device)
        specs = [RunSpec(n, n, ca, cb, sources=[], receivers=[], n_record=0) for _ in range(count)]
        return BatchedRuns(specs, dx=DX, dy=DX, dt=dt, device=device, **kw)

    big = batch(256, 16)     # 16 x 6 x 6 = 576 tiles of 48 rows, 48 columns
    assert (big.tile_rows, big.depth) == (32, 8)
    assert batch(256, 32).tile_rows == 64   # 1152 such tiles
    small = batch(128, 2)    # 2 x 3 x 3 = 18 such tiles, 2 x 16 x 3 = 96 of 8 rows
    assert (small.tile_rows, small.depth) == (8, 8)
    mid = batch(300, 1)      # 7 x 7 = 49 such tiles, 38 x 7 = 266 of 8 rows
    assert (mid.tile_rows, mid.depth) == (8, 8)
    assert batch(330, 1).tile_rows == 16  # 7 x 7 = 49 tiles of 48 rows, 42 x 7 = 294 of 8 rows
    assert (batch(256, 16, depth=6).tile_rows, batch(256, 16, tile_rows=48).depth) == (32, 8)
    big.advance(10)
    small.advance(10)
    assert big.steps_completed == small.steps_completed == 10


@pytest.mark.parametrize("comp,i,j,amp", [("Hz", 3, 5, 1.0), ("Ex", 0, 0, -0.37), ("Ey", 19, 11, 2.5e-3),
                                          ("Ez", 4, 4, 1.0)])
def test_single_cell_source_table_equals_general(device, comp, i, j, amp):
    """``SourceTable.cell`` (the direct build batched drivers use) holds the
    same records as the general constructor over one ``CellSource``; a
    component the 2D kernels do not update goes to the general path (inert)."""
    from waveforge_amd.engine import CellSource, SourceTable

    wave = torch.linspace(-1.0, 1.0, 37, device=device)
    a = SourceTable.cell(comp, wave, i, j, amp, 20, 12)
    b = SourceTable([CellSource(comp, wave, i, j, amp)], 20, 12, device, upload=False)
    assert (a.n_inj, a.n_grp, a.min_wave_len, len(a.inert)) == (b.n_inj, b.n_grp, b.min_wave_len, len(b.inert))
    if b.n_grp:
        assert a.host_inj.tobytes() == b.host_inj.tobytes()
        assert np.array_equal(a.host_grp, b.host_grp) and np.array_equal(a.group_cells, b.group_cells)

