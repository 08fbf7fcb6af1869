"""Host-side logic of the product on the CPU (no kernel launches): API
validation mirrors the reference, waveform tables and material coefficients
equal the oracle's restatement, launch scheduling invariants, and the product
refusing to step CPU tensors."""

import math

import numpy as np
import pytest
import torch

from oracle import waveforms_ref as wr
from oracle.fdtd2d_ref import ca_cb_from_maps, material_ca_cb, stable_dt
from waveforge_amd import core
from waveforge_amd.engine import division_range, schedule_launches, split_launches


def test_grid_validation_matches_reference():
    with pytest.raises(ValueError):
        core.YeeGrid(1, 8, 1e-3, 1e-3, device="cpu")
    with pytest.raises(ValueError):
        core.YeeGrid(8, 8, -1e-3, 1e-3, device="cpu")
    with pytest.raises(ValueError):
        core.YeeGrid(8, 8, 1e-3, 1e-3, courant=1.0, device="cpu")
    g = core.YeeGrid(8, 8, 1e-3, 1e-3, device="cpu")
    with pytest.raises(ValueError):
        core.YeeGrid(8, 8, 1e-3, 1e-3, dt=2 * g.dt, device="cpu")
    with pytest.warns(UserWarning):
        core.YeeGrid(8, 8, 1e-3, 1e-3, dt=0.01 * g.dt, device="cpu")
    assert g.dt == stable_dt(1e-3, 1e-3)
    assert g.cell_index(2, 3) == 2 * 8 + 3


def test_fdtd2d_construction_checks():
    g = core.YeeGrid(8, 8, 1e-3, 1e-3, device="cpu")
    f = core.FieldSet(g)
    b = core.MurABC(g, f.Hz)
    with pytest.raises(ValueError):
        core.FDTD2D(core.YeeGrid(3, 8, 1e-3, 1e-3, device="cpu"), f, b)
    with pytest.raises(ValueError):
        core.FDTD2D(g, f, core.MurABC(core.YeeGrid(9, 8, 1e-3, 1e-3, device="cpu"), f.Hz))
    with pytest.raises(ValueError):
        core.FDTD2D(g, f, b, Ca=torch.ones(8, 8, 1))


def test_launch_shape_arguments_are_checked():
    g = core.YeeGrid(16, 16, 1e-3, 1e-3, device="cpu")
    f = core.FieldSet(g)
    b = core.MurABC(g, f.Hz)
    for bad in (dict(depth=0), dict(depth=9), dict(depth="deep"), dict(tile=2.5)):
        with pytest.raises(ValueError):
            core.FDTD2D(g, f, b, None, **bad)
    core.FDTD2D(g, f, b, None, depth="auto", tile="auto")
    core.FDTD2D(g, f, b, None, depth=8, tile=48)


def test_product_refuses_cpu_tensors():
    g = core.YeeGrid(8, 8, 1e-3, 1e-3, device="cpu")
    f = core.FieldSet(g)
    sim = core.FDTD2D(g, f, core.MurABC(g, f.Hz))
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        sim.step()
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        sim.run(10)


WAVES = [
    (lambda: core.GaussianPulse(1.3, sigma=7e-12), lambda t: wr.gaussian(t, 1.3, 35e-12, 7e-12)),
    (lambda: core.SinusoidalSource(0.5, 3e9, 0.0), lambda t: wr.sinusoidal(t, 0.5, 3e9, 0.0)),
    (lambda: core.RickerWavelet(1.0, 1e9), lambda t: wr.ricker(t, 1.0, 1e9, 1.5e-9)),
    (lambda: core.ModulatedGaussian(1.0, 1e9, sigma=1 / (2 * math.pi * 0.5e9)),
     lambda t: wr.modulated_gaussian(t, 1.0, 1e9, 1 / (2 * math.pi * 0.5e9), 5 / (2 * math.pi * 0.5e9))),
    (lambda: core.Chirp(1.0, 1e8, 1e9, 1e-10, 3e-9), lambda t: wr.chirp(t, 1.0, 1e8, 1e9, 1e-10, 3e-9)),
    (lambda: core.UWBPulse(1.0, 1e9, 3e9), lambda t: wr.uwb(t, 1.0, 1e9, 3e9)),
]


@pytest.mark.parametrize("k", range(len(WAVES)))
def test_waveform_tables_equal_oracle(k):
    make, ref = WAVES[k]
    dt = stable_dt(1e-3, 1e-3)
    w = make().build(3000, dt, torch.device("cpu"))
    r = ref(wr.time_axis(3000, dt))
    assert w.dtype == torch.float32 and torch.equal(w, r)


def test_waveform_closed_forms():
    """Oracle waveforms against fp64 math-module evaluation (fp32 tolerance) and
    their defining properties: Gaussian peak value, Ricker zero crossings at
    t0 +- 1/(sqrt2 pi fp), UWB peak normalised to the amplitude."""
    dt = stable_dt(1e-3, 1e-3)
    t = wr.time_axis(4000, dt)
    g = wr.gaussian(t, 2.0, 100 * dt, 10 * dt)
    n = 100
    assert abs(float(g[n]) - 2.0) < 1e-6
    exact = np.array([2.0 * math.exp(-((k * dt - 100 * dt) ** 2) / (2 * (10 * dt) ** 2)) for k in range(400)])
    assert np.allclose(g[:400].numpy(), exact, rtol=2e-5, atol=1e-6)
    fp = 1e9
    r = wr.ricker(wr.time_axis(4000, dt, torch.float64), 1.0, fp, 1.5 / fp).numpy()
    zc = 1.5 / fp + 1 / (math.sqrt(2) * math.pi * fp)
    kz = int(zc / dt)
    assert r[kz] * r[kz + 1] <= 0.0
    u = wr.uwb(wr.time_axis(20000, 1e-13, torch.float64), 1.0, 1e9, 3e9).numpy()
    assert abs(np.abs(u).max() - 1.0) < 1e-4


def test_material_map_matches_closed_form():
    g = core.YeeGrid(20, 16, 1e-3, 1e-3, device="cpu")
    mm = core.MaterialMap(g)
    brain = core.Material("brain", 45.0, 0.8)
    mm.add_rectangle((2, 5), (0, 15), core.Material("slab", 4.0, 0.0))
    mm.add_circle((10, 8), 4, brain)
    ca, cb = mm.build()
    a, b = material_ca_cb(45.0, 0.8, g.dt)
    assert ca.shape == (20, 16, 1) and ca.dtype == torch.float32
    assert float(ca[10, 8, 0]) == np.float32(a) and float(cb[10, 8, 0]) == np.float32(b)
    assert float(ca[3, 0, 0]) == 1.0
    assert float(cb[0, 0, 0]) == np.float32(material_ca_cb(1.0, 0.0, g.dt)[1])
    eps = torch.full((20, 16), 45.0, dtype=torch.float64)
    sig = torch.full((20, 16), 0.8, dtype=torch.float64)
    c1, c2 = core.coefficients_from_maps(eps, sig, g.dt)
    o1, o2 = ca_cb_from_maps(eps, sig, g.dt)
    assert torch.equal(c1.squeeze(2), o1) and torch.equal(c2.squeeze(2), o2)
    with pytest.raises(ValueError):
        core.coefficients_from_maps(eps * 0.01, sig, g.dt)


def test_source_validation():
    g = core.YeeGrid(10, 10, 1e-3, 1e-3, device="cpu")
    w = core.GaussianPulse(1.0, sigma=5e-12)
    with pytest.raises(ValueError):
        core.PointSource(w, 10, 0, "Hz", grid=g, N_steps=10)
    with pytest.raises(ValueError):
        core.PointSource(w, 0, 0, "Bz", grid=g, N_steps=10)
    with pytest.raises(ValueError):
        core.LineSource(w, "x", 3, 5, 5, "Hz", grid=g, N_steps=10)
    with pytest.raises(ValueError):
        core.PlaneSource(w, "xz", 10, "Hz", grid=g, N_steps=10)


def test_split_launches_even():
    for n in range(2, 60):
        for d in range(1, 9):
            r = split_launches(n, d)
            if r is None:
                assert d == 1 and n % 2
                continue
            assert sum(r) == n and len(r) % 2 == 0 and max(r) <= d and min(r) >= 1


def test_schedule_launches_checks():
    for start in (0, 5, 99, 100):
        for n in (2, 3, 50, 101, 333):
            for d in (1, 3, 6, 8):
                for ce in (1, 10, 100):
                    L, checks = schedule_launches(start, n, d, ce)
                    if L is None:
                        continue
                    assert sum(x for x, _ in L) == n and len(L) % 2 == 0
                    step = start
                    for dd, slot in L:
                        step += dd
                        if slot >= 0:
                            assert checks[slot] == step and step % ce == 0
                    assert len({s for _, s in L if s >= 0}) == len(checks)


def test_division_range_for_millimetre_cells():
    """The verified range for dx = 1 mm (1.024 * 2^-10) is every |x| below
    2^86, subnormals and both zeros included; divisors whose scale factor
    2^(31-e) would leave the normal range get no fast path."""
    lo, span = division_range(float(np.float32(1e-3)))
    assert lo == 0
    assert span == (127 + 86) << 23
    assert division_range(2.0 ** 31) == (0, (127 + 127) << 23)
    assert division_range(2.0 ** 32) is None
    assert division_range(2.0 ** -95) == (0, 128 << 23)
    assert division_range(2.0 ** -96) is None
    assert division_range(-1e-3) is None


def test_snapshot_free_fused_step_condition():
    """The fused 3D step skips the Mur snapshot copy only when no H source cell
    lies on a Mur plane of its component (index 0, 1, n-2, n-1 across a
    tangential axis); E sources never matter."""
    from types import SimpleNamespace

    from waveforge_amd import _native
    from waveforge_amd.core.fdtd3d import snapshot_is_state
    from waveforge_amd.engine import INJ_DTYPE

    nx, ny, nz = 20, 16, 12

    def table(entries, nz=nz):
        rec = np.zeros(len(entries), dtype=INJ_DTYPE)
        for n, (comp, i, j, k) in enumerate(entries):
            rec[n]["comp"] = _native.WF_COMP3[comp]
            rec[n]["cell"] = (i * ny + j) * nz + k
        return SimpleNamespace(n_grp=len(entries), host_inj=rec)

    assert snapshot_is_state(SimpleNamespace(n_grp=0, host_inj=None), nx, ny, nz)
    assert snapshot_is_state(table([("Ez", 0, 0, 0), ("Ex", 1, 15, 11), ("Ey", 19, 1, 10)]), nx, ny, nz)
    assert snapshot_is_state(table([("Hx", 0, 5, 5), ("Hy", 7, 0, 5), ("Hz", 7, 7, 0)]), nx, ny, nz)  # normal axes
    assert snapshot_is_state(table([("Hx", 10, 2, 9), ("Hy", 17, 13, 2), ("Hz", 2, 13, 11)]), nx, ny, nz)
    for bad in [("Hx", 5, 1, 5), ("Hx", 5, 5, 10), ("Hy", 18, 5, 5), ("Hy", 5, 5, 0), ("Hz", 0, 5, 5),
                ("Hz", 5, 14, 5)]:
        assert not snapshot_is_state(table([("Ez", 5, 5, 5), bad]), nx, ny, nz), bad
    # Nz = 1: no z faces, so k = 0 is no Mur plane
    assert snapshot_is_state(table([("Hx", 5, 5, 0), ("Hy", 5, 5, 0)], nz=1), nx, ny, 1)
    assert not snapshot_is_state(table([("Hx", 5, 0, 0)], nz=1), nx, ny, 1)


def test_default_launch_shape_by_workload():
    """8 steps deep; 8-row tiles while there are at most 288 of them, else
    16-row tiles below 64 tiles of 48 rows, 32-row tiles below 1000 of them,
    then 64-row tiles (48-row for a run of 2^23 cells or more); smaller tiles
    where a run's detectors crowd one tile."""
    from types import SimpleNamespace

    import numpy as np
    import torch

    from waveforge_amd.core.fdtd2d import default_launch_shape
    from waveforge_amd.engine import TB_MAX_DETS_PER_TILE

    def runs(*shapes, det=None):
        return [SimpleNamespace(nx=a, ny=b, sources=None, det_cell=det, extra={}) for a, b in shapes]

    assert default_launch_shape(runs((4096, 4096))) == (48, 8, 1)   # 86 x 86 tiles, 2^24 cells
    assert default_launch_shape(runs((2048, 2048))) == (64, 8, 1)   # 43 x 43 tiles, 2^22 cells
    assert default_launch_shape(runs((1024, 1024))) == (32, 8, 1)   # 22 x 22 tiles
    assert default_launch_shape(runs(*[(256, 256)] * 16)) == (32, 8, 1)   # 36 tiles each
    assert default_launch_shape(runs(*[(256, 256)] * 64)) == (64, 8, 1)   # 2304 tiles
    assert default_launch_shape(runs((256, 256))) == (8, 8, 1)   # 32 x 6 = 192 tiles of 8 rows
    assert default_launch_shape(runs((128, 128), (200, 120))) == (8, 8, 1)   # 48 + 75 such tiles
    assert default_launch_shape(runs((300, 380))) == (16, 8, 1)   # 38 x 8 = 304 > 288; 7 x 8 = 56 < 64
    assert default_launch_shape(runs((384, 384))) == (32, 8, 1)   # 8 x 8 = 64 tiles of 48 rows
    # a 9 x 9 block of detectors (81 > 64 per tile) inside one 32-row tile at
    # every depth (rows 140-148), split 4 + 5 rows by the 16-row tile boundary
    # at row 144: (32, 8) and (32, 6) are refused, (16, 6) is taken
    side = int(np.sqrt(TB_MAX_DETS_PER_TILE)) + 1
    cells = [(140 + a) * 1024 + 100 + b for a in range(side) for b in range(side)]
    det = torch.tensor(cells, dtype=torch.int32)
    assert default_launch_shape(runs((1024, 1024), det=det)) == (16, 6, 1)


def test_grids_beyond_int32_cells_are_rejected():
    """Cells are int32 in the source and detector tables: 46341^2 > 2^31-1 is
    refused with a clear error, 46340^2 is accepted."""
    import pytest as _pytest

    from waveforge_amd.engine import check_grid_cells

    check_grid_cells(46340, 46340)
    check_grid_cells(4, 2 ** 29 - 1)
    for nx, ny in ((46341, 46341), (4, 2 ** 29)):
        with _pytest.raises(ValueError, match="2\\^31-cell index range"):
            check_grid_cells(nx, ny)


def test_tile_lists_match_brute_force():
    """TBPlan.tile_lists: per tile, exactly the source groups whose cell lies
    in the tile's region (core plus depth cells around it) and the detectors
    in its core, for runs of odd sizes, both tile widths and several depths."""
    import numpy as np
    import torch

    from waveforge_amd import _native
    from types import SimpleNamespace

    from waveforge_amd.engine import RunState, StepCoefficients, TBPlan

    rng = np.random.default_rng(4)
    cpu = torch.device("cpu")
    coef = StepCoefficients.from_grid(2e-12, 1e-3, 1e-3)
    runs = []
    for nx, ny, n_src, n_det in ((97, 150, 7, 12), (40, 33, 0, 3), (130, 71, 20, 0), (64, 64, 3, 64)):
        f = torch.zeros((3, nx, ny))
        cells = sorted({int(rng.integers(nx)) * ny + int(rng.integers(ny)) for _ in range(n_src)})
        # the group cells a SourceTable would hold (one Hz and one Ey group per cell)
        gc = np.array(cells + cells, dtype=np.int32)
        table = SimpleNamespace(n_grp=gc.size, n_inj=gc.size, group_cells=gc, inert=[], ptrs=lambda: (0, 0))
        rs = RunState(f[0], f[1], f[2], None, None, nx, ny, coef, sources=table)
        if n_det:
            det = rng.choice(nx * ny, size=n_det, replace=False).astype(np.int32)
            rs.det_cell = torch.from_numpy(det)
            rs.traces = torch.zeros((n_det, 4))
        runs.append(rs)
    for tile_i, cpl in ((48, 1), (16, 1), (32, 2)):
        plan = TBPlan(runs, tile_i, cpu, fast_div=False, cpl=cpl)
        for depth in (1, 5, 8):
            lists, grp, det = (x.numpy() for x in plan.tile_lists(depth))
            tiles, n = plan.tiles(depth)
            tiles = tiles.numpy()
            TI, TJ = tile_i, _native.tile_cols(depth, cpl)
            for t in range(n):
                k, tile = tiles[t]
                r = runs[k]
                ntj = -(-r.ny // TJ)
                I0, J0 = (tile // ntj) * TI, (tile % ntj) * TJ
                gc = r.sources.group_cells.astype(np.int64)
                want_g = sorted(g for g in range(gc.size)
                                if I0 - depth <= gc[g] // r.ny < I0 + TI + depth
                                and J0 - depth <= gc[g] % r.ny < J0 + TJ + depth)
                assert sorted(grp[lists[t, 0]:lists[t, 1]].tolist()) == want_g, (tile_i, cpl, depth, t)
                dc = [] if r.det_cell is None else r.det_cell.numpy().tolist()
                want_d = sorted(d for d, c in enumerate(dc) if I0 <= c // r.ny < I0 + TI and J0 <= c % r.ny < J0 + TJ)
                assert sorted(det[lists[t, 2]:lists[t, 3]].tolist()) == want_d, (tile_i, cpl, depth, t)
