"""Continuous-batching server: every request's traces and image are the same
served in the mix as served alone (bit for bit), equal the oracle for a small
request, and the HBM-footprint admission never exceeds its budget."""

import numpy as np
import pytest
import torch

from oracle import das_ref
from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource

pytestmark = pytest.mark.gpu


def _stream(n=6, seed=3):
    from waveforge_amd.serving import poisson_requests

    return poisson_requests(n, rate_hz=200.0, n_min=40, n_max=90, elements=(4, 6), seed=seed, steps_per_diag=2.5)


def test_mix_equals_alone_and_budget(device):
    from waveforge_amd.serving import ImagingServer, request_footprint
    from waveforge_amd.workloads import imaging_dt

    dt = imaging_dt()
    reqs = _stream()
    fps = [request_footprint(r.eps.shape[0], r.eps.shape[1], len(r.antennas), r.n_steps, 48) for r in reqs]
    budget = sorted(fps)[-1] + sorted(fps)[-2] + 1  # room for about two requests at a time
    srv = ImagingServer(dt=dt, device=device, hbm_budget=budget, quantum=48, depth=6)
    peak = [0]
    orig_admit = srv.admit

    def admit(req, now):
        orig_admit(req, now)
        peak[0] = max(peak[0], srv.used)
        assert srv.used <= budget

    srv.admit = admit
    stats = srv.serve(reqs, realtime=False)
    assert stats["requests"] == len(reqs) and stats["failed"] == 0
    assert peak[0] <= budget
    mixed = {r.rid: (r.traces.clone(), r.image.clone()) for r in srv.done}
    for r in _stream():
        alone = ImagingServer(dt=dt, device=device, quantum=48, depth=6, max_requests=1)
        alone.serve([r], realtime=False)
        a = alone.done[0]
        print(f"[serving] request {r.rid} grid={r.eps.shape} E={len(r.antennas)} steps={r.n_steps}", flush=True)
        assert torch.equal(a.traces, mixed[r.rid][0]), f"request {r.rid} traces differ alone vs mixed"
        assert torch.equal(a.image, mixed[r.rid][1]), f"request {r.rid} image differs alone vs mixed"


@pytest.mark.parametrize("fcfs_tiles", [1, 40, None])
def test_tile_budget_rounds_equal_alone(device, fcfs_tiles):
    """Rounds that advance only the oldest requests up to a tile budget (1:
    one request at a time, 40: a few, None: every admitted request) give
    every request the traces and image it gets alone; with a budget of one
    tile the requests finish in arrival order."""
    from waveforge_amd.serving import ImagingServer
    from waveforge_amd.workloads import imaging_dt

    dt = imaging_dt()
    srv = ImagingServer(dt=dt, device=device, quantum=48, depth=6, fcfs_tiles=fcfs_tiles)
    stats = srv.serve(_stream(), realtime=False)
    assert stats["requests"] == 6 and stats["failed"] == 0
    if fcfs_tiles == 1:
        assert [r.rid for r in srv.done] == sorted(r.rid for r in srv.done)
    mixed = {r.rid: r for r in srv.done}
    for r in _stream():
        alone = ImagingServer(dt=dt, device=device, quantum=48, depth=6, max_requests=1)
        alone.serve([r], realtime=False)
        a = alone.done[0]
        print(f"[serving] fcfs_tiles={fcfs_tiles} request {r.rid} grid={r.eps.shape} E={len(r.antennas)}", flush=True)
        assert torch.equal(a.traces, mixed[r.rid].traces), f"request {r.rid} traces differ"
        assert torch.equal(a.image, mixed[r.rid].image), f"request {r.rid} image differs"


@pytest.mark.parametrize("division", ["reciprocal", "ieee"])
def test_served_request_matches_oracle(device, division):
    from waveforge_amd.core.sources import RickerWavelet
    from waveforge_amd.serving import ImagingServer
    from waveforge_amd.workloads import imaging_dt
    from waveforge_amd.core.materials import coefficients_from_maps

    dt = imaging_dt()
    req = _stream(n=1, seed=11)[0]
    srv = ImagingServer(dt=dt, device=device, quantum=60, depth=5, division=division)
    srv.serve([req], realtime=False)
    r = srv.done[0]
    nx, ny = req.eps.shape
    ca, cb = coefficients_from_maps(req.eps, req.sigma, dt)
    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9).build(req.n_steps + 60, dt, torch.device("cpu"))
    S = np.zeros((len(req.antennas), len(req.antennas), req.n_steps), dtype=np.float32)
    for t, (i, j) in enumerate(req.antennas):
        o = OracleFDTD2D(nx, ny, 1e-3, 1e-3, dt, Ca=ca, Cb=cb, detectors=list(req.antennas), n_check=10 ** 9,
                         division=division, sources=[OracleSource("Hz", wave, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]))])
        o.run(req.n_steps)
        S[t] = o.trace_array().numpy()
    np.testing.assert_array_equal(r.traces.cpu().numpy(), S)
    img = das_ref.das_backprojection(S, req.antennas, nx, ny, 1e-3, dt, req.n_steps)
    np.testing.assert_array_equal(r.image.cpu().numpy(), img)


def _brute_traces(req, dt, division="reciprocal"):
    """The request's E x E x T traces from the oracle's C restatement (every TX
    a Hz point source with the 1 GHz Ricker table, all elements recording Hz),
    dividing as ``division`` does."""
    import ctypes
    from pathlib import Path

    from waveforge_amd.core.materials import coefficients_from_maps
    from waveforge_amd.core.sources import RickerWavelet
    from waveforge_amd.engine import StepCoefficients

    lib = ctypes.CDLL(str(Path(__file__).resolve().parent.parent / "oracle" / "build" / "libwf_brute.so"))
    P = ctypes.c_void_p
    fn = lib.wf_brute_run_f32_rcp if division == "reciprocal" else lib.wf_brute_run_f32
    fn.argtypes = [ctypes.c_int, ctypes.c_int] + [P] * 5 + [ctypes.c_float] * 6 + \
        [ctypes.c_int, ctypes.c_int] + [P] * 3 + [ctypes.c_int] + [P] * 2
    nx, ny = req.eps.shape
    ca, cb = coefficients_from_maps(req.eps, req.sigma, dt)
    ca = ca.reshape(nx, ny).numpy().astype(np.float32).copy()
    cb = cb.reshape(nx, ny).numpy().astype(np.float32).copy()
    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9).build(req.n_steps, dt, torch.device("cpu"))
    vals = np.ascontiguousarray((wave * torch.tensor([1.0])).numpy().astype(np.float32)[None, :])
    c = StepCoefficients.from_grid(dt, 1e-3, 1e-3)
    dcells = np.array([i * ny + j for i, j in req.antennas], dtype=np.int32)
    ptr = lambda a: a.ctypes.data_as(P)
    S = np.zeros((len(req.antennas), len(req.antennas), req.n_steps), dtype=np.float32)
    for t, (i, j) in enumerate(req.antennas):
        ex, ey, hz = (np.zeros((nx, ny), dtype=np.float32) for _ in range(3))
        tr = np.zeros((len(dcells), req.n_steps), dtype=np.float32)
        qx, qy = (c.rdx, c.rdy) if division == "reciprocal" else (c.dx, c.dy)
        fn(nx, ny, ptr(ex), ptr(ey), ptr(hz), ptr(ca), ptr(cb), c.dh, c.de, qx, qy, c.cmx,
           c.cmy, req.n_steps, 1, ptr(np.array([2], np.int32)),
           ptr(np.array([i * ny + j], np.int32)), ptr(vals), len(dcells), ptr(dcells), ptr(tr))
        S[t] = tr
    return S


def test_full_size_mix_alone_equals_mixed_and_oracle(device):
    """The serving workload's own size mix (grids 128^2..1024^2, E in {8, 16,
    32}, steps = 8 x grid diagonal): six requests of the seeded stream plus a
    1024^2 x 32-element one and a 128^2 x 8-element one, served together;
    each equals the same request served alone, bit for bit, and the smallest
    equals the oracle (traces and image)."""
    from waveforge_amd.serving import ImagingServer, make_request, poisson_requests
    from waveforge_amd.workloads import imaging_dt

    dt = imaging_dt()

    def stream():
        reqs = poisson_requests(6, 50.0, n_min=128, n_max=1024, elements=(8, 16, 32), seed=21, steps_per_diag=8.0)
        reqs.append(make_request(6, 0.05, 1024, 32, seed=777))
        reqs.append(make_request(7, 0.06, 128, 8, seed=778))
        return reqs

    reqs = stream()
    for r in reqs:
        print(f"[serving-c4] request {r.rid}: {r.eps.shape[0]}^2 x {len(r.antennas)} elements, {r.n_steps} steps",
              flush=True)
    assert (reqs[6].eps.shape, len(reqs[6].antennas), reqs[6].n_steps) == ((1024, 1024), 32, 11585)
    assert (reqs[7].eps.shape, len(reqs[7].antennas), reqs[7].n_steps) == ((128, 128), 8, 1448)
    srv = ImagingServer(dt=dt, device=device, hbm_budget=64e9, quantum=96)
    stats = srv.serve(reqs, realtime=False)
    assert stats["requests"] == 8 and stats["failed"] == 0
    mixed = {r.rid: (r.traces.clone(), r.image.clone()) for r in srv.done}
    for r in stream():
        alone = ImagingServer(dt=dt, device=device, quantum=96, max_requests=1)
        alone.serve([r], realtime=False)
        a = alone.done[0]
        assert torch.equal(a.traces, mixed[r.rid][0]), f"request {r.rid} traces differ alone vs mixed"
        assert torch.equal(a.image, mixed[r.rid][1]), f"request {r.rid} image differs alone vs mixed"
        print(f"[serving-c4] request {r.rid}: alone == mixed", flush=True)
    small = reqs[7]
    S = _brute_traces(small, dt)
    np.testing.assert_array_equal(mixed[7][0].cpu().numpy().view(np.int32), S.view(np.int32))
    nx, ny = small.eps.shape
    img = das_ref.das_backprojection(S, small.antennas, nx, ny, 1e-3, dt, small.n_steps)
    np.testing.assert_array_equal(mixed[7][1].cpu().numpy(), img)
    assert float(np.abs(S).max()) > 0


def _field_max(req, n_steps, dt, device):
    """max |Ex|, |Ey|, |Hz| over the request's TX runs after n_steps (the plain
    path's set-up, no stability checks)."""
    from waveforge_amd.core import FDTD2D, FieldSet, MurABC, PointSource, SourceCollection, YeeGrid
    from waveforge_amd.core.materials import coefficients_from_maps
    from waveforge_amd.core.sources import RickerWavelet

    nx, ny = req.eps.shape
    grid = YeeGrid(nx, ny, 1e-3, 1e-3, courant=0.99, dt=dt, device=device)
    ca, cb = coefficients_from_maps(req.eps, req.sigma, dt, device=device)
    m = 0.0
    for i, j in req.antennas:
        f = FieldSet(grid)
        src = PointSource(RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9), i, j, "Hz", grid=grid,
                          N_steps=n_steps)
        sim = FDTD2D(grid, f, MurABC(grid, f.Hz), SourceCollection([src]), Ca=ca, Cb=cb, n_check=10 ** 9)
        sim.run(n_steps)
        m = max(m, *(float(getattr(f, k).abs().max()) for k in ("Ex", "Ey", "Hz")))
    return m


def test_last_round_does_not_run_past_n_steps(device):
    """A request of 100 steps (not a multiple of the 48-step quantum) whose
    fields first exceed the stability threshold after its last step. The
    server stops every run at its own n_steps inside the last round and
    judges that state (rounds end at steps 48, 96, then 100), so the request
    is not failed, as on the plain path (checked at step 100); a whole last
    round (to step 144) would cross the threshold. Traces equal the plain path's."""
    import math

    from waveforge_amd.serving import ImagingServer, make_request, serve_plain
    from waveforge_amd.workloads import imaging_dt

    dt = imaging_dt()

    def req():
        r = make_request(0, 0.0, 64, 4, seed=5, steps_per_diag=1.0)
        r.n_steps = 100
        return r

    judged = max(_field_max(req(), n, dt, device) for n in (48, 96, 100))
    past = _field_max(req(), 144, dt, device)
    print(f"[serving] max|field| at steps 48/96/100: {judged:.3e}; at step 144: {past:.3e}", flush=True)
    assert past > 1.01 * judged
    thr = math.sqrt(judged * past)
    served = req()
    ImagingServer(dt=dt, device=device, quantum=48, depth=6, stability_threshold=thr).serve([served], realtime=False)
    plain = req()
    serve_plain([plain], dt=dt, device=device, realtime=False, n_check=100, stability_threshold=thr)
    assert not served.failed and not plain.failed
    assert torch.equal(served.traces, plain.traces) and torch.equal(served.image, plain.image)
    # the same request run one round further does fail: the threshold bites past step 100
    longer = req()
    longer.n_steps = 144
    ImagingServer(dt=dt, device=device, quantum=48, depth=6, stability_threshold=thr).serve([longer], realtime=False)
    assert longer.failed
