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


def test_served_request_matches_oracle(device):
    from waveforge_amd.core.sources import RickerWavelet
    from waveforge_amd.serving import ImagingServer
    from waveforge_amd.workloads import imaging_dt
    from waveforge_amd.core.materials import coefficients_from_maps

    dt = imaging_dt()
    req = _stream(n=1, seed=11)[0]
    srv = ImagingServer(dt=dt, device=device, quantum=60, depth=5)
    srv.serve([req], realtime=False)
    r = srv.done[0]
    nx, ny = req.eps.shape
    ca, cb = coefficients_from_maps(req.eps, req.sigma, dt)
    wave = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9).build(req.n_steps + 60, dt, torch.device("cpu"))
    S = np.zeros((len(req.antennas), len(req.antennas), req.n_steps), dtype=np.float32)
    for t, (i, j) in enumerate(req.antennas):
        o = OracleFDTD2D(nx, ny, 1e-3, 1e-3, dt, Ca=ca, Cb=cb, detectors=list(req.antennas), n_check=10 ** 9,
                         sources=[OracleSource("Hz", wave, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]))])
        o.run(req.n_steps)
        S[t] = o.trace_array().numpy()
    np.testing.assert_array_equal(r.traces.cpu().numpy(), S)
    img = das_ref.das_backprojection(S, req.antennas, nx, ny, 1e-3, dt, req.n_steps)
    np.testing.assert_array_equal(r.image.cpu().numpy(), img)
