"""The launch-shape autotuner: an admissible choice, cached per shape set,
and results that do not depend on the launch shape it picks (bit for bit
against the oracle and against the default shape)."""

import pytest
import torch

from cases import assert_bitwise, build_oracle, build_product, case

pytestmark = pytest.mark.gpu


def test_tuner_times_every_candidate_and_caches(device, tmp_path):
    from waveforge_amd import _native, autotune

    c = case(200, 300, steps=30, material=True, sources=[("point", "Hz", 50, 60, "gauss")], detectors=[(3, 4)])
    sim, _ = build_product(c, device, depth="auto", tile="auto")
    sim.run(30)
    res = sim._tuned
    assert sim.launch_shape == (res.tile_rows, res.depth, res.cpl)
    assert len(res.rates) == len(_native.tile_shapes()) * 8
    assert max(res.rates.values()) == res.rates[autotune.shape_key(res.tile_rows, res.cpl, res.depth)]
    again = autotune.tune_runs(sim._build_runs(), fast_div=True)
    assert again is res
    path = tmp_path / "tune.json"
    autotune.save(path)
    n = len(autotune._CACHE)
    autotune._CACHE.clear()
    assert autotune.load(path) == n
    assert autotune.tune_runs(sim._build_runs(), fast_div=True).as_dict() == res.as_dict()


def test_tuned_run_equals_oracle_and_default_shape(device):
    c = case(97, 130, steps=57, material=True, init=True,
             sources=[("point", "Hz", 40, 70, "gauss"), ("point", "Ey", 3, 100, "sin")],
             detectors=[(1, 1), (50, 60), (95, 128)])
    auto, det_a = build_product(c, device, depth="auto", tile="auto")
    auto.run(57)
    ref, det_r = build_product(c, device)
    ref.run(57)
    o = build_oracle(c)
    o.run(57)
    torch.cuda.synchronize()
    print(f"[autotune] picked {auto.launch_shape}", flush=True)
    for k in ("Ex", "Ey", "Hz"):
        assert_bitwise(getattr(auto.fields, k), o.fields[k], f"auto {k}")
        assert torch.equal(getattr(auto.fields, k), getattr(ref.fields, k))
    assert torch.equal(det_a.traces, det_r.traces)


def test_batched_runs_auto_shape(device):
    from waveforge_amd.mimo import BatchedRuns
    from test_gpu_batched import _specs
    from oracle.fdtd2d_ref import stable_dt

    dt = stable_dt(1e-3, 1e-3)
    specs, oracles = _specs(device, 40, dt)
    br = BatchedRuns(specs, dx=1e-3, dy=1e-3, dt=dt, device=device, tile_rows="auto", depth="auto")
    assert br.tuned is not None
    assert (br.tile_rows, br.depth, br.cpl) == (br.tuned.tile_rows, br.tuned.depth, br.tuned.cpl)
    br.advance(40)
    torch.cuda.synchronize()
    for k, o in enumerate(oracles):
        o.run(40)
        assert_bitwise(br.fields(k)[2], o.Hz, f"run {k} Hz")
        assert_bitwise(br.traces(k), o.trace_array(), f"run {k} traces")
