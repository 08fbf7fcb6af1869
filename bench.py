"""Benchmark: 2D TM FDTD Mcell-updates/s and HBM-roofline fraction on one MI355X.

Workload (default): the 4096x4096 fp32 blob medium — eps_r from 64 seeded
Gaussian blobs (peak 1-60, radius 20-200 cells) over 1, sigma = 0.02 (eps_r - 1),
1 GHz modulated-Gaussian Hz source at the centre, Mur on all edges, 64 Hz
detectors on a radius-1800-cell circle — 4000 timed steps after a warm-up, run
through ``FDTD2D.run`` exactly as a user would (stability check every 100
steps included). A "step" is one leapfrog step of the whole grid; the value is
cell-updates per second over the timed steps, with all inputs resident in HBM.
The run divides by the cell sizes as the reference's expressions do under torch
on a GPU (``division="reciprocal"``, the default of ``FDTD2D``); ``ieee_division``
reports the same timing for the form the reference computes on the CPU.

``roofline``: the dominant kernel is the temporally blocked stepper. Its
algorithmic traffic is 32 B per cell-update (read Ex, Ey, Hz, Ca, Cb, write
Ex, Ey, Hz); ``achieved`` = algorithmic bytes per launch / mean launch time,
both measured live with HIP events on the launch stream around every 4th
launch of the timed region (a uniform sample that keeps the event records'
own cost out of the timed stream; the warm-up runs with the same sampling),
against the 8 TB/s HBM peak. With temporal blocking the kernel moves fewer HBM
bytes than the algorithm nominally needs, so ``frac`` may exceed 1.

``cpu_baseline``: the oracle's torch-CPU restatement of the reference step
(same op sequence as the reference's eager PyTorch path) on this host's
cores, on a bounded sample of the same workload (a few steps of the same grid).

``python bench.py --gpus N --steps K --warmup W``; N > 1 exits: this build
runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s + %HBM roofline (2D TM fp32); p50/p99 per-image latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 32  # material path: read Ex, Ey, Hz, Ca, Cb; write Ex, Ey, Hz
EVENT_STRIDE = 4  # launches of the timed region bracketed by HIP events: every 4th


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4000)
    p.add_argument("--warmup", type=int, default=200)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--depth", type=int, default=None)
    p.add_argument("--tile", type=int, default=None, help="tile rows (tile columns are 64 - 2*depth)")
    p.add_argument("--autotune", action="store_true", help="let the autotuner pick tile rows and depth")
    p.add_argument("--division", default="reciprocal", choices=("reciprocal", "ieee"),
                   help="how the run divides by dx, dy: as torch on a GPU (default) or on the CPU")
    p.add_argument("--no-ieee-compare", action="store_true",
                   help="skip timing the IEEE-division form beside the headline")
    p.add_argument("--cpu-steps", type=int, default=8, help="timed steps of the CPU baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-imaging", action="store_true", help="skip the batched-imaging section")
    p.add_argument("--no-serving", action="store_true", help="skip the serving section")
    p.add_argument("--serving-requests", type=int, default=200)
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient section")
    p.add_argument("--adjoint-history", default="auto", choices=("auto", "full", "checkpoint"),
                   help="the adjoint section's state history (TraceProblem.history) for its main gradient")
    p.add_argument("--no-3d", action="store_true", help="skip the 3D section")
    p.add_argument("--no-das", action="store_true", help="skip the DAS and reference-CPU-case sections")
    p.add_argument("--size3d", type=int, default=256)
    p.add_argument("--steps3d", type=int, default=200)
    p.add_argument("--fd-cells", type=int, default=32, help="seeded cells of the central-difference check")
    p.add_argument("--serving-full", action="store_true",
                   help="run only the serving workload at its full size (2000 Poisson requests, grids 128..1024) "
                        "on the continuous server, with capacity and the plain path sampled; prints one JSON line")
    p.add_argument("--copy-probe", action="store_true",
                   help="run only a plain 512 MiB device copy (3 dispatches): the reference kernel of known "
                        "traffic used to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings; prints its rate")
    return p.parse_args()


def copy_probe() -> int:
    import torch

    n = 512 * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for k in range(3):
        dst.copy_(src)
        ev[k + 1].record()
    torch.cuda.synchronize()
    ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(3)]
    # the copy's read + write rate: the streaming bandwidth a kernel can expect on this stack
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "dispatches": 3,
                      "copy_ms": [round(m, 4) for m in ms],
                      "copy_gb_per_s": round(2 * n * 4 / (min(ms) / 1e3) / 1e9, 1)}))
    return 0


def cpu_model() -> str:
    """The host CPU's model name (``lscpu``'s "Model name", from /proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def cpu_baseline(wl, n_steps: int, slab=None, das=None) -> dict:
    """Time the oracle (torch CPU restatement of the reference step) on a sample
    of the headline run with every host core, and (``single_thread``) on a
    shorter sample with one. With ``slab`` (the reference's own CPU case, 256^2 slab,
    1000 steps) the oracle runs it in full with all cores and with one; with
    ``das`` (synthetic traces of a 384^2 16-element array) it times the NumPy
    restatement of the examples' backprojection onto the full grid and of the
    dataset generator's envelope DAS onto 40 x 40."""
    import numpy as np
    import torch

    from oracle.das_env_ref import compute_das_image
    from oracle.das_ref import das_backprojection as das_np
    from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource

    cores = os.cpu_count() or 1
    model = cpu_model()

    def oracle_for(w):
        g = w.grid
        wave = w.sim._sources._sources[0]._waveform_tensor.cpu()
        i, j = w.source_cell
        return OracleFDTD2D(g.Nx, g.Ny, g.dx, g.dy, g.dt, Ca=w.Ca.cpu(), Cb=w.Cb.cpu(),
                            sources=[OracleSource("Hz", wave, torch.tensor([i]), torch.tensor([j]),
                                                  torch.tensor([1.0]))],
                            detectors=w.detectors.cells)

    def rate(w, steps, warm, threads):
        torch.set_num_threads(threads)
        o = oracle_for(w)
        o.run(warm)
        t0 = time.perf_counter()
        o.run(steps)
        el = time.perf_counter() - t0
        return steps * w.grid.Nx * w.grid.Ny / el / 1e6, el

    g = wl.grid
    value, _ = rate(wl, n_steps, 2, cores)
    single, _ = rate(wl, 2, 1, 1)
    out = {
        "value": value,
        "unit": "Mcell-updates/s",
        "cores": cores,
        "cpu_model": model,
        "kind": "port",
        "sample": f"{g.Nx}x{g.Ny} blob medium, steps 3-{2 + n_steps} ({n_steps} timed steps) of the same run, "
                  f"oracle torch-CPU restatement, {cores} threads",
        "single_thread": {"value": round(single, 2), "unit": "Mcell-updates/s", "cores": 1,
                          "sample": f"{g.Nx}x{g.Ny} blob medium, steps 2-3 (2 timed steps), 1 thread"},
    }
    if slab is not None:
        r_all, s_all = rate(slab, slab.n_steps, 0, cores)
        r_one, s_one = rate(slab, slab.n_steps, 0, 1)
        out["slab256"] = {"workload": slab.description + " (the reference's CPU case), in full",
                     "all_cores": {"threads": cores, "seconds": round(s_all, 3), "mcell_updates_per_s": round(r_all, 1)},
                     "one_core": {"threads": 1, "seconds": round(s_one, 3), "mcell_updates_per_s": round(r_one, 1)}}
    torch.set_num_threads(cores)
    if das is not None:
        S, ants, n, dx, dt, geom = das["S"], das["ants"], das["n"], das["dx"], das["dt"], das["geom"]
        t0 = time.perf_counter()
        das_np(S, ants, n, n, dx, dt, S.shape[2])
        t_bp = time.perf_counter() - t0
        t0 = time.perf_counter()
        compute_das_image(S, ants, n, n, dx, dt, S.shape[2], 40, geom)
        t_env = time.perf_counter() - t0
        gathers = n * n * S.shape[0] * S.shape[1]
        out["das"] = {"sample": f"one {n}x{n} image of {S.shape[0]}x{S.shape[1]}x{S.shape[2]} synthetic traces, "
                                "NumPy restatements",
                      "backprojection_ms": round(t_bp * 1e3, 1),
                      "backprojection_gather_adds_per_s": round(gathers / t_bp, 1),
                      "envelope_40x40_ms": round(t_env * 1e3, 1)}
    return out


def das_case():
    """Synthetic traces (seeded normal, 16 TX x 16 RX x 2000 samples) for the
    16-element array of a 384^2 head phantom, with the phantom's layered-medium
    geometry; DAS timing does not depend on the trace values."""
    import numpy as np

    from waveforge_amd import dataset, workloads

    n = 384
    ph = workloads.head_phantom(n, 16, seed=3)
    S = np.random.default_rng(5).standard_normal((16, 16, 2000)).astype(np.float32)
    c = (n - 1) / 2.0
    geom = {"cx_cells": c, "cy_cells": c, "layers": dataset.head_layers(n)}
    return {"S": S, "ants": ph.antennas, "n": n, "dx": 1e-3, "dt": workloads.imaging_dt(), "geom": geom}


def das_suite(dev, case) -> dict:
    """Both DAS flavours on the GPU for one 384^2 16-element array (16 x 16 x
    2000 traces): the examples' per-pixel backprojection onto the full grid
    (``core.das_backprojection``, fp64 accumulation; H*W*N_tx*N_rx gather-adds
    per image) and the dataset generator's envelope DAS onto a 40 x 40 image
    with layered-medium delays (``antenna.EnvelopeDAS``: Hilbert envelope of
    every trace, then the delay-and-sum). Each is timed per call, traces
    already on the device, median of 10 calls after a warm-up."""
    import torch

    from waveforge_amd import core
    from waveforge_amd.antenna import EnvelopeDAS

    n, dt, dx = case["n"], case["dt"], case["dx"]
    S = torch.from_numpy(case["S"]).to(dev)
    g = core.YeeGrid(n, n, dx, dx, courant=0.99, device=dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(10):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    t_bp = timed(lambda: core.das_backprojection(S, case["ants"], g, S.shape[2], device=dev, return_tensor=True))
    env = EnvelopeDAS(case["ants"], (n, n), dx, dt, 40, case["geom"], device=dev)
    t_env = timed(lambda: env(S))
    gathers = n * n * S.shape[0] * S.shape[1]
    return {
        "workload": f"{n}x{n} grid, 16-element array, 16 x 16 x {S.shape[2]} synthetic traces on the device",
        "backprojection": {"image": [n, n], "ms_per_image": round(t_bp * 1e3, 3),
                           "gather_adds_per_s": round(gathers / t_bp, 1)},
        "envelope_das": {"image": [40, 40], "ms_per_image": round(t_env * 1e3, 3)},
    }


def slab_suite(dev, slab) -> dict:
    """The reference's own CPU case on the GPU: FDTD2D.run over the 256^2 slab
    for all 1000 steps (after one untimed run of a twin), wall time with the
    stability checks every 100 steps included."""
    import torch

    from waveforge_amd import workloads

    workloads.slab_256(dev).sim.run(slab.n_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    slab.sim.run(slab.n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"workload": slab.description, "seconds": round(el, 5),
            "mcell_updates_per_s": round(slab.n_steps * slab.grid.Nx * slab.grid.Ny / el / 1e6, 1)}


def das_batch_timing(phantoms, traces, images, dt, dev, reps: int = 10) -> dict:
    """The 32 DAS images of the imaging workload formed from its traces two
    ways: one ``wf_das2d`` launch per image, and ``wf_das2d_batch`` (32
    images in one launch). Times each with HIP events on the launch stream
    (mean of ``reps`` repetitions after one warm-up) and checks both equal
    the images ``image_phantoms`` returned, bit for bit."""
    import torch

    from waveforge_amd.core.imaging import DasBatch, run_das

    v = 299_792_458.0 / math.sqrt(40.0)
    jobs = []
    for p, (eps, _sig, ants) in enumerate(phantoms):
        nx, ny = eps.shape
        ax = torch.tensor([a[0] * 1e-3 for a in ants], dtype=torch.float64, device=dev)
        ay = torch.tensor([a[1] * 1e-3 for a in ants], dtype=torch.float64, device=dev)
        kw = dict(pitch_x=nx * 1e-3 / nx, pitch_y=ny * 1e-3 / ny, v_medium=v, dt=dt, gate_start=0, clip=True)
        jobs.append((traces[p], ax, ay, kw))
    outs = {m: [torch.empty_like(im) for im in images] for m in ("single", "batch")}

    def single():
        for (S, ax, ay, kw), img in zip(jobs, outs["single"]):
            run_das(S, ax, ay, img, **kw)

    def batch():
        b = DasBatch(dev)
        for (S, ax, ay, kw), img in zip(jobs, outs["batch"]):
            b.add(S, ax, ay, img, **kw)
        b.launch()

    res = {}
    for name, fn in (("single", single), ("batch", batch)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / reps
    same = all(torch.equal(a, c) and torch.equal(b, c)
               for a, b, c in zip(outs["single"], outs["batch"], images))
    n_img = len(images)
    return {
        "launches": {"single": n_img, "batch": -(-n_img // 32)},
        "ms_all_single_launches": round(res["single"], 3),
        "ms_all_batched": round(res["batch"], 3),
        "ms_per_image_single": round(res["single"] / n_img, 4),
        "ms_per_image_batched": round(res["batch"] / n_img, 4),
        "speedup": round(res["single"] / res["batch"], 2),
        "batched_equals_single_bitwise": bool(same),
    }


def imaging_suite(dev) -> dict:
    """Batched MIMO imaging: 32 seeded head phantoms (8 each of 150^2, 200^2,
    256^2, 384^2), 16 TX x 16 RX, 2000 steps, all 512 runs in one batched
    launch sequence (launch shape from the autotuner), then one DAS image per
    phantom; timed with HIP-graph replay of 240-step chunks and stepping
    eagerly, the faster of the two reported as ``batched_s``. Compared with the
    one-run-per-launch GPU paths on the same runs: FDTD2D.run per TX (temporally blocked, one run per launch) and
    FDTD2D.step per TX (three launches per step), each timed on a sample of
    2 runs per grid size and scaled to all 512 runs."""
    import torch

    from waveforge_amd import core, mimo, workloads

    dt = workloads.imaging_dt()
    steps = 2000
    phantoms = workloads.phantom_set(32)
    # warm-up on the same run set and path: division check, code objects, HIP-graph capture, and the
    # launch-shape autotuner's timing for these grids (cached per shape set, so the timed call reuses it)
    mimo.image_phantoms(phantoms, 480, dt=dt, device=dev, tile_rows="auto", depth="auto", use_graph=True, chunk=240)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    images, traces, br = mimo.image_phantoms(phantoms, steps, dt=dt, device=dev, use_graph=True, chunk=240,
                                             tile_rows="auto", depth="auto")
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - t0
    # the same call stepping eagerly (each launch enqueued from the host): the
    # launches are long enough here that the host keeps ahead, and the graph's
    # capture costs more than it saves
    mimo.image_phantoms(phantoms, 480, dt=dt, device=dev, tile_rows="auto", depth="auto")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    images_e, _, _ = mimo.image_phantoms(phantoms, steps, dt=dt, device=dev, tile_rows="auto", depth="auto")
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - t0
    graph_equals_eager = all(torch.equal(a, b) for a, b in zip(images, images_e))
    del images_e
    batched_s = min(graph_s, eager_s)
    cu = br.n_cells * steps
    st = br.status()
    # one-run-per-launch GPU paths on a sample, and alone == in-batch on the sampled runs
    run_s, step_s, same = [], [], True
    wave = core.RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9)
    for p in range(4):
        eps, sig, ants = phantoms[p]
        n = eps.shape[0]
        g = core.YeeGrid(n, n, 1e-3, 1e-3, courant=0.99, device=dev)
        Ca, Cb = core.coefficients_from_maps(eps, sig, dt, device=dev)
        for tx in range(2):
            for mode in ("run", "step"):
                f = core.FieldSet(g)
                det = core.DetectorSet(ants, steps, grid=g)
                src = core.PointSource(wave, ants[tx][0], ants[tx][1], "Hz", grid=g, N_steps=steps)
                sim = core.FDTD2D(g, f, core.MurABC(g, f.Hz), core.SourceCollection([src]), Ca=Ca, Cb=Cb,
                                  detectors=det, n_check=500)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                if mode == "run":
                    sim.run(steps)
                else:
                    for _ in range(steps):
                        sim.step()
                torch.cuda.synchronize()
                el = time.perf_counter() - t1
                (run_s if mode == "run" else step_s).append((n * n, el))
                same = same and torch.equal(det.traces, traces[p][tx])
    def projected(samples):
        per_size = {}
        for cells, el in samples:
            per_size.setdefault(cells, []).append(el)
        return sum(sum(v) / len(v) for v in per_size.values()) * 8 * 16  # 8 phantoms per size, 16 TX each
    plain_run = projected(run_s)
    plain_step = projected(step_s)
    das_t = das_batch_timing(phantoms, traces, images, dt, dev)
    peak = [tuple(int(x) for x in divmod(int(img.argmax()), img.shape[1])) for img in images[:4]]
    return {
        "workload": "32 head phantoms (8 x {150,200,256,384}^2), 16 TX x 16 RX, 2000 steps, 512 runs + 32 DAS images",
        "cell_updates": cu,
        "batched_s": round(batched_s, 4),
        "batched_graph_s": round(graph_s, 4),
        "batched_eager_s": round(eager_s, 4),
        "graph_equals_eager": bool(graph_equals_eager),
        "batched_mcell_updates_per_s": round(cu / batched_s / 1e6, 1),
        "images_per_s": round(32 / batched_s, 2),
        "per_image_latency_ms": round(batched_s * 1e3, 1),
        "plain_run_per_tx_projected_s": round(plain_run, 3),
        "plain_step_per_tx_projected_s": round(plain_step, 3),
        "speedup_vs_run_per_tx": round(plain_run / batched_s, 2),
        "speedup_vs_step_per_tx": round(plain_step / batched_s, 2),
        "launch_shape": [br.tile_rows, br.depth],
        "diverged_runs": int(sum(st["diverged"])),
        "alone_equals_in_batch": bool(same),
        "das_32_images": das_t,
        "sampled_runs": len(run_s),
        "peak_pixels_first4": peak,
    }


def serving_suite(dev, n_requests: int) -> dict:
    """The serving workload's own size mix: imaging requests on head phantoms
    with grids log-uniform in 128^2..1024^2, E in {8, 16, 32} elements (every
    element transmits, all receive), steps = 8 x grid diagonal, DAS each.

    * Capacity, every path on the same 48-request burst (all at t = 0): the
      continuous-batching server; the same server holding one request at a
      time (what continuous batching adds on its own); the plain
      one-run-per-launch GPU path (each TX its own FDTD2D.run, one request at
      a time).
    * Load curves: Poisson streams of ``n_requests`` at 50, 60, 70, 80 and 90 %
      of the continuous server's capacity (p50/p99 latency, arrival to DAS
      image ready), and of 32 requests at 50 and 80 % of the plain path's
      capacity (a rate it sustains); the plain path also on the first 16
      arrivals of the continuous 80 % stream (equal arrival rate: it cannot
      keep up, so its queue and p99 grow).
    * Throughput at equal p99: for each plain load point, the continuous
      server's best load point whose p99 is no worse, and the ratio.
    Throughput is compared in cell-updates/s."""
    import torch

    from waveforge_amd import serving, workloads

    dt = workloads.imaging_dt()
    mix = dict(n_min=128, n_max=1024, elements=(8, 16, 32), steps_per_diag=8.0)

    def reqs(n, rate, seed, burst=False):
        r = serving.poisson_requests(n, rate, seed=seed, **mix)
        if burst:
            for q in r:
                q.arrival = 0.0
        return r

    def server(max_requests=None):
        return serving.ImagingServer(dt=dt, device=dev, hbm_budget=128e9, quantum=192, max_requests=max_requests)

    def log(msg):
        print(f"[bench] serving: {msg}", file=sys.stderr, flush=True)

    warm = serving.poisson_requests(2, 100.0, n_min=64, n_max=96, seed=1)
    server().serve(warm, realtime=False)
    serving.serve_plain(serving.poisson_requests(1, 100.0, n_min=64, n_max=96, seed=1), dt=dt, device=dev,
                        realtime=False)
    torch.cuda.synchronize()
    log("capacity: 48-request burst, continuous")
    cap_c = server().serve(reqs(48, 1.0, 11, burst=True), realtime=False)
    log("capacity: the same burst, one request at a time")
    burst_1 = reqs(48, 1.0, 11, burst=True)
    cap_1 = server(1).serve(burst_1, realtime=False)
    log("capacity: the same burst, plain path")
    burst_p = reqs(48, 1.0, 11, burst=True)
    cap_p = serving.serve_plain(burst_p, dt=dt, device=dev, realtime=False)
    same = True
    ref = {q.rid: q for q in burst_1}
    for q in burst_p:
        same = same and torch.equal(q.image, ref[q.rid].image) and torch.equal(q.traces, ref[q.rid].traces)
    # load curves: each path under Poisson streams at fractions of its own capacity
    curve_c = {}
    for frac in (0.5, 0.6, 0.7, 0.8, 0.9):
        rate = frac * cap_c["images_per_s"]
        log(f"sustained: {n_requests} Poisson requests at {rate:.2f} images/s ({frac:.0%} of capacity), continuous")
        curve_c[frac] = dict(server().serve(reqs(n_requests, rate, 7), realtime=True), offered_images_per_s=rate)
    rate_c = curve_c[0.8]["offered_images_per_s"]
    sus_c = curve_c[0.8]
    log("plain path: the first 16 arrivals of the 80 % stream (equal arrival rate)")
    eq_p = serving.serve_plain(reqs(n_requests, rate_c, 7)[:16], dt=dt, device=dev, realtime=True)
    curve_p = {}
    for frac in (0.5, 0.8):
        rate = frac * cap_p["images_per_s"]
        log(f"plain path: 32 Poisson requests at {rate:.2f} images/s ({frac:.0%} of its capacity)")
        curve_p[frac] = dict(serving.serve_plain(reqs(32, rate, 7), dt=dt, device=dev, realtime=True),
                             offered_images_per_s=rate)
    sus_p = curve_p[0.8]
    ratio_sus = sus_c["mcell_updates_per_s"] / sus_p["mcell_updates_per_s"]
    # throughput at equal p99: for each plain load point, the continuous
    # server's highest-throughput load point whose p99 is no worse
    at_p99 = []
    for fp, p in curve_p.items():
        ok = [(fc, c) for fc, c in curve_c.items() if c["p99_ms"] <= p["p99_ms"]]
        row = {"plain_load": fp, "plain_p99_ms": round(p["p99_ms"], 1),
               "plain_mcell_updates_per_s": round(p["mcell_updates_per_s"], 1)}
        if ok:
            fc, c = max(ok, key=lambda x: x[1]["mcell_updates_per_s"])
            row.update(continuous_load=fc, continuous_p99_ms=round(c["p99_ms"], 1),
                       continuous_mcell_updates_per_s=round(c["mcell_updates_per_s"], 1),
                       throughput_ratio=round(c["mcell_updates_per_s"] / p["mcell_updates_per_s"], 2))
        at_p99.append(row)
    return {
        "workload": f"head phantoms, grid 128..1024 log-uniform, E in {{8,16,32}}, steps = 8*diag, DAS per request; "
                    f"sustained stream of {n_requests} Poisson requests",
        "capacity": {"continuous_48": cap_c, "one_request_at_a_time_48": cap_1, "plain_run_per_launch_48": cap_p},
        "capacity_ratio_vs_plain": round(cap_c["mcell_updates_per_s"] / cap_p["mcell_updates_per_s"], 2),
        "capacity_ratio_vs_one_at_a_time": round(cap_c["mcell_updates_per_s"] / cap_1["mcell_updates_per_s"], 2),
        "sustained": {"continuous": sus_c,
                      "plain_at_equal_arrival_rate_16": dict(eq_p, offered_images_per_s=rate_c),
                      "plain_at_80pct_of_own_capacity_32": sus_p},
        "load_curve_continuous": {f"{k:.0%}": v for k, v in curve_c.items()},
        "load_curve_plain": {f"{k:.0%}": v for k, v in curve_p.items()},
        "sustained_throughput_ratio": round(ratio_sus, 2),
        "p99_continuous_over_plain_at_equal_rate": round(sus_c["p99_ms"] / eq_p["p99_ms"], 3),
        "p99_continuous_over_plain_at_own_rate": round(sus_c["p99_ms"] / sus_p["p99_ms"], 3),
        "throughput_at_equal_p99": at_p99,
        "images_equal_across_paths": bool(same),
    }


def fdtd3d_suite(dev, n: int, steps: int = 200, warmup: int = 20) -> dict:
    """3D Yee stepper (FDTD3D.run) on an n^3 lossy medium: a coupling box with a
    layered sphere (MaterialMap3D), an Ez point source at the centre, Mur on
    all six faces, stability check every 100 steps. Algorithmic traffic 56 B
    per cell-update (read six fields, Ca, Cb; write six fields)."""
    import torch

    from waveforge_amd import core

    g = core.YeeGrid(n, n, 1e-3, 1e-3, Nz=n, dz=1e-3, device=dev)
    m = core.MaterialMap3D(g, default=core.Material("coupling", 20.0, 0.1))
    c = n // 2
    m.add_sphere((c, c, c), 0.40 * n, core.Material("skull", 12.0, 0.2))
    m.add_sphere((c, c, c), 0.36 * n, core.Material("brain", 45.0, 0.8))
    m.add_sphere((c + n // 10, c, c - n // 12), 0.05 * n, core.Material("clot", 60.0, 1.2))
    Ca, Cb = m.build3d()
    f = core.FieldSet(g)
    src = core.PointSource(core.ModulatedGaussian(fc=1e9, sigma=1.0 / (2.0 * 3.141592653589793 * 0.5e9)), c, c,
                           "Ez", k=c, grid=g, N_steps=steps + warmup)
    out = {"workload": f"{n}^3 lossy layered sphere (MaterialMap3D), Ez source, Mur on six faces, {steps} steps"}
    finals = []
    for tag, fused in (("per_stage", False), ("fused", True)):
        f = core.FieldSet(g)
        sim = core.FDTD3D(g, f, core.MurABC3D(g, f.Hx, f.Hy, f.Hz), core.SourceCollection([src]), Ca=Ca, Cb=Cb,
                          fused=fused)
        sim.run(warmup)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        sim.run(steps)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        rate = n ** 3 * steps / (ms / 1e3)
        out[tag] = {"ms_per_step": round(ms / steps, 4), "mcell_updates_per_s": round(rate / 1e6, 1),
                    "algorithmic_gb_per_s": round(rate * 56 / 1e9, 1), "roofline_frac": round(rate * 56 / 8e12, 4)}
        finals.append(f.Ez.clone())
        out["fields_finite"] = bool(all(torch.isfinite(getattr(f, k)).all() for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")))
    out["paths_equal"] = bool(torch.equal(finals[0], finals[1]))
    out["default_path"] = "fused"
    return out


def adjoint_suite(dev, n_fd: int, history: str = "auto") -> dict:
    """Trace-misfit gradient on the 256^2 head phantom, 16 TX x 16 RX, 3000
    steps, checkpoints every 100 steps: d_obs from the clot phantom, the model
    is the clot-free twin. Times the fp32 forward and forward + adjoint
    backward; states the fp32 traces' error against the fp64 GPU run; checks
    dL/d eps_r at ``n_fd`` seeded cells (and dL/d sigma at the first 8)
    against central differences (h = 1e-3 eps_r, 1e-3 sigma) of the fp64 loss."""
    import numpy as np
    import torch

    from waveforge_amd import workloads
    from waveforge_amd.adjoint import coefficients, misfit_gradient

    case = workloads.adjoint_case(dev)
    prob = case.problem
    dt = prob.dt
    eps, sig, d_obs = case.eps, case.sigma, case.d_obs
    n_tx, n = len(prob.tx), prob.nx
    cu = n_tx * n * n * prob.n_steps
    misfit_gradient(eps, sig, prob, d_obs)  # warm-up
    ca64, cb64 = coefficients(eps, sig, dt)
    ca, cb = ca64.float().contiguous(), cb64.float().contiguous()
    torch.cuda.synchronize()
    prob.history = "checkpoint"  # the plain forward: traces and a checkpoint every 100 steps
    t0 = time.perf_counter()
    d32, _ = prob.forward(ca, cb)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    prob.history = history
    full = prob.full_history()
    t0 = time.perf_counter()
    loss, ge, gs = misfit_gradient(eps, sig, prob, d_obs)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    print(f"[bench] adjoint: gradient ({'full history' if full else 'checkpoints'}) {t_grad * 1e3:.1f} ms",
          file=sys.stderr, flush=True)

    def grad_with(**kw):  # the same gradient by another path (warm-up, then timed)
        old = {k: getattr(prob, k) for k in kw}
        for k, v in kw.items():
            setattr(prob, k, v)
        misfit_gradient(eps, sig, prob, d_obs)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, g_e, g_s = misfit_gradient(eps, sig, prob, d_obs)
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        for k, v in old.items():
            setattr(prob, k, v)
        return el, bool(torch.equal(ge, g_e) and torch.equal(gs, g_s))


    # checkpoints + recompute with two-step adjoint launches, and with one step a launch
    if history == "checkpoint":  # the main gradient is already the checkpointed one
        t_grad_ck, same_ck = t_grad, True
        t_grad_one, same_one = grad_with(one_step=True)
    else:
        t_grad_ck, same_ck = grad_with(history="checkpoint")
        t_grad_one, same_one = grad_with(history="checkpoint", one_step=True)
    print("[bench] adjoint: checkpointed and one-step gradients done", file=sys.stderr, flush=True)
    out = {"workload": "256^2 head phantom (clot-free model, clot data), 16 TX x 16 RX, 3000 steps, "
                       "loss 0.5||d_sim - d_obs||^2",
           "loss": loss, "forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
           "gradient_path": ("full state history (%.1f GB), two adjoint steps per launch" % (prob.history_bytes() / 1e9)
                             if full else "checkpoints every 100 steps + recompute, two adjoint steps per launch"),
           "gradient_checkpoint_recompute_s": round(t_grad_ck, 4),
           "gradient_checkpoint_one_step_per_launch_s": round(t_grad_one, 4),
           "gradients_equal_bitwise_across_paths": bool(same_ck and same_one),
           "forward_mcell_updates_per_s": round(cu / t_fwd / 1e6, 1),
           "gradient_cell_updates": 3 * cu, "gradient_mcell_updates_per_s": round(3 * cu / t_grad / 1e6, 1)}
    if n_fd <= 0:  # the gradient paths only (a profiling form): no fp64 runs, no central differences
        return out
    t0 = time.perf_counter()
    loss64, ge64, gs64 = misfit_gradient(eps, sig, prob, d_obs, fp64_residual=True)
    torch.cuda.synchronize()
    t_grad64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    d64 = prob.forward_fp64(ca64, cb64)
    torch.cuda.synchronize()
    t_f64 = time.perf_counter() - t0
    fp64_err = float(torch.linalg.norm(d32.double() - d64) / torch.linalg.norm(d64))
    ge, gs = ge.double().cpu(), gs.double().cpu()
    ge64, gs64 = ge64.double().cpu(), gs64.double().cpu()
    resid_rel = float(torch.linalg.norm(d64 - d_obs.double()) / torch.linalg.norm(d64))
    d_obs64 = d_obs.double()

    def fd_loss(e, s):
        d = prob.forward_fp64(*coefficients(e, s, dt))
        return 0.5 * float(((d - d_obs64) ** 2).sum())

    rng = np.random.default_rng(12345)
    cells = [(int(rng.integers(2, n - 2)), int(rng.integers(2, n - 2))) for _ in range(n_fd)]
    checks = []
    t0 = time.perf_counter()
    print(f"[bench] adjoint: fp64 runs done; central differences at {n_fd} cells", file=sys.stderr, flush=True)
    for k, (i, j) in enumerate(cells):
        for name, base in (("eps", eps), ("sigma", sig)):
            if name == "sigma" and k >= 8:
                continue
            h = 1e-3 * float(base[i, j])
            p, m = base.clone(), base.clone()
            p[i, j] += h
            m[i, j] -= h
            lp = fd_loss(p, sig) if name == "eps" else fd_loss(eps, p)
            lm = fd_loss(m, sig) if name == "eps" else fd_loss(eps, m)
            checks.append((name, i, j, (lp - lm) / (2 * h)))
    t_fd = time.perf_counter() - t0
    out.update({"traces_rel_l2_vs_fp64": fp64_err, "residual_rel_l2": resid_rel,
                "gradient_fp64_residual_s": round(t_grad64, 4), "loss_fp64_residual": loss64,
                "fp64_forward_s": round(t_f64, 3), "fd_s": round(t_fd, 2)})
    for tag, grads in (("fp32", (ge, gs)), ("fp64_residual", (ge64, gs64))):
        for name, grad in zip(("eps", "sigma"), grads):
            rows = [(fd, float(grad[i, j])) for (nm, i, j, fd) in checks if nm == name]
            scale = float(grad.abs().max())
            rel = [abs(fd - ad) / max(abs(fd), 1e-300) for fd, ad in rows]
            out[f"fd_{name}_{tag}"] = {
                "cells": len(rows),
                "max_rel_err": max(rel),
                "median_rel_err": float(np.median(rel)),
                "max_abs_diff_over_grad_max": max(abs(fd - ad) for fd, ad in rows) / scale,
            }
    return out


def serving_full() -> int:
    """The serving workload at its full size: 2000 imaging requests (head
    phantoms, grid log-uniform in 128..1024, E in {8, 16, 32}, steps = 8 x grid
    diagonal, DAS each), Poisson arrivals at 80 % of the continuous server's
    capacity, served in real time. Capacity comes from a 64-request burst of
    the same mix (all at t = 0), served by both the continuous server and the
    plain one-run-per-launch path (the plain path's full stream would take
    hours); the capacity ratio is taken on that same burst."""
    import torch

    from waveforge_amd import serving, workloads

    dev = torch.device("cuda:0")
    dt = workloads.imaging_dt()

    def log(msg):
        print(f"[serving-full] {msg}", file=sys.stderr, flush=True)

    def server():
        return serving.ImagingServer(dt=dt, device=dev, hbm_budget=192e9, quantum=192)

    warm = serving.poisson_requests(2, 100.0, n_min=64, n_max=96, seed=1)
    server().serve(warm, realtime=False)
    serving.serve_plain(serving.poisson_requests(1, 100.0, n_min=64, n_max=96, seed=1), dt=dt, device=dev,
                        realtime=False)
    torch.cuda.synchronize()
    burst = serving.poisson_requests(64, 1.0, n_min=128, n_max=1024, seed=11)
    for q in burst:
        q.arrival = 0.0
    log("capacity burst: 64 requests")
    cap = server().serve(burst, realtime=False, progress=lambda d, a, t: log(f"burst {d} done, {t:.0f} s"))
    log(f"capacity {cap['images_per_s']:.2f} images/s, {cap['mcell_updates_per_s'] / 1e3:.1f} G cell-updates/s")
    plain = serving.poisson_requests(64, 1.0, n_min=128, n_max=1024, seed=11)
    for q in plain:
        q.arrival = 0.0
    log("plain path: the same 64 requests")
    cap_p = serving.serve_plain(plain, dt=dt, device=dev, realtime=False,
                                progress=lambda d, t: log(f"plain {d} done, {t:.0f} s"))
    log(f"plain {cap_p['mcell_updates_per_s'] / 1e3:.1f} G cell-updates/s")
    same = all(torch.equal(a.image, b.image) for a, b in zip(sorted(burst, key=lambda r: r.rid),
                                                             sorted(plain, key=lambda r: r.rid)))
    rate = 0.8 * cap["images_per_s"]
    stream = serving.poisson_requests(2000, rate, n_min=128, n_max=1024, seed=7)
    work = sum(q.cell_updates for q in stream)
    log(f"stream: 2000 requests at {rate:.2f} images/s, {work:.3e} cell-updates")
    sus = server().serve(stream, realtime=True,
                         progress=lambda d, a, t: log(f"stream {d} done, {a} admitted, {t:.0f} s"))
    print(json.dumps({
        "serving_full": {
            "workload": "2000 requests, head phantoms, grid 128..1024 log-uniform, E in {8,16,32}, "
                        "steps = 8*diag, DAS per request, Poisson arrivals at 80 % of capacity",
            "cell_updates": work,
            "capacity_burst_64": cap,
            "plain_run_per_launch_burst_64": cap_p,
            "images_equal_across_paths": bool(same),
            "offered_images_per_s": rate,
            "sustained": sus,
            "capacity_ratio_vs_plain_same_64": round(cap["mcell_updates_per_s"] / cap_p["mcell_updates_per_s"], 2),
        }}))
    return 0


def ieee_compare(dev, args, shape) -> dict:
    """The headline workload with ``division="ieee"`` (the quotient the
    reference computes on the CPU, through the verified fast quotient): the
    same warm-up, then the same number of timed steps (at most 800), same
    launch shape."""
    import torch

    from waveforge_amd import workloads

    steps = min(args.steps, 800)
    wl = workloads.blob_medium(dev, n=args.size, n_steps=args.warmup + steps, division="ieee",
                               tile=(shape[0], shape[2]), depth=shape[1])
    wl.sim.run(args.warmup)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    wl.sim.run(steps)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    res = {"division": "ieee", "steps": steps, "ms_per_step": round(ms / steps, 5),
           "mcell_updates_per_s": round(steps * args.size * args.size / (ms / 1e3) / 1e6, 1)}
    del wl
    torch.cuda.empty_cache()
    return res


def main() -> int:
    args = parse()
    if args.copy_probe:
        return copy_probe()
    if args.serving_full:
        return serving_full()
    if args.gpus != 1:
        print(json.dumps({"error": "waveforge_amd runs on one GPU; multi-GPU scaling is not part of this build",
                          "n_gpus": args.gpus}))
        return 2
    import torch

    from waveforge_amd import workloads
    from waveforge_amd.engine import TBPlan

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no ROCm GPU visible"}))
        return 1
    dev = torch.device("cuda:0")
    kw = {}
    if args.depth:
        kw["depth"] = args.depth
    if args.tile:
        kw["tile"] = args.tile
    if args.autotune:
        kw["depth"], kw["tile"] = "auto", "auto"
    total = args.warmup + args.steps
    print(f"[bench] building {args.size}x{args.size} blob medium ({total} steps)", file=sys.stderr, flush=True)
    wl = workloads.blob_medium(dev, n=args.size, n_steps=total, division=args.division, **kw)
    sim = wl.sim
    # the timing events of the timed region are created and recorded once
    # before it (the first record of a fresh HIP event costs far more than a
    # reuse; hundreds of them stalled the first timed launch by tens of ms)
    # every launch of a short timed region is bracketed (a mean over a few
    # launches is not one sample); a long one brackets every 4th
    from waveforge_amd import autotune
    from waveforge_amd.engine import schedule_launches

    sim._tb_setup()
    _, depth0, _ = sim.launch_shape
    planned, _ = schedule_launches(args.warmup, args.steps, depth0, sim._n_check, odd_ok=True)
    stride = 1 if planned is not None and len(planned) < 16 else EVENT_STRIDE
    n_events = 2 * (args.steps // stride + 64)
    pool = [torch.cuda.Event(enable_timing=True) for _ in range(n_events)]
    for ev in pool:
        ev.record()
    # the launch-shape tuner's table for this grid (reported below), timed on
    # scratch fields before the warm-up: some 60 ms of GPU work that also brings
    # the clock from idle to its working level, so a short timed region does not
    # measure the clock ramp
    print("[bench] autotuner table for this grid", file=sys.stderr, flush=True)
    tuned = autotune.tune_runs(sim._build_runs(), fast_div=sim._fast_div)
    TBPlan.event_log, TBPlan.event_stride = [], stride
    sim.run(args.warmup)
    torch.cuda.synchronize()
    TBPlan.event_pool = pool
    print(f"[bench] warm-up {args.warmup} steps done; timing {args.steps}", file=sys.stderr, flush=True)
    log: list = []
    TBPlan.event_log, TBPlan._event_count = log, 0
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    sim.run(args.steps)
    end.record()
    torch.cuda.synchronize()
    TBPlan.event_log, TBPlan.event_stride, TBPlan.event_pool = None, 1, None
    first_launch_ms = start.elapsed_time(log[0][2]) if log else None
    ms = start.elapsed_time(end)
    cells = args.size * args.size
    value = args.steps * cells / (ms / 1e3) / 1e6
    launch_ms = [e0.elapsed_time(e1) for (_, _, e0, e1) in log]
    launch_bytes = [d * n * BYTES_PER_CELL_UPDATE for (d, n, _, _) in log]
    mean_ms = sum(launch_ms) / max(1, len(launch_ms))
    mean_bytes = sum(launch_bytes) / max(1, len(launch_bytes))
    achieved = mean_bytes / (mean_ms / 1e3) / 1e9 if launch_ms else 0.0
    depths = sorted({d for (d, _, _, _) in log})
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "k_tb (temporally blocked stepper)",
        "launches_timed": len(log),
        "launch_sample": ("every launch of the timed region" if stride == 1
                          else f"every {stride}th launch of the timed region"),
        "mean_launch_us": round(mean_ms * 1e3, 2),
        "first_launch_offset_ms": None if first_launch_ms is None else round(first_launch_ms, 3),
        "bytes_per_launch_algorithmic": mean_bytes,
        "depths": depths,
    }
    if len(log) <= 16:
        roofline["launch_us"] = [[d, round(ms * 1e3, 1)] for (d, _, _, _), ms in zip(log, launch_ms)]
    traffic_file = ROOT / "profiles" / "k_tb_blob4096_traffic.json"
    valu_file = ROOT / "profiles" / "k_tb_blob4096_valu.json"
    tile_rows, depth, cpl = sim.launch_shape
    shape = (args.size, tile_rows, cpl)
    depths_timed = [d for (d, _, _, _) in log]
    dominant = max(set(depths_timed), key=depths_timed.count) if depths_timed else depth
    roofline["dominant_depth"] = dominant
    if traffic_file.exists():
        tinfo = json.loads(traffic_file.read_text())
        by_depth = {int(k): v for k, v in tinfo.get("hbm_bytes_by_depth", {}).items()}
        if (tinfo.get("grid"), tinfo.get("tile_rows"), tinfo.get("cpl", 1)) == shape and dominant in by_depth \
                and tinfo.get("division") == args.division:
            roofline["traffic"] = by_depth[dominant]
            roofline["traffic_source"] = (f"profiles/{tinfo['summary']} (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                          f"copy-kernel calibrated, per depth-{dominant} launch)")
            timed = [(by_depth[d], ms) for (d, _, _, _), ms in zip(log, launch_ms) if d in by_depth]
            # measured HBM bytes of the timed launches over their live launch time
            gbs = sum(b for b, _ in timed) / (sum(ms for _, ms in timed) / 1e3) / 1e9
            roofline["traffic_gb_per_s"] = round(gbs, 1)
            roofline["traffic_frac"] = round(gbs / HBM_PEAK_GBS, 4)
    if valu_file.exists():
        vinfo = json.loads(valu_file.read_text())
        by_depth = vinfo.get("by_depth", {})
        if (vinfo.get("grid"), vinfo.get("tile_rows"), vinfo.get("cpl", 1)) == shape and str(dominant) in by_depth \
                and vinfo.get("division") == args.division:
            roofline["valu"] = dict(by_depth[str(dominant)], depth=dominant, source="profiles/" + vinfo["summary"])
    tr = wl.detectors.traces
    finite = bool(torch.isfinite(tr).all()) and bool(torch.isfinite(sim.fields.Hz).all())
    ieee = None
    if args.division == "reciprocal" and not args.no_ieee_compare:
        print("[bench] the same run with IEEE division", file=sys.stderr, flush=True)
        ieee = ieee_compare(dev, args, sim.launch_shape)
    base = das_res = slab_res = None
    dcase = das_case()
    if not args.no_das:
        print("[bench] DAS and reference-CPU-case sections", file=sys.stderr, flush=True)
        das_res = das_suite(dev, dcase)
        slab_res = slab_suite(dev, workloads.slab_256(dev))
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline: {args.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        base = cpu_baseline(wl, args.cpu_steps, slab=workloads.slab_256(dev), das=dcase)
    imaging = serving_res = adjoint = fd3 = None
    if not args.no_imaging:
        print("[bench] batched imaging section", file=sys.stderr, flush=True)
        imaging = imaging_suite(dev)
    if not args.no_serving:
        print("[bench] serving section", file=sys.stderr, flush=True)
        serving_res = serving_suite(dev, args.serving_requests)
    if not args.no_adjoint:
        print("[bench] adjoint section", file=sys.stderr, flush=True)
        adjoint = adjoint_suite(dev, args.fd_cells, args.adjoint_history)
    if not args.no_3d:
        print("[bench] 3D section", file=sys.stderr, flush=True)
        fd3 = fdtd3d_suite(dev, args.size3d, steps=args.steps3d)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded blob medium)",
        "config": {
            "workload": f"blob medium {args.size}x{args.size} fp32, 64 blobs, sigma=0.02(eps_r-1), "
                        f"1 GHz modulated Gaussian at centre, Mur, 64 detectors",
            "grid": [args.size, args.size],
            "steps": args.steps,
            "tile_rows": tile_rows,
            "depth": depth,
            "cols_per_lane": cpl,
            "division": args.division,
        },
        "roofline": roofline,
        "autotune": {"choice": [tuned.tile_rows, tuned.depth, tuned.cpl], "used": [tile_rows, depth, cpl],
                     "rates_mcell_per_s": tuned.rates},
        "cpu_baseline": base,
        "fields_finite": finite,
        "ieee_division": ieee,
        "das": das_res,
        "slab256": slab_res,
        "imaging": imaging,
        "serving": serving_res,
        "adjoint": adjoint,
        "fdtd3d": fd3,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
