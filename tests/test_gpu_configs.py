"""The named workloads at their full sizes, end to end.

Each test checks the fp32 product against the oracle bit for bit wherever
the oracle finishes in seconds, checks size-independent properties where it
does not, and states the fp32 result's error against an FP64 run of the same
simulation (the oracle in float64 on the CPU, or the fp64 GPU step, which is
itself pinned bit for bit to that oracle).

* 256x256 dielectric slab, 1000 steps (Gaussian-derivative Hz pulse at
  (64, 128), detectors (100, 128) and (200, 128), dt = 0.99 dx/(c sqrt 2)).
* 4096x4096 blob medium, 4000 steps (64 Gaussian blobs, sigma = 0.02
  (eps_r - 1), 1 GHz modulated Gaussian at the centre, 64 detectors).
* Batched MIMO imaging: one head phantom of each size {150, 200, 256, 384}^2,
  16 TX x 16 RX, 2000 steps, then DAS.
* Adjoint: 256^2 head phantom, 16 TX x 16 RX, 3000 steps, checkpoints every
  100 steps, central differences at 32 seeded cells (h = 1e-3 eps_r).
"""

import math

import numpy as np
import pytest
import torch

from cases import assert_bitwise, oracle_from_sim, rel_l2

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[configs] {msg}", flush=True)


def test_slab_256_1000_steps(device):
    from waveforge_amd import workloads
    from waveforge_amd.fp64 import fp64_twin

    wl = workloads.slab_256(device)
    sim = wl.sim
    _log("slab 256^2: product run(1000)")
    sim.run(1000)
    torch.cuda.synchronize()
    o32 = oracle_from_sim(sim, torch.float32, n_check=100)
    o32.run(1000)
    for k in ("Ex", "Ey", "Hz"):
        assert_bitwise(getattr(sim.fields, k), o32.fields[k], f"slab {k}")
    assert_bitwise(wl.detectors.traces, o32.trace_array(), "slab traces")
    o64 = oracle_from_sim(sim, torch.float64)
    o64.run(1000)
    twin = fp64_twin(sim)
    twin.advance(1000)
    torch.cuda.synchronize()
    ex, ey, hz = (t.cpu() for t in twin.fields(0))
    assert torch.equal(hz, o64.Hz[..., 0]) and torch.equal(ex, o64.Ex[..., 0]) and torch.equal(ey, o64.Ey[..., 0])
    assert torch.equal(twin.traces(0).cpu(), o64.trace_array())
    e_tr = rel_l2(wl.detectors.traces.cpu(), o64.trace_array())
    e_hz = rel_l2(sim.fields.Hz.cpu(), o64.Hz)
    _log(f"slab: fp32 vs fp64 rel L2 traces {e_tr:.3e}, Hz {e_hz:.3e}")
    # the field left after 1000 steps is the small remainder the Mur edges did
    # not absorb, so its relative error is larger than the traces'
    assert e_tr < 1e-5 and e_hz < 1e-3
    # the pulse crossed the slab: both detectors saw it
    tr = wl.detectors.traces.cpu()
    assert float(tr[0].abs().max()) > 0 and float(tr[1].abs().max()) > 0


def test_blob_4096_4000_steps(device):
    """The full 4000 steps bit for bit against the oracle's C restatement
    (all rows in parallel on the host), fields and the traces of the 64
    detectors at r = 1800 plus an inner ring of 32 at r = 200 that the pulse
    crosses; then the error against the fp64 run."""
    import ctypes
    from pathlib import Path

    from waveforge_amd import workloads
    from waveforge_amd.fp64 import fp64_twin

    steps = 4000
    _log("blob 4096^2: building the medium (64 detectors at r = 1800, 32 at r = 200)")
    wl = workloads.blob_medium(device, inner_rings=((32, 200.0),))
    sim, det = wl.sim, wl.detectors
    sim.run(steps)
    torch.cuda.synchronize()
    assert sim.n_stability_checks == 40 and sim.launch_shape == (48, 8, 1)
    lib = ctypes.CDLL(str(Path(__file__).resolve().parent.parent / "oracle" / "build" / "libwf_brute.so"))
    P = ctypes.c_void_p
    # the run divides as the reference does on a GPU (x * fp32(1/dx)): the C
    # restatement's reciprocal entry, given fp32(1/dx), fp32(1/dy)
    assert sim._division == "reciprocal"
    fn = lib.wf_brute_run_f32_rcp
    fn.argtypes = [ctypes.c_int, ctypes.c_int] + [P] * 5 + [ctypes.c_float] * 6 + \
        [ctypes.c_int, ctypes.c_int] + [P] * 3 + [ctypes.c_int] + [P] * 2
    o = oracle_from_sim(sim, torch.float32)   # the scalars and tables the oracle uses
    n = 4096
    ex, ey, hz = (np.zeros((n, n), dtype=np.float32) for _ in range(3))
    ca = o.Ca.reshape(n, n).numpy().astype(np.float32).copy()
    cb = o.Cb.reshape(n, n).numpy().astype(np.float32).copy()
    src = o.sources[0]
    vals = np.ascontiguousarray((src.wave[:steps] * src.amp).numpy().astype(np.float32)[None, :])
    cells = np.array([int(src.i_idx[0]) * n + int(src.j_idx[0])], dtype=np.int32)
    dcells = np.array([i * n + j for i, j in det.cells], dtype=np.int32)
    tr = np.zeros((len(dcells), steps), dtype=np.float32)
    ptr = lambda a: a.ctypes.data_as(P)
    _log("blob: oracle (C restatement) for 4000 steps on the host")
    fn(n, n, ptr(ex), ptr(ey), ptr(hz), ptr(ca), ptr(cb), o.Dh, o.De, float(o._rdx), float(o._rdy), o.cx, o.cy,
                         steps, 1, ptr(np.array([2], np.int32)), ptr(cells), ptr(vals), len(dcells), ptr(dcells),
                         ptr(tr))
    del o
    for name, ref in (("Ex", ex), ("Ey", ey), ("Hz", hz)):
        assert_bitwise(getattr(sim.fields, name), torch.from_numpy(ref), f"blob step 4000 {name}")
    assert_bitwise(det.traces, torch.from_numpy(tr), "blob traces, both rings")
    inner = det.traces[64:].abs().max().item()
    _log(f"blob: 4000 steps equal; max |Hz| on the inner ring {inner:.3e}")
    assert inner > 1e-3 * float(np.abs(hz).max())   # the pulse crossed the inner ring
    _log("blob: fp64 run of the same simulation")
    twin = fp64_twin(sim)
    twin.advance(steps)
    torch.cuda.synchronize()
    e_in = rel_l2(det.traces[64:], twin.traces(0)[64:])
    e_hz = rel_l2(sim.fields.Hz.reshape(n, n), twin.fields(0)[2])
    e_ex = rel_l2(sim.fields.Ex.reshape(n, n), twin.fields(0)[0])
    m32, m64 = float(det.traces[:64].abs().max()), float(twin.traces(0)[:64].abs().max())
    hmax = float(twin.fields(0)[2].abs().max())
    _log(f"blob: fp32 vs fp64 rel L2 inner-ring traces {e_in:.3e}, Hz {e_hz:.3e}, Ex {e_ex:.3e}; "
         f"r = 1800 ring max|trace| fp32 {m32:.3e} fp64 {m64:.3e}, max|Hz| {hmax:.3e}")
    assert e_in < 1e-4 and e_hz < 1e-3 and e_ex < 1e-3
    # the pulse has not reached r = 1800 in 4000 steps: the fp64 traces there hold
    # only the scheme's precursor tail, far below fp32 resolution of the field
    assert m64 <= 1e-6 * hmax and m32 <= 1e-6 * hmax


def test_batched_imaging_one_phantom_per_size(device):
    from oracle.das_ref import das_backprojection as das_oracle
    from waveforge_amd import mimo, workloads
    from waveforge_amd.core.sources import RickerWavelet
    from waveforge_amd.fp64 import Fp64Runs
    from waveforge_amd.mimo import RunSpec

    dt = workloads.imaging_dt()
    steps = 2000
    phantoms = [workloads.head_phantom(n, 16, seed=100 + n, clot=True) for n in (150, 200, 256, 384)]
    ph = [(p.eps, p.sigma, p.antennas) for p in phantoms]
    _log("imaging: 64 runs batched with graph replay")
    images, traces, br = mimo.image_phantoms(ph, steps, dt=dt, device=device, use_graph=True, chunk=240)
    torch.cuda.synchronize()
    assert not any(br.status()["diverged"])
    wave64 = RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5 / 1e9).build(steps, dt, device, torch.float64)
    specs = []
    k = 0
    for eps, sig, ants in ph:
        n = eps.shape[0]
        for t in range(len(ants)):
            r = br.runs[k]
            specs.append(RunSpec(n, n, r.ca, r.cb, sources=[("Hz", ants[t][0], ants[t][1], wave64, 1.0)],
                                 receivers=list(ants), n_record=steps))
            k += 1
    f64 = Fp64Runs(specs, dx=1e-3, dy=1e-3, dt=dt, device=device)
    f64.advance(steps)
    torch.cuda.synchronize()
    k = 0
    v = 299_792_458.0 / math.sqrt(40.0)
    for p, (eps, sig, ants) in enumerate(ph):
        n = eps.shape[0]
        S64 = torch.stack([f64.traces(k + t) for t in range(len(ants))])
        k += len(ants)
        e_tr = rel_l2(traces[p], S64)
        img_same = das_oracle(traces[p].cpu().numpy(), ants, n, n, 1e-3, dt, steps, v_medium=v)
        img64 = das_oracle(S64.cpu().numpy(), ants, n, n, 1e-3, dt, steps, v_medium=v)
        got = images[p].cpu().numpy()
        assert np.array_equal(got, img_same), f"phantom {n}: DAS differs from the oracle on the same traces"
        e_img = float(np.linalg.norm(got - img64) / np.linalg.norm(img64))
        peak_same = np.unravel_index(int(got.argmax()), got.shape) == np.unravel_index(int(img64.argmax()),
                                                                                       img64.shape)
        _log(f"imaging {n}^2: fp32 vs fp64 rel L2 traces {e_tr:.3e}, image {e_img:.3e}, peak equal {peak_same}")
        assert e_tr < 1e-4 and e_img < 1e-4 and peak_same


def test_batched_imaging_32_phantoms_512_runs(device):
    """The batched-imaging workload at its real size: 32 seeded head phantoms
    (8 each of 150^2, 200^2, 256^2, 384^2), 16 TX x 16 RX, 2000 steps, all
    512 runs in one batched launch sequence (HIP-graph replay), then 32 DAS
    images. One seeded TX run per grid size equals a standalone FDTD2D run
    bit for bit (fields and traces), and every image equals the oracle's DAS
    of the same traces bit for bit."""
    from oracle.das_ref import das_backprojection as das_oracle
    from waveforge_amd import core, mimo, workloads

    dt = workloads.imaging_dt()
    steps = 2000
    phantoms = workloads.phantom_set(32)
    _log("imaging: 512 runs batched with graph replay")
    images, traces, br = mimo.image_phantoms(phantoms, steps, dt=dt, device=device, use_graph=True, chunk=240)
    torch.cuda.synchronize()
    assert len(br.runs) == 512 and not any(br.status()["diverged"])
    rng = np.random.default_rng(2024)
    wave = core.RickerWavelet(amplitude=1.0, peak_freq=1e9, t0=1.5e-9)
    for p in range(4):  # the first phantom of each size
        eps, sig, ants = phantoms[p]
        n = eps.shape[0]
        tx = int(rng.integers(len(ants)))
        g = core.YeeGrid(n, n, 1e-3, 1e-3, courant=0.99, device=device)
        Ca, Cb = core.coefficients_from_maps(eps, sig, dt, device=device)
        f = core.FieldSet(g)
        det = core.DetectorSet(ants, steps, grid=g)
        src = core.PointSource(wave, ants[tx][0], ants[tx][1], "Hz", grid=g, N_steps=steps)
        sim = core.FDTD2D(g, f, core.MurABC(g, f.Hz), core.SourceCollection([src]), Ca=Ca, Cb=Cb, detectors=det)
        sim.run(steps)
        torch.cuda.synchronize()
        k = 16 * p + tx
        ex, ey, hz = br.fields(k)
        for name, a, b in (("Ex", ex, f.Ex), ("Ey", ey, f.Ey), ("Hz", hz, f.Hz)):
            assert torch.equal(a.reshape(n, n).view(torch.int32), b.reshape(n, n).view(torch.int32)), \
                f"phantom {p} ({n}^2) TX {tx}: {name} differs from the standalone run"
        assert torch.equal(traces[p][tx], det.traces), f"phantom {p} TX {tx}: traces differ"
        _log(f"imaging: phantom {p} ({n}^2) TX {tx} equals its standalone FDTD2D run")
    v = 299_792_458.0 / math.sqrt(40.0)
    for p, (eps, sig, ants) in enumerate(phantoms):
        n = eps.shape[0]
        ref = das_oracle(traces[p].cpu().numpy(), ants, n, n, 1e-3, dt, steps, v_medium=v)
        assert np.array_equal(images[p].cpu().numpy(), ref), f"phantom {p}: DAS differs from the oracle"
    _log("imaging: 32 DAS images equal the oracle's")


def test_adjoint_head_256_3000_steps(device):
    from waveforge_amd import workloads
    from waveforge_amd.adjoint import coefficients, misfit_gradient

    _log("adjoint: building the case (clot data, clot-free model)")
    case = workloads.adjoint_case(device)
    prob, eps, sig, d_obs = case.problem, case.eps, case.sigma, case.d_obs
    dt = prob.dt
    ca64, cb64 = coefficients(eps, sig, dt)
    d32, _ = prob.forward(ca64.float().contiguous(), cb64.float().contiguous())
    d64 = prob.forward_fp64(ca64, cb64)
    e_tr = rel_l2(d32, d64)
    _log(f"adjoint: forward traces fp32 vs fp64 rel L2 {e_tr:.3e}")
    assert e_tr < 1e-5
    _, ge32, gs32 = misfit_gradient(eps, sig, prob, d_obs)
    # every adjoint path gives the same bits: the full state history (the
    # default here), checkpoints + recompute, and one adjoint step per launch
    assert prob.full_history()
    for kw in ({"history": "checkpoint"}, {"history": "checkpoint", "one_step": True}):
        for k, v in kw.items():
            setattr(prob, k, v)
        _, ge_p, gs_p = misfit_gradient(eps, sig, prob, d_obs)
        prob.history, prob.one_step = "auto", False
        assert torch.equal(ge_p, ge32) and torch.equal(gs_p, gs32), f"adjoint path {kw}"
    _log("adjoint: full history == checkpoint/recompute == one step per launch, bit for bit")
    _, ge, gs = misfit_gradient(eps, sig, prob, d_obs, fp64_residual=True)
    ge32, ge, gs = ge32.double().cpu(), ge.double().cpu(), gs.double().cpu()
    d_obs64 = d_obs.double()

    def loss(e, s):
        d = prob.forward_fp64(*coefficients(e, s, dt))
        return 0.5 * float(((d - d_obs64) ** 2).sum())

    n = prob.nx
    rng = np.random.default_rng(12345)
    cells = [(int(rng.integers(2, n - 2)), int(rng.integers(2, n - 2))) for _ in range(32)]
    worst, worst32 = 0.0, 0.0
    scale32 = float(ge32.abs().max())
    for k, (i, j) in enumerate(cells):
        for name, base, grad in (("eps", eps, ge), ("sigma", sig, gs)):
            if name == "sigma" and k >= 8:
                continue
            h = 1e-3 * float(base[i, j])
            p, m = base.clone(), base.clone()
            p[i, j] += h
            m[i, j] -= h
            fd = ((loss(p, sig) - loss(m, sig)) if name == "eps" else (loss(eps, p) - loss(eps, m))) / (2 * h)
            rel = abs(fd - float(grad[i, j])) / abs(fd)
            worst = max(worst, rel)
            if name == "eps":
                worst32 = max(worst32, abs(fd - float(ge32[i, j])) / scale32)
    _log(f"adjoint: 32 eps + 8 sigma cells, worst rel err vs fp64 central differences {worst:.3e} "
         f"(fp32-residual gradient: worst |diff| / max|grad| {worst32:.3e})")
    assert worst < 1e-3
    assert worst32 < 2e-2
