"""Hand-written adjoint vs autograd through the oracle's restatement (fp64)
and vs central differences of the HIP forward.

The loss is 0.5 ||d_sim - d_obs||^2 over all TX x RX traces; gradients are
taken with respect to per-cell eps_r and sigma through Ca/Cb. Checkpoints are
placed so the last segment is shorter than the others.
"""

import math

import numpy as np
import pytest
import torch

from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource, stable_dt

pytestmark = pytest.mark.gpu

EPS0 = 8.8541878128e-12


def _media(nx, ny, seed):
    g = torch.Generator().manual_seed(seed)
    eps = 2.0 + 20.0 * torch.rand((nx, ny), generator=g, dtype=torch.float64)
    sig = 0.5 * torch.rand((nx, ny), generator=g, dtype=torch.float64)
    return eps, sig


def _setup(device, nx=28, ny=24, steps=90, ck=25, depth=5, tx=((6, 5), (20, 17)), rx=((3, 20), (24, 4), (14, 12))):
    from waveforge_amd.adjoint import TraceProblem
    from waveforge_amd.core.sources import RickerWavelet

    dt = stable_dt(1e-3, 1e-3)
    wave = RickerWavelet(1.0, 30e9).build(steps + 4, dt, torch.device("cpu"))
    prob = TraceProblem(nx, ny, 1e-3, dt, list(tx), list(rx), wave.to(device), steps, checkpoint=ck, depth=depth)
    return prob, wave, dt


@pytest.mark.parametrize("rx", [((3, 20), (24, 4), (14, 12)),
                                ((3, 20), (14, 12), (3, 20), (0, 0), (27, 23), (0, 12), (14, 12))])
def test_adjoint_matches_autograd_through_oracle(device, rx):
    """Also with receivers sharing cells and on the Mur faces and corners."""
    from waveforge_amd.adjoint import coefficients, misfit_gradient

    prob, wave, dt = _setup(device, rx=rx)
    eps, sig = _media(prob.nx, prob.ny, 0)
    g = torch.Generator().manual_seed(9)
    d_obs = 1e-3 * torch.randn((len(prob.tx), len(prob.rx), prob.n_steps), generator=g, dtype=torch.float64)
    loss, ge, gs = misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device))

    e = eps.clone().requires_grad_(True)
    s = sig.clone().requires_grad_(True)
    ca, cb = coefficients(e, s, dt)
    d = []
    for (i, j) in prob.tx:
        src = OracleSource("Hz", wave.double(), torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0], dtype=torch.float64))
        o = OracleFDTD2D(prob.nx, prob.ny, 1e-3, 1e-3, dt, Ca=ca.unsqueeze(2), Cb=cb.unsqueeze(2), sources=[src],
                         detectors=list(prob.rx), dtype=torch.float64, n_check=10 ** 9, functional=True)
        o.run(prob.n_steps)
        d.append(o.trace_array())
    d = torch.stack(d)
    ref_loss = 0.5 * ((d - d_obs) ** 2).sum()
    ref_loss.backward()
    ref_l = float(ref_loss.detach())
    assert abs(loss - ref_l) / ref_l < 1e-5
    for got, ref, nm in ((ge, e.grad, "eps"), (gs, s.grad, "sigma")):
        got = got.detach().cpu().double()
        err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
        print(f"[adjoint] d loss / d {nm}: rel L2 vs fp64 autograd = {err:.3e}", flush=True)
        assert err < 1e-4, (nm, err)


def test_adjoint_vs_central_differences_of_the_fp64_forward(device):
    """dL/d eps_r at single cells vs (L(eps+h) - L(eps-h)) / 2h, h = 1e-3 eps_r,
    with L evaluated by the fp64 GPU forward (fp32 rounding would swamp the
    quotient at weakly sensitive cells). The adjoint differentiates the fp32
    forward, so agreement is to fp32 accuracy relative to the gradient's scale."""
    from waveforge_amd.adjoint import coefficients, misfit_gradient

    prob, wave, dt = _setup(device, nx=64, ny=60, steps=260, ck=100, depth=6,
                            tx=((10, 10), (50, 45), (30, 52)), rx=((5, 30), (58, 30), (32, 5), (32, 55)))
    eps, sig = _media(prob.nx, prob.ny, 3)
    eps_d, sig_d = eps.to(device), sig.to(device)
    base = prob.forward_fp64(*coefficients(eps_d, sig_d, dt))
    d_obs = base * 0.8
    _, ge, gs = misfit_gradient(eps_d, sig_d, prob, d_obs)
    ge = ge.double().cpu()
    scale = float(ge.abs().max())

    def loss(e):
        d = prob.forward_fp64(*coefficients(e, sig_d, dt))
        return 0.5 * float(((d - d_obs) ** 2).sum())

    rng = np.random.default_rng(4)
    cells = [(int(rng.integers(8, 56)), int(rng.integers(8, 52))) for _ in range(8)]
    cells.append(tuple(int(v) for v in np.unravel_index(int(ge.abs().argmax()), ge.shape)))
    for (i, j) in cells:
        h = 1e-3 * float(eps[i, j])
        ep, em = eps_d.clone(), eps_d.clone()
        ep[i, j] += h
        em[i, j] -= h
        fd = (loss(ep) - loss(em)) / (2 * h)
        adj = float(ge[i, j])
        tol = 2e-3 * abs(fd) + 1e-5 * scale
        print(f"[adjoint] cell ({i},{j}): FD {fd:.6e} adjoint {adj:.6e} |diff| {abs(fd - adj):.2e} tol {tol:.2e}",
              flush=True)
        assert abs(fd - adj) <= tol, (i, j, fd, adj)


def test_receiver_map_folds_records_bit_for_bit(device):
    """With distinct receiver cells, recording each step's residuals inside the
    previous step's kernel gives the same gradient bits as one record launch
    per step."""
    from waveforge_amd.adjoint import misfit_gradient

    prob, wave, dt = _setup(device, steps=120, ck=50, rx=((3, 20), (24, 4), (14, 12), (0, 5)))
    eps, sig = _media(prob.nx, prob.ny, 1)
    g = torch.Generator().manual_seed(2)
    d_obs = 1e-3 * torch.randn((len(prob.tx), len(prob.rx), prob.n_steps), generator=g, dtype=torch.float64)
    _, ge1, gs1 = misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device))
    prob.receiver_map = False
    _, ge0, gs0 = misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device))
    assert torch.equal(ge1, ge0) and torch.equal(gs1, gs0)


@pytest.mark.parametrize("tile,depth", [(32, 8), (32, 6), (48, 8), (48, 3)])
def test_recompute_history_equals_oracle_states(device, tile, depth):
    """The recompute's state-history launches write the state after every one
    of their steps, bit for bit as the fp32 oracle steps from the same start."""
    from cases import assert_bitwise
    from waveforge_amd.adjoint import coefficients

    prob, wave, dt = _setup(device, nx=70, ny=66, steps=60, ck=30, tx=((6, 5), (50, 40)), rx=((3, 20),))
    prob.recompute_tile_rows, prob.recompute_depth = tile, depth
    eps, sig = _media(prob.nx, prob.ny, 2)
    ca64, cb64 = coefficients(eps.double(), sig.double(), dt)
    ca, cb = ca64.float().contiguous(), cb64.float().contiguous()
    prob.history = "checkpoint"
    _, ckpts = prob.forward(ca.to(device), cb.to(device))
    s0, L = 30, 30
    H = prob._recompute(prob._history(ca.to(device), cb.to(device)), ckpts[1], s0, L)
    torch.cuda.synchronize()
    for r, (i, j) in enumerate(prob.tx):
        src = OracleSource("Hz", wave, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]))
        o = OracleFDTD2D(prob.nx, prob.ny, 1e-3, 1e-3, dt, Ca=ca.unsqueeze(2), Cb=cb.unsqueeze(2), sources=[src],
                         n_check=10 ** 9)
        o.run(s0)
        for k in range(1, L + 1):
            o.step()
            for ci, name in enumerate(("Ex", "Ey", "Hz")):
                assert_bitwise(H[k, r, ci], o.fields[name], f"run {r} step {s0 + k} {name}")


@pytest.mark.parametrize("nx,ny,steps,ck,rx", [
    (75, 131, 121, 45, ((3, 20), (74, 4), (14, 130), (0, 0), (40, 66))),
    (17, 57, 40, 13, ((0, 56), (16, 0), (8, 28), (8, 28))),
    (5, 6, 30, 7, ((4, 5), (1, 1))),
    (40, 40, 24, 24, None),
])
def test_adjoint_launch_forms_give_the_same_bits(device, nx, ny, steps, ck, rx):
    """The adjoint step's launch forms give the same gradient bits: two steps
    per launch with the verified fast division by dx, dy (the default; the
    last step of an odd segment alone) and with IEEE division, one step per
    launch (stages through LDS), and two launches per step;
    also for segments of odd length, grids that are not multiples of any of
    their tiles or smaller than one, receivers on the faces and corners or
    sharing a cell, and no receivers (rx None: the adjoint call gets no
    residuals and starts from a seeded adjoint state instead, the gradient
    of <lam0, u_n>)."""
    from waveforge_amd.adjoint import misfit_gradient

    tx = ((nx // 2, ny // 3), (nx - 2, 1))
    prob, wave, dt = _setup(device, nx=nx, ny=ny, steps=steps, ck=ck, tx=tx, rx=rx or ((1, 1),))
    eps, sig = _media(prob.nx, prob.ny, 5)
    g = torch.Generator().manual_seed(3)
    if rx:
        d_obs = 1e-3 * torch.randn((len(prob.tx), len(prob.rx), prob.n_steps), generator=g, dtype=torch.float64)
        assert prob._divisions() is not None  # the default form takes the fast division
        runs = [misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device))]
        for attr, val in (("fast_div", False), ("one_step", True), ("two_pass", True)):
            old = getattr(prob, attr)
            setattr(prob, attr, val)
            runs.append(misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device)))
            setattr(prob, attr, old)
        grads = [(ge, gs) for _, ge, gs in runs]
    else:
        grads = [_seeded_adjoint(prob, eps, sig, dt, device, flags) for flags in (4, 0, 2, 1)]
    print(f"[adjoint] {nx}x{ny} steps={steps} ck={ck} rx={rx}", flush=True)
    assert bool(grads[0][0].abs().max() > 0) and bool(grads[0][1].abs().max() > 0)
    for k, (ge, gs) in enumerate(grads[1:], 1):
        assert torch.equal(grads[0][0], ge) and torch.equal(grads[0][1], gs), f"launch form {k}"


def _seeded_adjoint(prob, eps, sig, dt, device, flags):
    """dL/dCa, dL/dCb through one wf_adjoint_step call over all of the
    problem's steps (one segment) with no receivers: the adjoint state starts
    from a seeded field instead of trace residuals."""
    import ctypes

    from waveforge_amd import _native
    from waveforge_amd.adjoint import WfAdjArgs, coefficients

    ca64, cb64 = coefficients(eps.double(), sig.double(), dt)
    ca, cb = ca64.float().contiguous().to(device), cb64.float().contiguous().to(device)
    prob.history = "checkpoint"
    _, ckpts = prob.forward(ca, cb)
    n_tx, nx, ny = len(prob.tx), prob.nx, prob.ny
    gen = torch.Generator().manual_seed(11)
    lam = [torch.randn((n_tx, nx, ny), generator=gen).to(device) for _ in range(3)]
    z = lambda dt=torch.float32: torch.zeros((n_tx, nx, ny), dtype=dt, device=device)  # noqa: E731
    lex2, ley2, lhz2, lp = z(), z(), z(), z()
    gca, gcb = z(torch.float64), z(torch.float64)
    hist = prob._history(ca, cb)
    L = prob.n_steps
    prob._recompute(hist, ckpts[0], 0, L)
    H = hist[0]
    a = WfAdjArgs()
    a.ca, a.cb = ca.data_ptr(), cb.data_ptr()
    a.lex, a.ley, a.lhz = (t.data_ptr() for t in lam)
    a.lex2, a.ley2, a.lhz2, a.lp = lex2.data_ptr(), ley2.data_ptr(), lhz2.data_ptr(), lp.data_ptr()
    a.gca, a.gcb = gca.data_ptr(), gcb.data_ptr()
    a.flags = flags
    if flags & _native.WF_ADJ_FAST_DIV:
        (a.div_rx, _, a.div_span_x), (a.div_ry, _, a.div_span_y) = prob._divisions()
    a.run_stride, a.state_stride = nx * ny, 3 * nx * ny
    a.nx, a.ny, a.n_runs, a.n_rx = nx, ny, n_tx, 0
    c = prob.coef
    a.dh, a.dx, a.dy, a.cmx, a.cmy = c.dh, c.dx, c.dy, c.cmx, c.cmy
    step_bytes = n_tx * 3 * nx * ny * 4
    k = L - 1
    a.step, a.n_steps, a.hist_stride = k, L, step_bytes // 4
    a.ex_prev = H.data_ptr() + k * step_bytes
    a.ey_prev = H.data_ptr() + k * step_bytes + nx * ny * 4
    a.hz_next = H.data_ptr() + (k + 1) * step_bytes + 2 * nx * ny * 4
    lib = _native.load()
    _native.check(lib.wf_adjoint_step(ctypes.byref(a), ctypes.c_void_p(_native.stream_handle(device))),
                  "wf_adjoint_step")
    torch.cuda.synchronize()
    return gca.sum(0), gcb.sum(0)

@pytest.mark.parametrize("nx,ny,steps,ck", [(75, 131, 121, 45), (40, 52, 64, 16), (20, 20, 9, 4)])
def test_full_history_equals_checkpoint_recompute(device, nx, ny, steps, ck):
    """The full-history path (one forward writing every step's state, the
    adjoint straight through it) gives the checkpointed path's traces and
    gradients bit for bit, and auto mode picks it within its budget and the
    checkpoints beyond."""
    from waveforge_amd.adjoint import FullHistory, coefficients, misfit_gradient

    prob, wave, dt = _setup(device, nx=nx, ny=ny, steps=steps, ck=ck, tx=((nx // 2, ny // 3), (nx - 3, 2)),
                            rx=((3, 5), (nx - 1, ny // 2), (nx // 2, ny // 3), (0, 0)))
    eps, sig = _media(prob.nx, prob.ny, 7)
    ca64, cb64 = coefficients(eps.double(), sig.double(), dt)
    ca, cb = ca64.float().contiguous().to(device), cb64.float().contiguous().to(device)
    g = torch.Generator().manual_seed(4)
    d_obs = 1e-3 * torch.randn((len(prob.tx), len(prob.rx), prob.n_steps), generator=g, dtype=torch.float64)
    out = {}
    for mode in ("full", "checkpoint"):
        prob.history = mode
        tr, saved = prob.forward(ca, cb)
        assert isinstance(saved, FullHistory) == (mode == "full")
        out[mode] = (tr,) + tuple(misfit_gradient(eps.to(device), sig.to(device), prob, d_obs.to(device))[1:])
    print(f"[adjoint-history] {nx}x{ny} steps={steps} checkpoint={ck}", flush=True)
    for k, name in enumerate(("traces", "dL/deps", "dL/dsigma")):
        assert torch.equal(out["full"][k], out["checkpoint"][k]), name
    assert bool(out["full"][1].abs().max() > 0)
    prob.history = "auto"
    prob.history_budget = prob.history_bytes()
    assert prob.full_history()
    prob.history_budget = prob.history_bytes() - 1
    assert not prob.full_history()
