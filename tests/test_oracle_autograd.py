"""The oracle's functional (autograd-friendly) step: identical to the in-place
restatement bit for bit, and its autograd gradient agrees with central
differences, so it can serve as the gradient reference of the HIP adjoint."""

import torch

from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource, ca_cb_from_maps, stable_dt

NX, NY, DX = 14, 11, 1e-3


def _sources(dtype):
    w = torch.sin(torch.arange(60, dtype=dtype) * 0.31)
    return [
        OracleSource("Hz", w, torch.tensor([5]), torch.tensor([4]), torch.tensor([1.0], dtype=dtype)),
        OracleSource("Ey", 0.3 * w, torch.tensor([3, 4, 0]), torch.tensor([6, 6, 2]),
                     torch.tensor([1.0, 0.5, 2.0], dtype=dtype)),
        OracleSource("Hz", w, torch.tensor([0]), torch.tensor([3]), torch.tensor([1.0], dtype=dtype)),  # edge: inert
    ]


def _run(dtype, functional, ca=None, cb=None, n=40):
    dt = stable_dt(DX, DX)
    o = OracleFDTD2D(NX, NY, DX, DX, dt, Ca=ca, Cb=cb, sources=_sources(dtype), detectors=[(2, 2), (11, 8)],
                     dtype=dtype, n_check=10 ** 9, functional=functional)
    o.run(n)
    return o


def _maps(seed):
    g = torch.Generator().manual_seed(seed)
    eps = 1.5 + 9 * torch.rand((NX, NY), generator=g, dtype=torch.float64)
    sig = 0.4 * torch.rand((NX, NY), generator=g, dtype=torch.float64)
    return eps, sig


def test_functional_step_is_bitwise_the_inplace_step():
    eps, sig = _maps(0)
    ca, cb = ca_cb_from_maps(eps, sig, stable_dt(DX, DX))
    for dtype in (torch.float32, torch.float64):
        for mats in ((None, None), (ca.unsqueeze(2), cb.unsqueeze(2))):
            a = _run(dtype, False, *mats)
            b = _run(dtype, True, *mats)
            for k in ("Ex", "Ey", "Hz"):
                assert torch.equal(a.fields[k], b.fields[k]), (dtype, k)
            assert torch.equal(a.trace_array(), b.trace_array())


def test_functional_autograd_matches_central_differences():
    """d(sum w*traces)/dCa and /dCb at a few cells vs (f(+h)-f(-h))/2h in fp64;
    the loss is linear-quadratic in each coefficient so h = 1e-4*|C| leaves a
    truncation error far below the tolerance."""
    eps, sig = _maps(1)
    ca32, cb32 = ca_cb_from_maps(eps, sig, stable_dt(DX, DX))
    ca0, cb0 = ca32.double(), cb32.double()
    wts = torch.randn((2, 40), generator=torch.Generator().manual_seed(5), dtype=torch.float64)

    def f(ca, cb):
        return float((_run(torch.float64, True, ca.unsqueeze(2), cb.unsqueeze(2)).trace_array() * wts).sum())

    ca = ca0.clone().requires_grad_(True)
    cb = cb0.clone().requires_grad_(True)
    o = _run(torch.float64, True, ca.unsqueeze(2), cb.unsqueeze(2))
    (o.trace_array() * wts).sum().backward()
    for (i, j) in ((6, 5), (1, 1), (12, 9), (3, 8)):
        for which, base, grad in (("Ca", ca0, ca.grad), ("Cb", cb0, cb.grad)):
            h = 1e-4 * abs(float(base[i, j]))
            p, m = base.clone(), base.clone()
            p[i, j] += h
            m[i, j] -= h
            fp = f(p, cb0) if which == "Ca" else f(ca0, p)
            fm = f(m, cb0) if which == "Ca" else f(ca0, m)
            fd = (fp - fm) / (2 * h)
            g = float(grad[i, j])
            assert abs(fd - g) <= 1e-6 * max(abs(g), 1e-30) + 1e-12 * abs(fp), (which, i, j, fd, g)
