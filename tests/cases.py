"""Matched product / oracle set-ups for the parity tests.

A case is a small dict: grid size, time step, optional random Ca/Cb, optional
random initial fields, a list of sources and detector cells. ``build_oracle``
makes the CPU oracle of it, ``build_product`` the HIP-backed ``FDTD2D``; both
receive the same waveform tables, coefficients and initial fields.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from oracle.fdtd2d_ref import OracleFDTD2D, OracleSource, stable_dt

DX = 1e-3


def case(nx, ny, *, steps=40, material=False, init=False, sources=(), detectors=(), seed=0, courant=0.98,
         batch=1, n_check=100, threshold=1e10):
    return dict(nx=nx, ny=ny, steps=steps, material=material, init=init, sources=list(sources),
                detectors=list(detectors), seed=seed, courant=courant, batch=batch, n_check=n_check,
                threshold=threshold)


def _dt(c):
    return stable_dt(DX, DX, c["courant"])


def waveform_table(kind: str, n: int, dt: float) -> torch.Tensor:
    """CPU fp32 waveform tables (same construction as the product's ``Waveform.build``)."""
    t = torch.arange(n, dtype=torch.float32) * dt
    if kind == "gauss":
        sigma = 6 * dt
        return 1.0 * torch.exp(-((t - 30 * dt) ** 2) / (2.0 * sigma ** 2))
    if kind == "dgauss":
        tau = t - 25 * dt
        w = 8 * dt
        return 0.7 * (-(tau / w)) * torch.exp(-(tau ** 2) / (2.0 * w ** 2))
    if kind == "sin":
        return 0.3 * torch.sin(2 * math.pi * 20e9 * t)
    raise ValueError(kind)


def _materials(c, gen):
    if not c["material"]:
        return None, None
    nx, ny = c["nx"], c["ny"]
    dt = _dt(c)
    eps = 1.0 + 30.0 * torch.rand((nx, ny), generator=gen, dtype=torch.float64)
    sig = 2.0 * torch.rand((nx, ny), generator=gen, dtype=torch.float64)
    alpha = sig * dt / (2.0 * 8.8541878128e-12 * eps)
    ca = ((1.0 - alpha) / (1.0 + alpha)).float()
    cb = ((dt / (8.8541878128e-12 * eps)) / (1.0 + alpha)).float()
    return ca.unsqueeze(2), cb.unsqueeze(2)


def _extreme_field(shape, gen):
    """Values spread over binades 2^-149 (subnormal) .. 2^80 with random signs,
    a tenth of them +0 and a tenth -0: exercises every branch of the division
    (subnormal and near-subnormal dividends, both zeros, and, after a step or
    two of growth, dividends above the fast quotient's range)."""
    e = torch.randint(-149, 81, shape, generator=gen).double()
    m = 1.0 + torch.rand(shape, generator=gen, dtype=torch.float64)
    sign = torch.where(torch.rand(shape, generator=gen) < 0.5, -1.0, 1.0).double()
    v = (sign * torch.ldexp(m, e)).float()
    u = torch.rand(shape, generator=gen)
    v = torch.where(u < 0.1, torch.zeros_like(v), v)
    return torch.where((u >= 0.1) & (u < 0.2), torch.full_like(v, -0.0), v)


def _init_fields(c, gen):
    if not c["init"]:
        return None
    shape = (c["nx"], c["ny"], 1) if c["batch"] == 1 else (c["batch"], c["nx"], c["ny"], 1)
    if c["init"] == "extreme":
        return {k: _extreme_field(shape, gen) for k in ("Ex", "Ey", "Hz")}
    return {k: (torch.rand(shape, generator=gen) - 0.5) * s for k, s in (("Ex", 1e-2), ("Ey", 1e-2), ("Hz", 1e-4))}


def _source_cells(spec, nx, ny):
    kind = spec[0]
    if kind == "point":
        _, comp, i, j, wave = spec
        return comp, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]), wave, "cells"
    if kind == "line":
        _, comp, axis, pos, start, stop, wave, amp = spec
        n = stop - start
        if axis == "x":
            i, j = torch.arange(start, stop), torch.full((n,), pos)
        else:
            i, j = torch.full((n,), pos), torch.arange(start, stop)
        # the profile as Python floats (doubles): each side rounds it to its grid dtype
        a = torch.ones(n, dtype=torch.float64) if amp is None else torch.tensor(amp, dtype=torch.float64)
        return comp, i, j, a, wave, "cells"
    if kind == "plane":
        _, comp, plane, pos, wave = spec
        return comp, None, None, None, wave, f"plane_{plane}"
    raise ValueError(kind)


def build_oracle(c, dtype=torch.float32, division="ieee"):
    gen = torch.Generator().manual_seed(c["seed"])
    ca, cb = _materials(c, gen)
    init = _init_fields(c, gen)
    dt = _dt(c)
    n_wave = c["steps"] + 8
    srcs = []
    for spec in c["sources"]:
        comp, i, j, a, wave, kind = _source_cells(spec, c["nx"], c["ny"])
        w = waveform_table(wave, n_wave, dt).to(dtype)
        if kind == "cells":
            srcs.append(OracleSource(comp, w, i.long(), j.long(), a.to(dtype)))
        else:
            srcs.append(OracleSource(comp, w, kind=kind, position=spec[3]))
    o = OracleFDTD2D(c["nx"], c["ny"], DX, DX, dt, Ca=ca, Cb=cb, sources=srcs, detectors=c["detectors"],
                     batch=c["batch"], dtype=dtype, n_check=c["n_check"], stability_threshold=c["threshold"],
                     division=division)
    if init is not None:
        for k, v in init.items():
            o.fields[k].copy_(v.to(dtype))
    return o


def build_product(c, device, *, depth=None, tile=None, n_det_steps=None, fast_div=True, dtype=torch.float32,
                  division="ieee"):
    from waveforge_amd import core
    from waveforge_amd.core.sources import Waveform

    class _Table(Waveform):
        """Waveform returning a fixed CPU table (the test's shared input)."""

        def __init__(self, kind):
            self.kind = kind

        def _compute(self, t):
            raise NotImplementedError

        def build(self, N_steps, dt, device, dtype=torch.float32):
            return waveform_table(self.kind, N_steps, dt).to(device=device, dtype=dtype)

        peak_time = property(lambda self: 0.0)
        bandwidth = property(lambda self: 0.0)
        amplitude = property(lambda self: 1.0)

    gen = torch.Generator().manual_seed(c["seed"])
    ca, cb = _materials(c, gen)
    init = _init_fields(c, gen)
    grid = core.YeeGrid(c["nx"], c["ny"], DX, DX, courant=c["courant"], device=device, dtype=dtype)
    fields = core.FieldSet(grid, batch_size=c["batch"])
    if init is not None:
        for k, v in init.items():
            getattr(fields, k).copy_(v.to(device))
    boundary = core.MurABC(grid, fields.Hz)
    n_wave = c["steps"] + 8
    srcs = []
    for spec in c["sources"]:
        kind = spec[0]
        if kind == "point":
            _, comp, i, j, wave = spec
            srcs.append(core.PointSource(_Table(wave), i, j, comp, grid=grid, N_steps=n_wave))
        elif kind == "line":
            _, comp, axis, pos, start, stop, wave, amp = spec
            srcs.append(core.LineSource(_Table(wave), axis, pos, start, stop, comp, grid=grid, N_steps=n_wave,
                                        amplitude_profile=amp))
        else:
            _, comp, plane, pos, wave = spec
            srcs.append(core.PlaneSource(_Table(wave), plane, pos, comp, grid=grid, N_steps=n_wave))
    coll = core.SourceCollection(srcs) if srcs else None
    det = None
    if c["detectors"]:
        det = core.DetectorSet(c["detectors"], n_det_steps or c["steps"], grid=grid, batch_size=c["batch"])
    kw = {}
    if ca is not None:
        kw = dict(Ca=ca.to(device), Cb=cb.to(device))
    sim = core.FDTD2D(grid, fields, boundary, coll, stability_threshold=c["threshold"], n_check=c["n_check"],
                      detectors=det, depth=depth, tile=tile, fast_div=fast_div, division=division, **kw)
    return sim, det


def assert_bitwise(product: torch.Tensor, oracle: torch.Tensor, what: str) -> None:
    oracle = oracle.detach().cpu()
    p = product.detach().cpu().reshape(oracle.shape)
    same = torch.equal(p, oracle)
    if same and p.dtype == torch.float32:  # bit patterns too: +0 and -0 differ
        same = torch.equal(p.contiguous().view(torch.int32), oracle.contiguous().view(torch.int32))
    if not same:
        diff = (p - oracle).abs()
        idx = int(diff.reshape(-1).argmax())
        raise AssertionError(
            f"{what}: {int((p != oracle).sum())} of {oracle.numel()} values differ; max |diff| "
            f"{float(diff.max()):.3e} at flat index {idx} (product {float(p.reshape(-1)[idx])!r}, "
            f"oracle {float(oracle.reshape(-1)[idx])!r})"
        )


def rel_l2(a, b) -> float:
    a = torch.as_tensor(a, dtype=torch.float64).reshape(-1)
    b = torch.as_tensor(b, dtype=torch.float64).reshape(-1)
    den = float(torch.linalg.norm(b))
    return float(torch.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def oracle_from_sim(sim, dtype=torch.float32, n_check=10 ** 9, division=None):
    """The oracle of a product ``FDTD2D`` built from the same inputs: Ca/Cb as
    stored, point/line sources with the product's fp32 waveform tables (fp32)
    or the tables rebuilt in fp64 by the same ``Waveform`` (fp64, as the
    reference builds them on a float64 grid), detectors at the same cells."""
    g = sim.grid
    srcs = []
    for src in sim._source_list():
        if dtype == torch.float32:
            w = src._waveform_tensor.detach().cpu()
        else:
            w = src._waveform.build(src._waveform_tensor.numel(), g.dt, torch.device("cpu"), dtype)
        srcs.append(OracleSource(src.component, w, src._i_idx.cpu(), src._j_idx.cpu(),
                                 src._amp.detach().cpu().to(dtype)))
    det = sim._detectors
    cells = [] if det is None else [tuple(c) for c in det.cells]
    ca = None if sim._Ca is None else sim._Ca.detach().cpu().reshape(g.Nx, g.Ny, 1)
    cb = None if sim._Cb is None else sim._Cb.detach().cpu().reshape(g.Nx, g.Ny, 1)
    if division is None:  # the product's own division (fp64 runs divide as IEEE)
        division = sim._division if dtype == torch.float32 else "ieee"
    return OracleFDTD2D(g.Nx, g.Ny, g.dx, g.dy, g.dt, Ca=ca, Cb=cb, sources=srcs, detectors=cells, dtype=dtype,
                        n_check=n_check, division=division)
