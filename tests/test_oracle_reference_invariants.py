"""The reference's own 2D test assertions (``tests/test_fdtd2d.py``), restated
against the CPU oracle, plus closed-form checks that pin the oracle's physics.

The reference ships no golden vectors; its tests are sign, bound and
known-answer checks, each cited below. The GPU product runs the same list in
``test_gpu_reference_suite.py``.
"""

import math

import numpy as np
import pytest
import torch

from oracle.fdtd2d_ref import (
    C0,
    EPS0,
    MU0,
    OracleDivergedError,
    OracleFDTD2D,
    OracleSource,
    ca_cb_from_maps,
    material_ca_cb,
    mur_coefficients,
    stable_dt,
)
from oracle import waveforms_ref as wr

DX = 1e-3


def sim(n, dt=None, **kw):
    dt = stable_dt(DX, DX) if dt is None else dt
    return OracleFDTD2D(n, n, DX, DX, dt, **kw)


def point(n_steps, dt, i, j, sigma_steps, amp=1.0, comp="Hz"):
    w = wr.gaussian(wr.time_axis(n_steps, dt), amp, 5 * sigma_steps * dt, sigma_steps * dt)
    return OracleSource(comp, w, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0]))


def test_cfl_dt_is_stable():  # reference tests/test_fdtd2d.py:57-63
    dt = stable_dt(DX, DX)
    assert dt <= 1.0 / (C0 * math.sqrt(2 / DX ** 2)) * 1.001
    assert math.isclose(stable_dt(DX, DX, 0.99), 0.99 * DX / (C0 * math.sqrt(2.0)), rel_tol=1e-15)


def test_mur_coefficient_in_open_interval():  # :83-89
    cx, cy = mur_coefficients(stable_dt(DX, DX), DX, DX)
    assert -1.0 < cx < 0.0 and -1.0 < cy < 0.0
    # closed form at courant 0.99: C = (0.99/sqrt2 - 1)/(0.99/sqrt2 + 1)
    c99 = mur_coefficients(stable_dt(DX, DX, 0.99), DX, DX)[0]
    k = 0.99 / math.sqrt(2.0)
    assert math.isclose(c99, (k - 1) / (k + 1), rel_tol=1e-12)
    assert abs(c99 - (-0.176446)) < 1e-6


def test_fields_start_at_zero_energy():  # :97-107
    o = sim(32)
    for k in ("Ex", "Ey", "Hz"):
        assert float(o.fields[k].abs().max()) == 0.0


def test_free_space_stable_500_steps():  # :130-140
    o = sim(32, n_check=10)
    o.sources = [point(600, o.dt, 16, 16, 20, amp=0.01)]
    o.run(500)
    assert o.last_field_max < 1e6


def test_divergence_error_raised():  # :143-156
    o = OracleFDTD2D(8, 8, DX, DX, stable_dt(DX, DX), n_check=1, stability_threshold=1.0)
    o.Hz[:] = 2.0
    with pytest.raises(OracleDivergedError):
        o.step()


def test_faraday_sign():  # :179-197
    o = OracleFDTD2D(8, 8, DX, DX, stable_dt(DX, DX), n_check=1000)
    o.Ex[:, 4, 0] = 1.0
    o.step()
    assert float(o.Hz[3, 3, 0]) > 0.0


def test_ampere_ex_sign():  # :200-217
    o = OracleFDTD2D(8, 8, DX, DX, stable_dt(DX, DX), n_check=1000)
    o.Hz[:, 4, 0] = 1.0
    o.step()
    assert float(o.Ex[3, 4, 0]) > 0.0


def test_ampere_ey_sign():  # :220-237
    o = OracleFDTD2D(8, 8, DX, DX, stable_dt(DX, DX), n_check=1000)
    o.Hz[4, :, 0] = 1.0
    o.step()
    assert float(o.Ey[4, 3, 0]) < 0.0


def _energy(o):  # reference FieldSet energy constants (src/core/fields.py:77-78)
    e = (o.Ex.double() ** 2 + o.Ey.double() ** 2).sum() * (8.854e-12 * 0.5)
    h = (o.Hz.double() ** 2).sum() * (1.2566e-6 * 0.5)
    return float(e + h) * DX ** 3


def test_energy_increases_with_source():  # :240-254
    o = sim(32, n_check=1000)
    o.sources = [point(200, o.dt, 16, 16, 30)]
    e0 = _energy(o)
    o.run(50)
    assert _energy(o) > e0


def test_pulse_propagates_outward():  # :262-276
    o = sim(64, n_check=1000)
    o.sources = [point(300, o.dt, 32, 32, 20)]
    o.run(100)
    assert float(o.Hz[32 - 22:32 - 18, 32, 0].abs().sum()) > 0.0


def test_mur_absorbs_outgoing():  # :279-297
    o = sim(64, n_check=1000)
    o.sources = [point(600, o.dt, 32, 32, 15)]
    o.run(80)
    p1 = float(o.Hz.abs().max())
    o.run(300)
    assert float(o.Hz.abs().max()) < p1


def test_point_injection_touches_one_cell():  # :348-375
    o = sim(32, n_check=1000)
    src = point(200, o.dt, 10, 10, 50)
    src.inject(o.fields, 199)
    assert float(o.Hz[10, 10, 0]) != 0.0 and float(o.Hz[10, 11, 0]) == 0.0
    nz = int((o.Hz != 0).sum())
    assert nz == 1


def test_material_coefficients_closed_form():
    """Ca/Cb of the lossy update (src/core/materials.py:61-82) at the head media
    values quoted for dt = 2.335068e-12 s."""
    dt = 0.99 * DX / (C0 * math.sqrt(2.0))
    assert math.isclose(dt, 2.335068e-12, rel_tol=1e-6)
    expect = {(20.0, 0.1): (0.9986822, 1.31775e-2), (12.0, 0.2): (0.9956142, 2.19289e-2),
              (45.0, 0.8): (0.9953225, 5.84684e-3), (60.0, 1.2): (0.9947394, 4.38385e-3)}
    for (eps, sig), (ca, cb) in expect.items():
        a, b = material_ca_cb(eps, sig, dt)
        assert abs(a - ca) < 2e-7 and abs(b - cb) / cb < 2e-5
    ca_t, cb_t = ca_cb_from_maps(torch.tensor([[45.0]]), torch.tensor([[0.8]]), dt)
    assert ca_t.dtype == torch.float32 and float(ca_t) == np.float32(material_ca_cb(45.0, 0.8, dt)[0])


def test_lossless_medium_reduces_to_free_space_update():
    """Ca = 1, Cb = dt/eps0 is the free-space update up to fp32 rounding."""
    n = 24
    dt = stable_dt(DX, DX)
    ca = torch.ones(n, n, 1)
    cb = torch.full((n, n, 1), dt / EPS0)
    a = sim(n, Ca=ca, Cb=cb, n_check=10 ** 9)
    b = sim(n, n_check=10 ** 9)
    for o in (a, b):
        o.sources = [point(100, dt, 12, 11, 6)]
        o.run(60)
    assert torch.allclose(a.Hz, b.Hz, rtol=1e-5, atol=1e-9)


def test_mirror_symmetry_about_the_diagonal():
    """A centred Hz source in a square vacuum grid: the Yee scheme maps onto
    itself under i <-> j with Ex <-> -Ey, so Hz = Hz^T and Ex = -Ey^T exactly
    (fp64). A transposed operand or a sign slip in only one of the curl
    updates, or in one Mur edge, breaks the identity."""
    n = 41
    dt = stable_dt(DX, DX)
    w = wr.gaussian(wr.time_axis(200, dt, torch.float64), 1.0, 30 * dt, 6 * dt)
    src = OracleSource("Hz", w, torch.tensor([20]), torch.tensor([20]), torch.tensor([1.0], dtype=torch.float64))
    o = OracleFDTD2D(n, n, DX, DX, dt, sources=[src], dtype=torch.float64, n_check=10 ** 9)
    o.run(120)  # long enough for the wave to reach and reflect off every Mur edge
    Hz, Ex, Ey = o.Hz[..., 0], o.Ex[..., 0], o.Ey[..., 0]
    assert float(Hz[0].abs().max()) > 0.0
    assert torch.equal(Hz, Hz.T)
    assert torch.equal(Ex, -Ey.T)


def test_superposition_of_sources():
    """The update is linear: fields from sources A and B together equal the sum
    of the separate runs (fp64, to rounding)."""
    n = 30
    dt = stable_dt(DX, DX)
    gen = torch.Generator().manual_seed(1)
    ca = (0.9 + 0.1 * torch.rand(n, n, 1, generator=gen)).double()
    cb = (dt / EPS0 / (1 + 30 * torch.rand(n, n, 1, generator=gen))).double()

    def src(i, j, s):
        w = wr.gaussian(wr.time_axis(100, dt, torch.float64), 1.0, 5 * s * dt, s * dt)
        return OracleSource("Hz", w, torch.tensor([i]), torch.tensor([j]), torch.tensor([1.0], dtype=torch.float64))

    runs = []
    for srcs in ([src(8, 9, 4)], [src(20, 15, 6)], [src(8, 9, 4), src(20, 15, 6)]):
        o = OracleFDTD2D(n, n, DX, DX, dt, Ca=ca, Cb=cb, sources=srcs, dtype=torch.float64, n_check=10 ** 9)
        o.run(80)
        runs.append(o.Hz.clone())
    assert torch.allclose(runs[0] + runs[1], runs[2], rtol=1e-12, atol=1e-12 * float(runs[2].abs().max()))
