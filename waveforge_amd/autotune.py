"""Launch-shape autotuner for the temporally blocked stepper.

The stepper's free parameters are the tile shape (core height ``tile_rows``
and columns per lane ``cpl``, one of the shapes the library reports through
``wf_tb_tile_shapes``) and the temporal depth (steps per launch, 1..8; the
tile width follows as 64*cpl - 2*depth). Their best values depend on the set of grids in flight (a
4096^2 grid and 512 small grids fill the 256 CUs differently) and on whether
the runs carry per-cell coefficients. ``tune_runs`` times every admissible
shape and depth on scratch copies of the caller's runs (same shapes, same Ca/Cb, fresh
field buffers, no sources or detectors), with HIP events around each launch
on the launch stream, and returns the fastest. Results are cached per
(device, grid shapes, materials, division mode, and the depths each shape
admits for the runs' sources and detectors) for the process, and can be
saved to and loaded from a JSON file.

Admissible means the caller's own runs meet the kernel's per-tile source and
detector limits at every depth up to the candidate's, since a run of ``n``
steps at depth ``d`` also launches shallower blocks.
"""

from __future__ import annotations

import json
from collections import Counter
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import torch

from . import _native
from .engine import RunState, TBPlan, tb_fits_any_tile, tb_limits_ok

DEPTHS = tuple(range(1, _native.TB_MAX_DEPTH + 1))


@dataclass
class TuneResult:
    tile_rows: int
    depth: int
    rates: dict = field(default_factory=dict)   # shape_key(rows, cpl, depth) -> Mcell-updates/s
    cpl: int = 1

    def as_dict(self) -> dict:
        return {"tile_rows": self.tile_rows, "depth": self.depth, "cpl": self.cpl, "rates": self.rates}


def shape_key(tile_rows: int, cpl: int, depth: int) -> str:
    """Key of a shape in ``TuneResult.rates``: ``"rows,depth"`` for one column
    per lane, ``"rowsc2,depth"`` for two."""
    return f"{tile_rows},{depth}" if cpl == 1 else f"{tile_rows}c{cpl},{depth}"


_CACHE: dict = {}


def _admissible_depth(runs: Sequence[RunState], ti: int, cpl: int, max_depth: int) -> int:
    """Largest depth d <= max_depth such that the runs meet the per-tile
    source and detector limits at every depth up to d (0 if none)."""
    hard = [r for r in runs if not tb_fits_any_tile(r)]  # the others fit every shape
    ok = 0
    for d in range(1, max_depth + 1):
        if not all(tb_limits_ok(r, ti, d, cpl) for r in hard):
            break
        ok = d
    return ok


def _key(runs: Sequence[RunState], device, fast_div: bool, admissible, depths) -> str:
    """Cache key: device, grid shapes, materials, division mode, cell size,
    the candidate depths and, per candidate shape, the largest admissible
    depth for these runs' sources and detectors (so a denser set of the same
    grids is tuned again rather than handed a shape it cannot run)."""
    grids = sorted(Counter((r.nx, r.ny) for r in runs).items())
    mats = int(runs[0].ca is not None)
    c = runs[0].coef
    name = torch.cuda.get_device_name(device)
    division = "reciprocal" if runs[0].reciprocal else ("fast" if fast_div else "ieee")
    return json.dumps([name, grids, mats, division, c.dx, c.dy, sorted(depths), admissible])


def _scratch(runs: Sequence[RunState], device) -> list[RunState]:
    g = torch.Generator(device=device).manual_seed(0)
    out = []
    for r in runs:
        f = [(torch.rand((r.nx, r.ny), generator=g, device=device) - 0.5) * 1e-2 for _ in range(6)]
        s = RunState(f[0], f[1], f[2], r.ca, r.cb, r.nx, r.ny, r.coef, reciprocal=r.reciprocal)
        s.second = (f[3], f[4], f[5])
        out.append(s)
    return out


def tune_runs(runs: Sequence[RunState], device=None, *, fast_div: bool = True,
              tile_rows: Optional[Sequence] = None, depths: Sequence[int] = DEPTHS,
              launches: int = 6, use_cache: bool = True) -> TuneResult:
    """Fastest ``(tile_rows, cpl, depth)`` for advancing ``runs`` together.
    ``tile_rows`` restricts the shapes (ints: one column per lane, or
    ``(rows, cpl)`` pairs)."""
    if not runs:
        raise ValueError("no runs to tune")
    device = torch.device(device) if device is not None else runs[0].ex.device
    shapes = ([(t, 1) if isinstance(t, int) else (int(t[0]), int(t[1])) for t in tile_rows]
              if tile_rows is not None else _native.tile_shapes())
    admissible = [(ti, cpl, _admissible_depth(runs, ti, cpl, max(depths))) for ti, cpl in shapes]
    key = _key(runs, device, fast_div, admissible, depths)
    if use_cache and key in _CACHE:
        return _CACHE[key]
    scratch = _scratch(runs, device)
    stream = _native.stream_handle(device)
    launches += launches % 2
    best, rates = None, {}
    for ti, cpl, ok_depth in admissible:
        if ok_depth == 0:
            continue
        plan = TBPlan(scratch, ti, device, fast_div=fast_div, cpl=cpl)
        for d in depths:
            if d > ok_depth:
                continue
            plan.tiles(d)
            for k in range(2):  # warm-up: code object, tile table, caches
                plan.launch(d, k % 2, 0, stream)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record()
            for k in range(launches):
                plan.launch(d, k % 2, 0, stream)
            ev1.record()
            ev1.synchronize()
            sec = ev0.elapsed_time(ev1) / 1e3
            rate = plan.n_cells * d * launches / sec / 1e6
            rates[shape_key(ti, cpl, d)] = round(rate, 1)
            if best is None or rate > best[0]:
                best = (rate, ti, d, cpl)
    if best is None:
        raise ValueError("no admissible tile shape: too many sources or detectors per tile")
    res = TuneResult(best[1], best[2], rates, best[3])
    _CACHE[key] = res
    return res


def save(path) -> None:
    """Write the process cache to a JSON file."""
    Path(path).write_text(json.dumps({k: v.as_dict() for k, v in _CACHE.items()}, indent=1))


def load(path) -> int:
    """Merge tuning results from a JSON file into the process cache; returns
    the number of entries read."""
    data = json.loads(Path(path).read_text())
    for k, v in data.items():
        _CACHE[k] = TuneResult(int(v["tile_rows"]), int(v["depth"]), dict(v.get("rates", {})), int(v.get("cpl", 1)))
    return len(data)
