"""Point detectors recorded inside the stepping kernels.

The reference records receivers in its callers' step loops:
``sig_buf[:, step] = fields.Hz[rx_i, rx_j, 0]`` after each ``sim.step()``
(``examples/brain_mimo_imaging.py:196-198``, ``examples/02_dielectric_slab.py:52-55``,
``src/datasets/brain/antenna.py:173-198``). ``DetectorSet`` holds the same
cells and a device trace buffer that the HIP kernels fill as they step:
``traces[..., d, n - start_step]`` is Hz at ``cells[d]`` after step ``n``
(0-based), i.e. the value the reference loop stores in column ``n``.
On 3D grids cells are ``(i, j, k)`` and ``component`` may be any of the six
(the dataset generator records Ez, ``src/datasets/brain/antenna.py:173-198``);
the 2D TM stepper records Hz only.
"""

from __future__ import annotations

import numpy as np
import torch


COMPONENT_CODES = {"Ex": 0, "Ey": 1, "Hz": 2, "Ez": 3, "Hx": 4, "Hy": 5}


class DetectorSet:
    """Detectors of one field component at integer cells ``(i, j)`` or
    ``(i, j, k)``, recording ``n_steps`` samples."""

    def __init__(self, cells, n_steps: int, *, grid, batch_size: int = 1, start_step: int = 0,
                 component: str = "Hz") -> None:
        if component not in COMPONENT_CODES:
            raise ValueError(f"component must be one of {tuple(COMPONENT_CODES)}, got {component!r}")
        cells = [tuple(int(v) for v in c) for c in cells]
        Nz = getattr(grid, "Nz", 1)
        for c in cells:
            if len(c) == 2:
                c3 = (c[0], c[1], 0)
            elif len(c) == 3:
                c3 = c
            else:
                raise ValueError(f"detector cell {c} must be (i, j) or (i, j, k)")
            i, j, k = c3
            if not (0 <= i < grid.Nx and 0 <= j < grid.Ny and 0 <= k < Nz):
                raise ValueError(f"detector cell {c} outside grid ({grid.Nx}, {grid.Ny}, {Nz})")
        if n_steps < 1:
            raise ValueError(f"n_steps must be >= 1, got {n_steps!r}")
        self.cells = cells
        self.n_steps = int(n_steps)
        self.start_step = int(start_step)
        self.batch_size = int(batch_size)
        shape = (len(cells), self.n_steps) if batch_size == 1 else (batch_size, len(cells), self.n_steps)
        # recorded in the grid's dtype, as the reference's callers' buffers hold
        # its fields (float64 on a float64 grid)
        self.traces = torch.zeros(shape, dtype=getattr(grid, "dtype", torch.float32), device=grid.device)
        c3 = [c if len(c) == 3 else (c[0], c[1], 0) for c in cells]
        flat = [(i * grid.Ny + j) * Nz + k for i, j, k in c3]
        self.cell_index = torch.tensor(flat, dtype=torch.int32, device=grid.device)
        self._cell_index_host = np.array(flat, dtype=np.int64)  # the same cells, for host-side tables
        self.component = component
        self.component_index = torch.full((len(cells),), COMPONENT_CODES[component], dtype=torch.int32,
                                          device=grid.device)

    def __len__(self) -> int:
        return len(self.cells)

    def numpy(self):
        return self.traces.detach().cpu().numpy()
