"""Hz point detectors recorded inside the stepping kernels.

The reference records receivers in its callers' step loops:
``sig_buf[:, step] = fields.Hz[rx_i, rx_j, 0]`` after each ``sim.step()``
(``examples/brain_mimo_imaging.py:196-198``, ``examples/02_dielectric_slab.py:52-55``,
``src/datasets/brain/antenna.py:173-198``). ``DetectorSet`` holds the same
cells and a device trace buffer that the HIP kernels fill as they step:
``traces[..., d, n - start_step]`` is Hz at ``cells[d]`` after step ``n``
(0-based), i.e. the value the reference loop stores in column ``n``.
"""

from __future__ import annotations

import torch


class DetectorSet:
    """Hz detectors at integer cells ``(i, j)`` recording ``n_steps`` samples."""

    def __init__(self, cells, n_steps: int, *, grid, batch_size: int = 1, start_step: int = 0) -> None:
        cells = [(int(i), int(j)) for i, j in cells]
        for i, j in cells:
            if not (0 <= i < grid.Nx and 0 <= j < grid.Ny):
                raise ValueError(f"detector cell ({i}, {j}) outside grid ({grid.Nx}, {grid.Ny})")
        if n_steps < 1:
            raise ValueError(f"n_steps must be >= 1, got {n_steps!r}")
        self.cells = cells
        self.n_steps = int(n_steps)
        self.start_step = int(start_step)
        self.batch_size = int(batch_size)
        shape = (len(cells), self.n_steps) if batch_size == 1 else (batch_size, len(cells), self.n_steps)
        self.traces = torch.zeros(shape, dtype=torch.float32, device=grid.device)
        flat = [i * grid.Ny + j for i, j in cells]
        self.cell_index = torch.tensor(flat, dtype=torch.int32, device=grid.device)

    def __len__(self) -> int:
        return len(self.cells)

    def numpy(self):
        return self.traces.detach().cpu().numpy()
