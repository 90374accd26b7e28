"""Result container of the 3D FDTD solver (src/neurosim/state.py:387-413).

Same field names and shapes as the reference's ``EMFieldHistory3D``:
``t`` has shape ``(n_snapshots,)``, each field ``(n_snapshots, nx, ny, nz)``,
``grid_x/y/z`` the cell coordinates.  Values are torch tensors.  The field
tensors are views of x-fastest device storage, permuted to the reference's
``[i, j, k]`` index order (use ``.contiguous()`` for a z-contiguous copy).
"""

from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass(frozen=True)
class EMFieldHistory3D:
    t: torch.Tensor
    ex: torch.Tensor
    ey: torch.Tensor
    ez: torch.Tensor
    hx: torch.Tensor
    hy: torch.Tensor
    hz: torch.Tensor
    grid_x: torch.Tensor
    grid_y: torch.Tensor
    grid_z: torch.Tensor
