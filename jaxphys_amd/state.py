"""Result containers of the FDTD solvers (src/neurosim/state.py:271-290,
387-413).

Same field names and shapes as the reference's ``EMFieldHistory`` (2D TM:
``ez, hx, hy`` of shape ``(n_snapshots, nx, ny)``) and ``EMFieldHistory3D``:
``t`` has shape ``(n_snapshots,)``, each field ``(n_snapshots, nx, ny, nz)``,
``grid_x/y/z`` the cell coordinates.  Values are torch tensors.  The field
tensors are views of x-fastest device storage, permuted to the reference's
``[i, j, k]`` index order (use ``.contiguous()`` for a z-contiguous copy).
"""

from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass(frozen=True)
class EMFieldHistory:
    """2D TM history (state.py:271-290): ``t`` (n_snapshots,), ``ez, hx, hy``
    (n_snapshots, nx, ny), ``grid_x`` (nx,), ``grid_y`` (ny,)."""

    t: torch.Tensor
    ez: torch.Tensor
    hx: torch.Tensor
    hy: torch.Tensor
    grid_x: torch.Tensor
    grid_y: torch.Tensor


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
