"""Field inspection for the FDTD results: slices, Poynting flux, plots.

The reference's visualisation (src/neurosim/viz/fields.py:32-167) plots the
2D ``EMFieldHistory`` only; this module keeps its ``plot_field_snapshot``
for the 2D grid and adds the 3D counterparts: ``field_slice`` (a plane of
an ``EMFieldHistory3D`` snapshot), ``plot_field_slice`` and
``poynting_flux`` (the power through a Yee plane of a running solver,
accumulated in float64 on the GPU by ``yee_poynting_flux_fp64``).
Plotting needs matplotlib; without it the plot functions raise
``VisualizationError`` as the reference does (viz/fields.py:25-29).
"""

from __future__ import annotations

from typing import Any

import torch

from jaxphys_amd.exceptions import VisualizationError

try:  # optional, as in the reference
    import matplotlib.pyplot as plt

    HAS_MATPLOTLIB = True
except ImportError:
    HAS_MATPLOTLIB = False

_AXES = {"x": 0, "y": 1, "z": 2}
_COMPONENTS_3D = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _check_matplotlib() -> None:
    if not HAS_MATPLOTLIB:
        raise VisualizationError(
            "matplotlib required for plotting; it is not installed in this environment")


def field_slice(fields, component: str = "Ez", axis: str = "z", index: int | None = None,
                step: int = -1) -> torch.Tensor:
    """One plane of one snapshot of an ``EMFieldHistory3D``: the component
    at ``index`` along ``axis`` (default: the middle plane), as a 2D tensor
    in the reference's index order (the two remaining axes in x, y, z
    order)."""
    if component not in _COMPONENTS_3D:
        raise VisualizationError(
            f"Unknown component '{component}'. Available: {list(_COMPONENTS_3D)}")
    if axis not in _AXES:
        raise VisualizationError(f"Unknown axis '{axis}'. Available: ['x', 'y', 'z']")
    data = getattr(fields, component.lower())[step]          # (nx, ny, nz)
    a = _AXES[axis]
    n = data.shape[a]
    index = n // 2 if index is None else int(index)
    if not 0 <= index < n:
        raise VisualizationError(f"index {index} outside [0, {n}) along {axis}")
    return data.select(a, index)


def poynting_flux(solver, axis: str = "z", plane: int | None = None) -> float:
    """Power (W in SI units) through the Yee plane normal to ``axis`` halfway
    between E planes ``plane`` and ``plane + 1`` (default: the middle), from
    the solver's current E and H: dx^2 * sum (E x H) . n, accumulated in
    float64 on the GPU."""
    from jaxphys_amd import _native

    if axis not in _AXES:
        raise VisualizationError(f"Unknown axis '{axis}'. Available: ['x', 'y', 'z']")
    a = _AXES[axis]
    shape = (solver.nx, solver.ny, solver.nz)
    plane = shape[a] // 2 - 1 if plane is None else int(plane)
    s = _native.poynting_flux_fp64(shape, solver.fields, a, plane)
    return float(s.item()) * solver.dx * solver.dx


def plot_field_slice(fields, component: str = "Ez", axis: str = "z",
                     index: int | None = None, step: int = -1,
                     figsize: tuple[float, float] = (8, 6), cmap: str = "RdBu_r") -> Any:
    """Plot a plane of a 3D snapshot (symmetric colour scale, physical axes)."""
    _check_matplotlib()
    data = field_slice(fields, component, axis, index, step).detach().float().cpu()
    rest = [ax for ax in "xyz" if ax != axis]
    g0 = getattr(fields, f"grid_{rest[0]}").cpu()
    g1 = getattr(fields, f"grid_{rest[1]}").cpu()
    vmax = float(data.abs().max()) or 1.0
    fig, ax = plt.subplots(figsize=figsize)
    im = ax.imshow(data.T.numpy(), origin="lower", cmap=cmap, vmin=-vmax, vmax=vmax,
                   extent=(float(g0[0]), float(g0[-1]), float(g1[0]), float(g1[-1])))
    ax.set_xlabel(f"{rest[0]} (m)")
    ax.set_ylabel(f"{rest[1]} (m)")
    ax.set_title(f"{component} at {axis} index {index}, t = {float(fields.t[step]):.2e} s")
    fig.colorbar(im, ax=ax, label=f"{component} (a.u.)")
    fig.tight_layout()
    return fig


def plot_field_snapshot(fields, step: int = -1, component: str = "Ez",
                        figsize: tuple[float, float] = (8, 6), cmap: str = "RdBu_r") -> Any:
    """Plot a snapshot of the 2D TM history (viz/fields.py:32-93)."""
    _check_matplotlib()
    component_map = {"Ez": fields.ez, "Hx": fields.hx, "Hy": fields.hy}
    if component not in component_map:
        raise VisualizationError(
            f"Unknown component '{component}'. Available: {list(component_map.keys())}")
    data = component_map[component][step].detach().float().cpu()
    vmax = float(data.abs().max()) or 1.0
    fig, ax = plt.subplots(figsize=figsize)
    gx, gy = fields.grid_x.cpu(), fields.grid_y.cpu()
    im = ax.imshow(data.T.numpy(), origin="lower", cmap=cmap, vmin=-vmax, vmax=vmax,
                   extent=(float(gx[0]), float(gx[-1]), float(gy[0]), float(gy[-1])))
    ax.set_xlabel("x (m)", fontsize=12)
    ax.set_ylabel("y (m)", fontsize=12)
    ax.set_title(f"{component} at t = {float(fields.t[step]):.2e} s", fontsize=13)
    fig.colorbar(im, ax=ax, label=f"{component} (a.u.)")
    fig.tight_layout()
    return fig
