"""Field inspection for the FDTD results: slices, Poynting flux, plots.

The reference's visualisation (src/neurosim/viz/fields.py:32-167) plots the
2D ``EMFieldHistory`` only; this module keeps its ``plot_field_snapshot``
and ``animate_field`` for the 2D grid (``animate_field`` also takes an
``EMFieldHistory3D`` and animates one plane of it) and adds the 3D counterparts: ``field_slice`` (a plane of
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
    import matplotlib.animation as animation
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


def animate_field(fields, component: str = "Ez", save: str | None = None, fps: int = 30,
                  figsize: tuple[float, float] = (8, 6), cmap: str = "RdBu_r", *,
                  axis: str = "z", index: int | None = None) -> Any:
    """Animate a field component over the saved snapshots
    (viz/fields.py:94-167): one frame per snapshot, symmetric colour scale
    fixed by the largest magnitude over all frames, title with the time.
    A 2D ``EMFieldHistory`` takes Ez, Hx or Hy as the reference does; an
    ``EMFieldHistory3D`` (extension) takes any of Ex..Hz and animates the
    plane ``index`` along ``axis`` (keyword-only; default the middle).
    ``save`` writes the animation with matplotlib's writers.  Returns the
    ``FuncAnimation``."""
    _check_matplotlib()
    if hasattr(fields, "grid_z"):
        if component not in _COMPONENTS_3D:
            raise VisualizationError(f"Unknown component '{component}'")
        if axis not in _AXES:
            raise VisualizationError(f"Unknown axis '{axis}'. Available: ['x', 'y', 'z']")
        full = getattr(fields, component.lower())
        a = _AXES[axis]
        n = full.shape[1 + a]
        index = n // 2 if index is None else int(index)
        if not 0 <= index < n:
            raise VisualizationError(f"index {index} outside [0, {n}) along {axis}")
        data = full.select(1 + a, index)
        rest = [ax for ax in "xyz" if ax != axis]
        g0 = getattr(fields, f"grid_{rest[0]}")
        g1 = getattr(fields, f"grid_{rest[1]}")
        labels = (f"{rest[0]} (m)", f"{rest[1]} (m)")
    else:
        component_map = {"Ez": fields.ez, "Hx": fields.hx, "Hy": fields.hy}
        if component not in component_map:
            raise VisualizationError(f"Unknown component '{component}'")
        data = component_map[component]
        g0, g1 = fields.grid_x, fields.grid_y
        labels = ("x (m)", "y (m)")
    data = data.detach().float().cpu()
    t = fields.t.detach().cpu()
    vmax = float(data.abs().max()) if data.numel() else 0.0
    if vmax == 0:
        vmax = 1.0
    g0, g1 = g0.cpu(), g1.cpu()
    fig, ax = plt.subplots(figsize=figsize)
    im = ax.imshow(data[0].T.numpy(), origin="lower", cmap=cmap, vmin=-vmax, vmax=vmax,
                   extent=(float(g0[0]), float(g0[-1]), float(g1[0]), float(g1[-1])))
    ax.set_xlabel(labels[0])
    ax.set_ylabel(labels[1])
    title = ax.set_title(f"{component} at t = 0")
    fig.colorbar(im, ax=ax)

    def update(frame: int) -> list:
        im.set_data(data[frame].T.numpy())
        title.set_text(f"{component} at t = {float(t[frame]):.2e} s")
        return [im, title]

    anim = animation.FuncAnimation(fig, update, frames=int(data.shape[0]),
                                   interval=1000 / fps, blit=False)
    if save:
        anim.save(save, fps=fps, dpi=100)
    return anim
