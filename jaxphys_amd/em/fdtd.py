"""2D TM FDTD solver (Ez, Hx, Hy) with the reference's construction/run API.

Drop-in for ``neurosim.em.fdtd`` (src/neurosim/em/fdtd.py:44-386):
``EMGrid(size, resolution, boundary, pml_layers)``, ``add_source(PlaneWave)``,
``add_conductor(Wall)`` and ``simulate(t_span, dt, save_every)`` take the
same arguments, validate them the same way (the pydantic ``EMConfig`` of
config.py:96-110 and the same ``ConfigurationError`` messages) and return
the same ``EMFieldHistory`` (torch tensors).  The time loop runs in
libtm2d_hip.so on the GPU in float32 (the reference's float32 path, with
x64 switched off): one fused kernel per step, see csrc/tm2d.hip.

Extension (keyword-only, default = reference behaviour):
``simulate(..., n_steps=None)`` -- an explicit step count.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np
import torch

from jaxphys_amd.config import EMConfig
from jaxphys_amd.em import sources as _src
from jaxphys_amd.em.constants import C0, EPS0, MU0
from jaxphys_amd.exceptions import ConfigurationError
from jaxphys_amd.state import EMFieldHistory

logger = logging.getLogger(__name__)


@dataclass(frozen=True)
class PlaneWave:
    """Plane wave source (fdtd.py:44-56): every step adds
    ``amplitude * sin(2 pi frequency t)`` to the whole row ``Ez[:, y]``.

    Attributes:
        frequency: Wave frequency in Hz.
        y: Source y-position (grid index).
        amplitude: Peak electric field amplitude (V/m).
    """

    frequency: float
    y: int
    amplitude: float = 1.0


@dataclass(frozen=True)
class Wall:
    """Conducting wall with optional gap (fdtd.py:59-71): Ez is forced to 0
    on row ``y`` except for ``gap_start <= x < gap_end`` (when both are set).
    """

    y: int
    gap_start: int | None = None
    gap_end: int | None = None


def sigma_profile_f32(n: int, n_pml: int, dx: float) -> np.ndarray:
    """fdtd.py:159-189, one axis, on the float32 path: ``sigma_max =
    0.8*(3+1)/(dx*jnp.sqrt(1.0))`` is a float32 scalar, each value the
    float32 product ``sigma_max*((n_pml-i)/n_pml)**3``, written in
    increasing ``i`` with the mirror write second (later writes win)."""
    f32 = np.float32
    s = np.zeros(n, dtype=f32)
    if n_pml == 0:
        return s
    smax = f32(f32(0.8 * (3 + 1)) / (f32(dx) * f32(1.0)))
    for i in range(n_pml):
        val = f32(smax * f32(((n_pml - i) / n_pml) ** 3))
        s[i] = val
        s[n - 1 - i] = val
    return s


class EMGrid:
    """2D electromagnetic FDTD simulation grid (fdtd.py:74-386) on a HIP device.

    Example:
        >>> grid = EMGrid(size=(200, 200), resolution=0.01)
        >>> grid.add_source(PlaneWave(frequency=3e9, y=20))
        >>> grid.add_conductor(Wall(y=100, gap_start=90, gap_end=110))
        >>> fields = grid.simulate(t_span=(0, 1e-8), dt=1e-11)

    Args:
        size: Grid dimensions (nx, ny) in cells.
        resolution: Cell size in meters.
        boundary: "absorbing", "periodic" or "reflecting".
        pml_layers: Number of damped layers (absorbing boundary only).
    """

    def __init__(
        self,
        size: tuple[int, int] = (200, 200),
        resolution: float = 0.01,
        boundary: str = "absorbing",
        pml_layers: int = 10,
        *,
        device=None,
        dtype=torch.float32,
    ) -> None:
        self._nx, self._ny = size
        self._dx = resolution
        # fdtd.py:96-101: the pydantic model validates the arguments
        self._config = EMConfig(
            resolution=resolution,
            courant_number=0.5,
            boundary=boundary,
            pml_layers=pml_layers,
        )
        self._sources: list[PlaneWave] = []
        self._conductors: list[Wall] = []
        self._device = device

        # fdtd.py:105-107
        if self._nx < 10 or self._ny < 10:
            raise ConfigurationError(f"Grid must be at least 10x10, got {size}")
        if dtype not in (torch.float32, np.float32, "float32"):
            raise ConfigurationError(
                f"the HIP FDTD step computes in float32; dtype={dtype!r} is not supported")

    @property
    def size(self) -> tuple[int, int]:
        """Grid dimensions (nx, ny)."""
        return (self._nx, self._ny)

    def add_source(self, source: PlaneWave) -> None:
        """Add a plane wave source to the grid (fdtd.py:120-130)."""
        if source.y < 0 or source.y >= self._ny:
            raise ConfigurationError(
                f"Source y={source.y} out of grid bounds [0, {self._ny})"
            )
        self._sources.append(source)

    def add_conductor(self, wall: Wall) -> None:
        """Add a conducting wall (with optional slit) to the grid (fdtd.py:132-141)."""
        if wall.y < 0 or wall.y >= self._ny:
            raise ConfigurationError(
                f"Wall y={wall.y} out of grid bounds [0, {self._ny})"
            )
        self._conductors.append(wall)

    # ------------------------------------------------------------------
    def _conductor_mask(self) -> np.ndarray | None:
        """fdtd.py:143-157 as uint8 in the x-fastest layout (ny, nx); None
        without walls."""
        if not self._conductors:
            return None
        nx, ny = self._nx, self._ny
        mask = np.zeros((ny, nx), dtype=np.uint8)
        for wall in self._conductors:
            row = np.ones(nx, dtype=bool)
            if wall.gap_start is not None and wall.gap_end is not None:
                gap = np.arange(nx)
                row = (gap < wall.gap_start) | (gap >= wall.gap_end)
            mask[wall.y, :] = row
        return mask

    def _damping(self, dt: float) -> np.ndarray:
        """fdtd.py:229, 339: ``exp(-(0.5*(sigma_x+sigma_y))*dt/EPS0)`` per
        cell in float32, x-fastest (ny, nx).  A one-off table: the kernel
        reads it only in the band where it differs from 1."""
        f32 = np.float32
        p = self._config.pml_layers
        sx = sigma_profile_f32(self._nx, p, self._dx)[None, :]
        sy = sigma_profile_f32(self._ny, p, self._dx)[:, None]
        avg = (f32(0.5) * (sx + sy)).astype(f32)
        return np.exp(((-avg) * f32(dt)) / f32(EPS0)).astype(f32)

    def build_solver(self, dt: float, n_steps: int, t_start: float = 0.0, device=None):
        """Device state for this grid (extension; used by ``simulate`` and
        the benchmarks)."""
        from jaxphys_amd import _native

        dev = device or self._device or _native.require_device()
        return TM2DSolver(self, dt, n_steps, t_start, dev)

    def simulate(
        self,
        t_span: tuple[float, float] = (0.0, 1e-8),
        dt: float | None = None,
        save_every: int = 10,
        *,
        n_steps: int | None = None,
    ) -> EMFieldHistory:
        """Run the FDTD simulation (fdtd.py:193-386).

        Args:
            t_span: (t_start, t_end) in seconds.
            dt: Time step. If None, computed from the CFL condition.
            save_every: Save field snapshot every N steps.
            n_steps: extension -- explicit step count (overrides t_span).

        Returns:
            EMFieldHistory with time-series of field snapshots.
        """
        dx = self._dx
        nx, ny = self._nx, self._ny
        # fdtd.py:211; jnp.sqrt(2.0) is float32 on the float32 path
        dt = float(0.99 * dx / (C0 * float(np.sqrt(np.float32(2.0))))) if dt is None else float(dt)

        t_start, t_end = t_span
        if n_steps is None:
            n_steps = int((t_end - t_start) / dt)
        n_steps = int(n_steps)

        if not self._sources:
            raise ConfigurationError(
                "At least one source must be added before simulation"
            )

        logger.info(
            "Starting FDTD simulation: grid=%dx%d, dt=%.2e, n_steps=%d",
            nx, ny, float(dt), n_steps,
        )

        solver = self.build_solver(dt, n_steps, t_start)
        dev = solver.device
        # fdtd.py:362-368: ys[0::save_every] (every step when save_every <= 1)
        saves = list(range(0, n_steps, save_every)) if save_every > 1 else list(range(n_steps))
        hist = [torch.empty((len(saves), ny, nx), dtype=torch.float32, device=dev)
                for _ in range(3)]
        cur = 0
        for q, n in enumerate(saves):
            solver.run(cur, n + 1 - cur)
            cur = n + 1
            for c in range(3):
                hist[c][q].copy_(solver.fields[c])
        solver.run(cur, n_steps - cur)
        self.last_solver = solver

        # fdtd.py:306, 355: t = t_c + dt with t_c = t_start + n*dt (float32)
        f32 = np.float32
        nn = np.asarray(saves, dtype=f32)
        t_c = (f32(t_start) + (nn * f32(dt)).astype(f32)).astype(f32)
        t = torch.from_numpy((t_c + f32(dt)).astype(f32)).to(dev)
        view = [h.permute(0, 2, 1) for h in hist]
        # fdtd.py:234-235
        gx = torch.arange(nx, dtype=torch.float32, device=dev) * dx
        gy = torch.arange(ny, dtype=torch.float32, device=dev) * dx
        return EMFieldHistory(t=t, ez=view[0], hx=view[1], hy=view[2], grid_x=gx, grid_y=gy)


class TM2DSolver:
    """GPU state and stepping of one 2D TM grid (internal layout (ny, nx),
    two buffer sets for the ping-pong step)."""

    def __init__(self, grid: EMGrid, dt: float, n_steps: int, t_start: float, device):
        from jaxphys_amd import _tm2d

        _tm2d.lib()
        self.device = torch.device(device)
        self.nx, self.ny = grid.size
        self.dt = float(dt)
        cfg = grid._config
        nx, ny = self.nx, self.ny
        dev = self.device
        d = _tm2d.Tm2dDesc()
        d.nx, d.ny = nx, ny
        d.boundary = {"absorbing": _tm2d.ABSORBING, "periodic": _tm2d.PERIODIC,
                      "reflecting": _tm2d.REFLECTING}[cfg.boundary]
        d.npml = cfg.pml_layers if cfg.boundary == "absorbing" else 0
        # fdtd.py:240-241 (Python float64, rounded to float32 at use)
        d.ch = float(np.float32(self.dt / (float(MU0) * grid._dx)))
        d.ce = float(np.float32(self.dt / (float(EPS0) * grid._dx)))
        self.damp = None
        if d.npml > 0:
            self.damp = torch.from_numpy(grid._damping(self.dt)).to(dev)
        d.damp = _ptr(self.damp)
        mask = grid._conductor_mask()
        self.conductor = None if mask is None else torch.from_numpy(mask).to(dev)
        d.conductor = _ptr(self.conductor)
        n_tab = max(1, int(n_steps))
        cols = [_src.reference_sine(s.frequency, s.amplitude, n_tab, t_start, self.dt)
                for s in grid._sources]
        table = (np.stack(cols, axis=1) if cols else np.zeros((n_tab, 0), np.float32))
        self.src_table = torch.from_numpy(np.ascontiguousarray(table, np.float32)).to(dev)
        self.src_rows = torch.tensor([int(s.y) for s in grid._sources], dtype=torch.int32,
                                     device=dev)
        d.n_src = len(cols)
        d.table_steps = n_tab
        d.src_rows = _ptr(self.src_rows) if cols else None
        d.src_table = _ptr(self.src_table) if cols else None
        self.desc = d
        shape = (ny, nx)
        self._bufs = [[torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(3)]
                      for _ in range(2)]
        self._cur = 0

    @property
    def fields(self):
        """The current (Ez, Hx, Hy) tensors, internal layout (ny, nx)."""
        return self._bufs[self._cur]

    def run(self, n0: int, nsteps: int, stream=None) -> None:
        """Steps n0 .. n0+nsteps-1."""
        from jaxphys_amd import _tm2d

        if nsteps <= 0:
            return
        a, b = self._cur, 1 - self._cur
        _tm2d.run(self.desc, self._bufs[a], self._bufs[b], n0, nsteps, stream)
        self._cur = a if nsteps % 2 == 0 else b

    def set_state(self, ez, hx, hy) -> None:
        """Load a state given in the reference's (nx, ny) index order."""
        for f, v in zip(self.fields, (ez, hx, hy)):
            f.copy_(torch.as_tensor(v, dtype=torch.float32).T.contiguous())

    def bytes_per_step(self) -> int:
        """Algorithmic HBM bytes of one step: Ez, Hx, Hy read and written
        (24 B/cell), the damping factor of each cell in the absorbing band
        and one conductor byte per cell when walls exist."""
        nx, ny = self.nx, self.ny
        b = 24 * nx * ny
        p = self.desc.npml
        if p > 0:
            band = nx * ny - max(0, nx - 2 * p) * max(0, ny - 2 * p)
            b += 4 * band
        if self.conductor is not None:
            b += nx * ny
        return b


def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


__all__ = ["EMGrid", "PlaneWave", "Wall", "TM2DSolver", "sigma_profile_f32"]
