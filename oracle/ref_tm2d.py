"""CPU oracle: NumPy restatement of the reference 2D TM FDTD solver.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP product
path is compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
does.

It restates ``src/neurosim/em/fdtd.py`` of the reference (sushaan-k/jaxphys,
package ``neurosim``) operation by operation, in the reference's layout
``F[i, j]`` with shape ``(nx, ny)`` (y contiguous):

* constants ``C0, MU0, EPS0``                 -- fdtd.py:38-41
* ``PlaneWave`` / ``Wall``                    -- fdtd.py:44-71
* ``_build_conductor_mask``                   -- fdtd.py:143-157
* ``_build_pml_sigma``                        -- fdtd.py:159-191
* ``simulate`` prologue (dt, n_steps)         -- fdtd.py:193-247
* ``add_sources`` (whole row ``y``)           -- fdtd.py:249-254
* reflecting / periodic edge copies           -- fdtd.py:256-296
* periodic and sliced H / E updates           -- fdtd.py:306-335
* damping, edge copies, sources, conductors   -- fdtd.py:337-353
* ys stacking and subsampling                 -- fdtd.py:355-386

Precision follows ``src/neurosim/__init__.py:20-24``: ``np.float64`` is the
reference default (x64 on); ``np.float32`` is the reference with x64 off,
where Python scalars are weakly typed and round to float32 where they meet a
float32 array.  The one place where this differs from the 3D solver:
``sigma_max = 0.8*(3+1)/(dx*jnp.sqrt(1.0))`` (fdtd.py:174) is a float32
array scalar on the float32 path, so the profile values are float32
products ``f32(sigma_max) * f32(((n-i)/n)**3)``.

Pinning (``tests/test_oracle_tm2d.py``): ``simulate_bruteforce`` (per-cell
loops written from the index ranges of fdtd.py:306-353) agrees bit for bit
in float64; closed forms (profile values, dt, the row source after one
step, the conductor mask); invariants of the lossless periodic scheme
(conserved discrete energy, the Yee dispersion relation of one Fourier
mode, time reversal, linearity); the reflecting copies' mirror identities.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# fdtd.py:38-41 (jnp.pi is the Python float math.pi)
C0 = 299792458.0
MU0 = 4.0e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)


@dataclass(frozen=True)
class Wave:
    """fdtd.py:44-56 (PlaneWave): Ez[:, y] += amplitude*sin(2 pi f t)."""

    frequency: float
    y: int
    amplitude: float = 1.0


@dataclass(frozen=True)
class Wall:
    """fdtd.py:59-71: conductor row y, open on [gap_start, gap_end)."""

    y: int
    gap_start: int | None = None
    gap_end: int | None = None


def default_dt(dx: float, dtype) -> float:
    """fdtd.py:211: ``0.99*dx/(C0*float(jnp.sqrt(2.0)))`` (float32 sqrt on
    the float32 path)."""
    s2 = float(np.sqrt(np.asarray(2.0, dtype=dtype)))
    return float(0.99 * dx / (C0 * s2))


def conductor_mask(nx: int, ny: int, walls) -> np.ndarray:
    """fdtd.py:143-157: rows in insertion order; a wall with both gap ends
    set is open for ``gap_start <= x < gap_end``."""
    mask = np.zeros((nx, ny), dtype=bool)
    for w in walls:
        row = np.ones(nx, dtype=bool)
        if w.gap_start is not None and w.gap_end is not None:
            g = np.arange(nx)
            row = (g < w.gap_start) | (g >= w.gap_end)
        mask[:, w.y] = row
    return mask


def sigma_1d(n: int, n_pml: int, dx: float, dtype) -> np.ndarray:
    """fdtd.py:174-189, one axis: sequential writes, the mirror write second."""
    s = np.zeros(n, dtype=dtype)
    if n_pml == 0:
        return s
    # 0.8*(3+1) / (dx * jnp.sqrt(1.0)): a dtype scalar
    smax = dtype(0.8 * (3 + 1)) / (dtype(dx) * np.sqrt(dtype(1.0)))
    smax = dtype(smax)
    for i in range(n_pml):
        val = dtype(smax * dtype(((n_pml - i) / n_pml) ** 3))
        s[i] = val
        s[n - 1 - i] = val
    return s


def damping(nx: int, ny: int, n_pml: int, dx: float, dt: float, dtype) -> np.ndarray:
    """fdtd.py:229, 339: ``exp(-(0.5*(sigma_x+sigma_y))*dt/EPS0)`` per cell."""
    sx = sigma_1d(nx, n_pml, dx, dtype)[:, None]
    sy = sigma_1d(ny, n_pml, dx, dtype)[None, :]
    avg = (dtype(0.5) * (sx + sy)).astype(dtype)
    return np.exp(((-avg) * dtype(dt)) / dtype(EPS0)).astype(dtype)


def source_values(waves, n_steps: int, t_start: float, dt: float, dtype) -> np.ndarray:
    """fdtd.py:245-254 with t_c = t_start + n*dt (fdtd.py:306):
    ``amplitude * sin(omega * t_c)`` per step and wave, in ``dtype``."""
    n = np.arange(n_steps, dtype=dtype)
    t_c = (dtype(t_start) + (n * dtype(dt)).astype(dtype)).astype(dtype)
    out = np.zeros((n_steps, len(waves)), dtype=dtype)
    for s, w in enumerate(waves):
        om = dtype(2.0 * math.pi * w.frequency)
        out[:, s] = (dtype(w.amplitude) * np.sin((om * t_c).astype(dtype))).astype(dtype)
    return out


def _edge_copy(f: np.ndarray, mode: str) -> None:
    """fdtd.py:256-296 on one array (in place, the reference's order: the
    two x edge rows first, then the two y edge columns)."""
    if mode == "reflecting":
        f[0, :] = f[1, :]
        f[-1, :] = f[-2, :]
        f[:, 0] = f[:, 1]
        f[:, -1] = f[:, -2]
    else:
        f[0, :] = f[-2, :]
        f[-1, :] = f[1, :]
        f[:, 0] = f[:, -2]
        f[:, -1] = f[:, 1]


def step(ez, hx, hy, n: int, *, boundary: str, ch, ce, damp, mask, src_vals, waves, dtype):
    """One step of fdtd.py:304-353; returns new (ez, hx, hy)."""
    d = dtype
    if boundary == "periodic":
        hx_n = (hx - d(ch) * (np.roll(ez, -1, axis=1) - ez)).astype(d)
        hy_n = (hy + d(ch) * (np.roll(ez, -1, axis=0) - ez)).astype(d)
        dhy = hy_n - np.roll(hy_n, 1, axis=0)
        dhx = hx_n - np.roll(hx_n, 1, axis=1)
        ez_n = (ez + d(ce) * (dhy - dhx)).astype(d)
    else:
        hx_n = hx.copy()
        hx_n[:, :-1] = hx[:, :-1] - d(ch) * (ez[:, 1:] - ez[:, :-1])
        hy_n = hy.copy()
        hy_n[:-1, :] = hy[:-1, :] + d(ch) * (ez[1:, :] - ez[:-1, :])
        dhy = hy_n[1:, :] - hy_n[:-1, :]
        dhx = hx_n[:, 1:] - hx_n[:, :-1]
        ez_n = ez.copy()
        ez_n[1:, 1:] = ez[1:, 1:] + d(ce) * (dhy[:, 1:] - dhx[1:, :])
    if boundary == "absorbing":
        ez_n = (ez_n * damp).astype(d)
    elif boundary in ("reflecting", "periodic"):
        for f in (ez_n, hx_n, hy_n):
            _edge_copy(f, boundary)
    for s, w in enumerate(waves):
        ez_n[:, w.y] = ez_n[:, w.y] + src_vals[n, s]
    if mask is not None:
        ez_n = np.where(mask, d(0.0), ez_n).astype(d)
    return ez_n, hx_n, hy_n


def simulate(size, dx: float, boundary: str, n_pml: int, waves, walls=(),
             t_span=(0.0, 1e-8), dt=None, save_every: int = 10, dtype=np.float64,
             n_steps=None, init=None):
    """fdtd.py:193-386.  Returns (t, {"ez","hx","hy": (n_snap, nx, ny)}).
    ``init`` (extension for tests): initial (ez, hx, hy) instead of zeros."""
    nx, ny = size
    if dt is None:
        dt = default_dt(dx, dtype)
    dt = float(dt)
    if n_steps is None:
        n_steps = int((t_span[1] - t_span[0]) / dt)
    t_start = t_span[0]
    ch = dt / (float(MU0) * dx)
    ce = dt / (float(EPS0) * dx)
    if init is None:
        ez = np.zeros((nx, ny), dtype)
        hx = np.zeros((nx, ny), dtype)
        hy = np.zeros((nx, ny), dtype)
    else:
        ez, hx, hy = (np.array(a, dtype=dtype) for a in init)
    mask = conductor_mask(nx, ny, walls) if walls else None
    damp = damping(nx, ny, n_pml, dx, dt, dtype) if boundary == "absorbing" else None
    vals = source_values(waves, n_steps, t_start, dt, dtype)
    saves = range(0, n_steps, save_every) if save_every > 1 else range(n_steps)
    keep = set(saves)
    out = {"ez": [], "hx": [], "hy": []}
    ts = []
    for n in range(n_steps):
        ez, hx, hy = step(ez, hx, hy, n, boundary=boundary, ch=ch, ce=ce, damp=damp, mask=mask,
                          src_vals=vals, waves=waves, dtype=dtype)
        if n in keep:
            out["ez"].append(ez.copy())
            out["hx"].append(hx.copy())
            out["hy"].append(hy.copy())
            t_c = dtype(dtype(t_start) + dtype(dtype(n) * dtype(dt)))
            ts.append(dtype(t_c + dtype(dt)))
    res = {k: (np.stack(v) if v else np.zeros((0, nx, ny), dtype)) for k, v in out.items()}
    return np.asarray(ts, dtype=dtype), res


def simulate_bruteforce(size, dx: float, boundary: str, n_pml: int, waves, walls=(),
                        dt=None, n_steps: int = 10, dtype=np.float64, init=None):
    """Per-cell loops written from the index ranges of fdtd.py:306-353
    (pins ``simulate``); returns the final (ez, hx, hy)."""
    nx, ny = size
    if dt is None:
        dt = default_dt(dx, dtype)
    ch = dtype(dt / (float(MU0) * dx))
    ce = dtype(dt / (float(EPS0) * dx))
    if init is None:
        ez = np.zeros((nx, ny), dtype)
        hx = np.zeros((nx, ny), dtype)
        hy = np.zeros((nx, ny), dtype)
    else:
        ez, hx, hy = (np.array(a, dtype=dtype) for a in init)
    mask = conductor_mask(nx, ny, walls) if walls else np.zeros((nx, ny), bool)
    damp = damping(nx, ny, n_pml, dx, dt, dtype)
    vals = source_values(waves, n_steps, 0.0, dt, dtype)
    per = boundary == "periodic"

    def src_i(i):  # where the edge copies take cell i's value from
        if boundary == "reflecting":
            return 1 if i == 0 else (nx - 2 if i == nx - 1 else i)
        if per:
            return nx - 2 if i == 0 else (1 if i == nx - 1 else i)
        return i

    def src_j(j):
        if boundary == "reflecting":
            return 1 if j == 0 else (ny - 2 if j == ny - 1 else j)
        if per:
            return ny - 2 if j == 0 else (1 if j == ny - 1 else j)
        return j

    for n in range(n_steps):
        hxn = np.empty_like(hx)
        hyn = np.empty_like(hy)
        ezn = np.empty_like(ez)
        for i in range(nx):
            for j in range(ny):
                if per or j <= ny - 2:
                    hxn[i, j] = hx[i, j] - ch * (ez[i, (j + 1) % ny] - ez[i, j])
                else:
                    hxn[i, j] = hx[i, j]
                if per or i <= nx - 2:
                    hyn[i, j] = hy[i, j] + ch * (ez[(i + 1) % nx, j] - ez[i, j])
                else:
                    hyn[i, j] = hy[i, j]
        for i in range(nx):
            for j in range(ny):
                if per or (i >= 1 and j >= 1):
                    ezn[i, j] = ez[i, j] + ce * ((hyn[i, j] - hyn[i - 1, j])
                                                 - (hxn[i, j] - hxn[i, j - 1]))
                else:
                    ezn[i, j] = ez[i, j]
                if boundary == "absorbing":
                    ezn[i, j] = ezn[i, j] * damp[i, j]
        fe, fx, fy = np.empty_like(ez), np.empty_like(hx), np.empty_like(hy)
        for i in range(nx):
            for j in range(ny):
                a, b = src_i(i), src_j(j)
                fe[i, j], fx[i, j], fy[i, j] = ezn[a, b], hxn[a, b], hyn[a, b]
                for s, w in enumerate(waves):
                    if j == w.y:
                        fe[i, j] = fe[i, j] + vals[n, s]
                if mask[i, j]:
                    fe[i, j] = dtype(0.0)
        ez, hx, hy = fe, fx, fy
    return ez, hx, hy
