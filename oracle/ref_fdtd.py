"""CPU oracle: NumPy restatement of the reference 3D FDTD step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP product
path is compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
does.

It restates ``src/neurosim/em/fdtd3d.py`` of the reference (sushaan-k/jaxphys,
package ``neurosim``) operation by operation, in the reference's own array
layout ``F[i, j, k]`` with shape ``(nx, ny, nz)`` (z contiguous):

* constants ``C0, MU0, EPS0``               -- fdtd3d.py:35-38
* ``_build_epsilon_r``                      -- fdtd3d.py:132-137
* ``_build_pml_sigma_3d``                   -- fdtd3d.py:139-167
* ``simulate`` prologue (dt, n_steps)       -- fdtd3d.py:185-231
* source injection ``add_sources``          -- fdtd3d.py:233-245
* periodic curls (``jnp.roll``)             -- fdtd3d.py:247-278, 290-301
* sliced H / E updates                      -- fdtd3d.py:303-359
* sponge damping (E only)                   -- fdtd3d.py:361-366
* step carry, ys stacking, subsampling      -- fdtd3d.py:371-417

Precision follows the reference's policy (``src/neurosim/__init__.py:20-24``):
``dtype=np.float64`` is the reference default (x64 on); ``dtype=np.float32``
is the reference with x64 switched off, where Python scalars are weakly typed
and round to float32 at the point of use.  Every scalar below is rounded with
``dtype(...)`` at exactly that point so the float32 operation order matches.

Pinning (how this restatement is itself checked, in ``tests/``):
* ``simulate_bruteforce`` -- per-cell Python loops written from the index
  ranges of fdtd3d.py:303-359 -- must agree bit-for-bit in float64;
* closed forms: sigma profile values, the source value at the source cell
  after steps 1 and 2, dt / n_steps arithmetic;
* invariants of the lossless periodic scheme: conserved discrete energy,
  preserved discrete divergence, Yee dispersion relation of a single
  Fourier mode, time reversal, linearity, mirror symmetry.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# fdtd3d.py:35-38 (jnp.pi is the Python float math.pi)
C0 = 299792458.0
MU0 = 4.0e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)


@dataclass(frozen=True)
class Src:
    """fdtd3d.py:41-55 (PointSource3D)."""

    frequency: float
    position: tuple
    amplitude: float = 1.0
    polarization: str = "z"


@dataclass(frozen=True)
class Mat:
    """fdtd3d.py:58-68 (DielectricRegion); epsilon_r may be an array."""

    mask: np.ndarray
    epsilon_r: object = 1.0


def default_dt(dx: float, dtype) -> float:
    """fdtd3d.py:190-192: ``0.99*dx/(C0*float(jnp.sqrt(3.0)))``.

    ``jnp.sqrt(3.0)`` is evaluated in the default float type, so with x64
    off the divisor uses float32(sqrt(3)).
    """
    s3 = float(np.sqrt(np.asarray(3.0, dtype=dtype)))
    return float(0.99 * dx / (C0 * s3))


def n_steps_for(t_span, dt: float) -> int:
    """fdtd3d.py:194-195 (float64 Python arithmetic, floor via int())."""
    t_start, t_end = t_span
    return int((t_end - t_start) / dt)


def build_eps_r(size, materials, dtype) -> np.ndarray:
    """fdtd3d.py:132-137: later materials overwrite earlier ones."""
    eps = np.ones(size, dtype=dtype)
    for m in materials:
        eps = np.where(np.asarray(m.mask, dtype=bool),
                       np.asarray(m.epsilon_r, dtype=dtype), eps).astype(dtype)
    return eps


def build_sigma_1d(n: int, n_pml: int, dx: float, dtype) -> np.ndarray:
    """fdtd3d.py:149-157: sequential writes, the mirror write second."""
    s = np.zeros(n, dtype=dtype)
    if n_pml == 0:
        return s
    sigma_max = 0.8 * 4.0 / (dx * 1.0)
    for i in range(n_pml):
        val = sigma_max * ((n_pml - i) / n_pml) ** 3
        s[i] = val
        s[n - 1 - i] = val
    return s


def sigma_avg(size, n_pml: int, dx: float, dtype) -> np.ndarray:
    """fdtd3d.py:139-167 and :221: ``(sigma_x + sigma_y + sigma_z) / 3.0``."""
    nx, ny, nz = size
    sx = build_sigma_1d(nx, n_pml, dx, dtype)[:, None, None]
    sy = build_sigma_1d(ny, n_pml, dx, dtype)[None, :, None]
    sz = build_sigma_1d(nz, n_pml, dx, dtype)[None, None, :]
    return (((sx + sy) + sz) / dtype(3.0)).astype(dtype)


def damping_factor(sig_avg: np.ndarray, dt: float, dtype) -> np.ndarray:
    """fdtd3d.py:363: ``jnp.exp(-sigma_avg * dt / float(EPS0))``."""
    return np.exp(((-sig_avg) * dtype(dt)) / dtype(EPS0)).astype(dtype)


def source_values(sources, n: int, t_start: float, dt: float, dtype):
    """fdtd3d.py:227, 236-238, 288 for step ``n``.

    ``t_c = t_start + step_idx*dt`` (step_idx an int32 carry, so the product
    is taken in the default float type), ``val = A*sin(omega*t_c)``.
    """
    t_c = dtype(t_start) + dtype(dtype(n) * dtype(dt))
    out = []
    for s in sources:
        omega = 2.0 * math.pi * s.frequency
        out.append(dtype(dtype(s.amplitude) * np.sin(dtype(dtype(omega) * t_c))))
    return out, t_c


def _curl_e_periodic(ex, ey, ez):
    """fdtd3d.py:249-264."""
    cx = (np.roll(ez, -1, axis=1) - ez) - (np.roll(ey, -1, axis=2) - ey)
    cy = (np.roll(ex, -1, axis=2) - ex) - (np.roll(ez, -1, axis=0) - ez)
    cz = (np.roll(ey, -1, axis=0) - ey) - (np.roll(ex, -1, axis=1) - ex)
    return cx, cy, cz


def _curl_h_periodic(hx, hy, hz):
    """fdtd3d.py:266-278."""
    cx = (hz - np.roll(hz, 1, axis=1)) - (hy - np.roll(hy, 1, axis=2))
    cy = (hx - np.roll(hx, 1, axis=2)) - (hz - np.roll(hz, 1, axis=0))
    cz = (hy - np.roll(hy, 1, axis=0)) - (hx - np.roll(hx, 1, axis=1))
    return cx, cy, cz


def step(fields, eps_r, damp, boundary, ch, ce, dtype):
    """One reference step without sources (fdtd3d.py:290-366).

    ``fields`` is ``[ex, ey, ez, hx, hy, hz]``; returns new arrays.
    ``damp`` is the per-cell damping factor or ``None`` (no damping).
    """
    ex, ey, ez, hx, hy, hz = fields
    chv = dtype(ch)
    cev = dtype(ce)
    if boundary == "periodic":
        cx, cy, cz = _curl_e_periodic(ex, ey, ez)
        hx = hx - chv * cx
        hy = hy - chv * cy
        hz = hz - chv * cz
        cx, cy, cz = _curl_h_periodic(hx, hy, hz)
        cb = cev / eps_r
        ex = ex + cb * cx
        ey = ey + cb * cy
        ez = ez + cb * cz
    else:
        hx = hx.copy()
        hy = hy.copy()
        hz = hz.copy()
        # fdtd3d.py:305-330
        hx[:, :-1, :-1] = hx[:, :-1, :-1] - chv * (
            (ez[:, 1:, :-1] - ez[:, :-1, :-1]) - (ey[:, :-1, 1:] - ey[:, :-1, :-1]))
        hy[:-1, :, :-1] = hy[:-1, :, :-1] - chv * (
            (ex[:-1, :, 1:] - ex[:-1, :, :-1]) - (ez[1:, :, :-1] - ez[:-1, :, :-1]))
        hz[:-1, :-1, :] = hz[:-1, :-1, :] - chv * (
            (ey[1:, :-1, :] - ey[:-1, :-1, :]) - (ex[:-1, 1:, :] - ex[:-1, :-1, :]))
        ex = ex.copy()
        ey = ey.copy()
        ez = ez.copy()
        # fdtd3d.py:334-359
        ex[:, 1:, 1:] = ex[:, 1:, 1:] + (cev / eps_r[:, 1:, 1:]) * (
            (hz[:, 1:, 1:] - hz[:, :-1, 1:]) - (hy[:, 1:, 1:] - hy[:, 1:, :-1]))
        ey[1:, :, 1:] = ey[1:, :, 1:] + (cev / eps_r[1:, :, 1:]) * (
            (hx[1:, :, 1:] - hx[1:, :, :-1]) - (hz[1:, :, 1:] - hz[:-1, :, 1:]))
        ez[1:, 1:, :] = ez[1:, 1:, :] + (cev / eps_r[1:, 1:, :]) * (
            (hy[1:, 1:, :] - hy[:-1, 1:, :]) - (hx[1:, 1:, :] - hx[1:, :-1, :]))
    if damp is not None:
        ex = ex * damp
        ey = ey * damp
        ez = ez * damp
    return [ex, ey, ez, hx, hy, hz]


def add_sources(fields, sources, vals):
    """fdtd3d.py:236-244 (insertion order; repeated cells accumulate)."""
    ex, ey, ez = fields[0], fields[1], fields[2]
    for s, v in zip(sources, vals):
        ix, iy, iz = s.position
        if s.polarization == "z":
            ez[ix, iy, iz] += v
        elif s.polarization == "x":
            ex[ix, iy, iz] += v
        else:
            ey[ix, iy, iz] += v
    return fields


def simulate(size, dx, boundary, pml_layers, sources, materials,
             t_span=(0.0, 5e-9), dt=None, save_every=20, dtype=np.float64,
             n_steps=None, keep=None):
    """fdtd3d.py:169-417 without stacking every step.

    Returns ``(t, snaps)`` where ``snaps`` maps a component name to an array
    ``(n_snap, nx, ny, nz)``; identical to the reference's subsampled ``ys``.
    ``keep`` optionally restricts which components are stored.
    """
    dtype = np.dtype(dtype).type
    if dt is None:
        dt = default_dt(dx, dtype)
    dt = float(dt)
    t_start = t_span[0]
    if n_steps is None:
        n_steps = n_steps_for(t_span, dt)
    ch = dt / (float(MU0) * dx)
    ce = dt / (float(EPS0) * dx)
    eps_r = build_eps_r(size, materials, dtype)
    damp = None
    if boundary == "absorbing":
        damp = damping_factor(sigma_avg(size, pml_layers, dx, dtype), dt, dtype)
    fields = [np.zeros(size, dtype=dtype) for _ in range(6)]
    if save_every > 1:
        save = set(range(0, n_steps, save_every))
    else:
        save = set(range(n_steps))
    names = ["ex", "ey", "ez", "hx", "hy", "hz"]
    keep = names if keep is None else list(keep)
    snaps = {k: [] for k in keep}
    ts = []
    for n in range(n_steps):
        fields = step(fields, eps_r, damp, boundary, ch, ce, dtype)
        vals, t_c = source_values(sources, n, t_start, dt, dtype)
        fields = add_sources(fields, sources, vals)
        if n in save:
            for k in keep:
                snaps[k].append(fields[names.index(k)].copy())
            ts.append(dtype(t_c + dtype(dt)))
    out = {k: (np.stack(v) if v else np.zeros((0,) + tuple(size), dtype))
           for k, v in snaps.items()}
    return np.asarray(ts, dtype=dtype), out


def simulate_bruteforce(size, dx, boundary, pml_layers, sources, materials,
                        n_steps, dt):
    """Per-cell float64 loops written from the index ranges of
    fdtd3d.py:303-366 and the roll directions of :249-278.  Tiny grids only.
    Returns the final ``[ex, ey, ez, hx, hy, hz]``."""
    nx, ny, nz = size
    ch = dt / (MU0 * dx)
    ce = dt / (EPS0 * dx)
    eps = build_eps_r(size, materials, np.float64)
    sx = build_sigma_1d(nx, pml_layers, dx, np.float64)
    sy = build_sigma_1d(ny, pml_layers, dx, np.float64)
    sz = build_sigma_1d(nz, pml_layers, dx, np.float64)
    E = [np.zeros(size) for _ in range(3)]
    H = [np.zeros(size) for _ in range(3)]
    per = boundary == "periodic"
    for n in range(n_steps):
        En = [a.copy() for a in E]
        Hn = [a.copy() for a in H]
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    ip, jp, kp = (i + 1) % nx, (j + 1) % ny, (k + 1) % nz
                    if per or (j < ny - 1 and k < nz - 1):
                        Hn[0][i, j, k] = H[0][i, j, k] - ch * (
                            (E[2][i, jp, k] - E[2][i, j, k]) - (E[1][i, j, kp] - E[1][i, j, k]))
                    if per or (i < nx - 1 and k < nz - 1):
                        Hn[1][i, j, k] = H[1][i, j, k] - ch * (
                            (E[0][i, j, kp] - E[0][i, j, k]) - (E[2][ip, j, k] - E[2][i, j, k]))
                    if per or (i < nx - 1 and j < ny - 1):
                        Hn[2][i, j, k] = H[2][i, j, k] - ch * (
                            (E[1][ip, j, k] - E[1][i, j, k]) - (E[0][i, jp, k] - E[0][i, j, k]))
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    im, jm, km = (i - 1) % nx, (j - 1) % ny, (k - 1) % nz
                    cb = ce / eps[i, j, k]
                    if per or (j >= 1 and k >= 1):
                        En[0][i, j, k] = E[0][i, j, k] + cb * (
                            (Hn[2][i, j, k] - Hn[2][i, jm, k]) - (Hn[1][i, j, k] - Hn[1][i, j, km]))
                    if per or (i >= 1 and k >= 1):
                        En[1][i, j, k] = E[1][i, j, k] + cb * (
                            (Hn[0][i, j, k] - Hn[0][i, j, km]) - (Hn[2][i, j, k] - Hn[2][im, j, k]))
                    if per or (i >= 1 and j >= 1):
                        En[2][i, j, k] = E[2][i, j, k] + cb * (
                            (Hn[1][i, j, k] - Hn[1][im, j, k]) - (Hn[0][i, j, k] - Hn[0][i, jm, k]))
                    if boundary == "absorbing":
                        d = np.exp(-((sx[i] + sy[j] + sz[k]) / 3.0) * dt / EPS0)
                        for c in range(3):
                            En[c][i, j, k] *= d
        t_c = n * dt
        for s in sources:
            c = 2 if s.polarization == "z" else (0 if s.polarization == "x" else 1)
            En[c][s.position] += s.amplitude * math.sin(2.0 * math.pi * s.frequency * t_c)
        E, H = En, Hn
    return E + H
