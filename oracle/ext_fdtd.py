"""CPU oracle for the extensions of the reference FDTD step.

TEST INFRASTRUCTURE ONLY (the checker; the product never imports it).

The reference (``src/neurosim/em/fdtd3d.py``) has a sponge boundary, no
material conductivity and only ``A*sin(omega*t)`` sources.  The HIP product
adds, behind the same ``EMGrid3D`` API:

* a convolutional PML (``boundary="cpml"``): Roden & Gedney, "Convolution
  PML (CPML)", Microw. Opt. Technol. Lett. 27 (2000); Taflove & Hagness,
  "Computational Electrodynamics" 3rd ed., ch. 7.  Polynomial grading of
  order ``m`` for sigma and kappa, linear alpha, ``sigma_max =
  -(m+1) ln(R0) / (2 eta0 L)`` with ``L = npml*dx``.  Recursive-convolution
  coefficients ``b = exp(-(sigma/kappa + alpha) dt/eps0)``, ``c = sigma
  (b-1) / (sigma kappa + kappa^2 alpha)``.  psi variables are kept in
  "difference" units (multiplied by dx) so that the update reads
  ``E = Ca E + Cb ((dH1 / kappa1 + psi1) - (dH2 / kappa2 + psi2))``.
* lossy media: ``Ca = (1-x)/(1+x)``, ``Cb = (dt/(eps0 eps_r dx))/(1+x)`` with
  ``x = sigma dt / (2 eps0 eps_r)`` (Taflove & Hagness eq. 3.31); where
  sigma == 0 the reference's ``Cb = ce/eps_r`` (float division in the working
  type) and ``Ca = 1`` are used, so lossless cells are bitwise the reference.
* pulse waveforms given as per-step value tables (``waveform_table``).

This module restates that scheme in NumPy in the reference layout
``(nx, ny, nz)``; the index ranges of the curl updates are exactly those of
fdtd3d.py:303-359 (the same planes stay un-updated and act as walls).  It is
pinned in ``tests/`` by: equality with ``ref_fdtd`` in the reference modes;
an independent torch-autograd float64 restatement (``autograd_fdtd``); the
closed-form coefficient values; CPML reflection well below the sponge's
against a large-domain run; energy decay in lossy media at the closed-form
rate.
"""

from __future__ import annotations

import math

import numpy as np

from oracle.ref_fdtd import C0, EPS0, MU0

ETA0 = math.sqrt(MU0 / EPS0)


# --------------------------------------------------------------------------
# coefficients
# --------------------------------------------------------------------------

def cpml_depth(n: int, npml: int, half: bool) -> np.ndarray:
    """Normalised depth d in [0, 1] of each node along one axis.

    Nodes live at position ``i`` (E-type derivative, ``half=False``) or
    ``i + 1/2`` (H-type derivative, ``half=True``).  Only indices in the two
    storage slabs ``[0, npml)`` and ``[n-npml, n)`` get a nonzero depth; the
    low interface sits at position ``npml`` and the high one at
    ``n - 1 - npml``.
    """
    d = np.zeros(n, dtype=np.float64)
    if npml <= 0:
        return d
    for i in range(n):
        x = i + (0.5 if half else 0.0)
        if i < npml:
            d[i] = (npml - x) / npml
        elif i >= n - npml:
            d[i] = (x - (n - 1 - npml)) / npml
    return np.clip(d, 0.0, 1.0)


def cpml_coeffs(n, npml, dx, dt, m=3.0, R0=1e-6, kappa_max=1.0, alpha_max=0.0):
    """Per-axis CPML tables ``{'e': (b, c, kinv), 'h': (b, c, kinv)}``, float64."""
    out = {}
    if npml > 0:
        sigma_max = -(m + 1.0) * math.log(R0) / (2.0 * ETA0 * npml * dx)
    else:
        sigma_max = 0.0
    for key, half in (("e", False), ("h", True)):
        d = cpml_depth(n, npml, half)
        in_slab = np.zeros(n, dtype=bool)
        if npml > 0:
            in_slab[:npml] = True
            in_slab[n - npml:] = True
        sig = sigma_max * d**m
        kap = 1.0 + (kappa_max - 1.0) * d**m
        alp = np.where(in_slab, alpha_max * (1.0 - d), 0.0)
        b = np.exp(-(sig / kap + alp) * dt / EPS0)
        den = sig * kap + kap * kap * alp
        c = np.where(sig > 0, sig * (b - 1.0) / np.where(den > 0, den, 1.0), 0.0)
        out[key] = (b, c, 1.0 / kap)
    return out


def e_coeffs(eps_r, sigma, dt, dx, dtype):
    """(Ca, Cb) per cell.  Lossless cells: Ca=1, Cb=dtype(ce)/eps_r."""
    dtype = np.dtype(dtype).type
    ce = dt / (float(EPS0) * dx)
    eps_t = np.asarray(eps_r, dtype=dtype)
    cb = (dtype(ce) / eps_t).astype(dtype)
    ca = np.ones_like(cb)
    if sigma is not None and np.any(np.asarray(sigma) != 0):
        sig = np.asarray(sigma, dtype=np.float64)
        e64 = np.asarray(eps_r, dtype=np.float64)
        x = sig * dt / (2.0 * EPS0 * e64)
        lossy = sig != 0
        ca = np.where(lossy, (1.0 - x) / (1.0 + x), ca).astype(dtype)
        cb = np.where(lossy, (ce / e64) / (1.0 + x), cb).astype(dtype)
    return ca, cb


# --------------------------------------------------------------------------
# waveforms (per-step source values; step n's value is added after E^{n+1})
# --------------------------------------------------------------------------

def waveform_table(kind, n_steps, *, amplitude=1.0, t0=0.0, spread=1.0,
                   wavelength=None, courant=None, frequency=None, dt=None):
    """float64 values v[n], n = 0..n_steps-1, time measured in steps."""
    n = np.arange(n_steps, dtype=np.float64)
    if kind == "gaussian_derivative":
        u = (n - t0) / spread
        v = -u * np.exp(-0.5 * u * u)
    elif kind == "gaussian":
        u = (n - t0) / spread
        v = np.exp(-0.5 * u * u)
    elif kind == "gaussian_sine":
        u = (n - t0) / spread
        v = np.sin(2.0 * math.pi * courant * n / wavelength) * np.exp(-0.5 * u * u)
    elif kind == "sine":
        v = np.sin(2.0 * math.pi * frequency * (n * dt))
    else:
        raise ValueError(kind)
    return amplitude * v


# --------------------------------------------------------------------------
# stepping
# --------------------------------------------------------------------------

class ExtState:
    """Fields and CPML psi arrays (full-size, zero outside the slabs)."""

    def __init__(self, size, dtype):
        self.f = [np.zeros(size, dtype=dtype) for _ in range(6)]
        self.psi = {}


PSI_NAMES = ["exy", "exz", "eyz", "eyx", "ezx", "ezy",
             "hxy", "hxz", "hyz", "hyx", "hzx", "hzy"]


def make_stepper(size, dx, dt, boundary, eps_r, sigma=None, npml=0,
                 cpml_kw=None, dtype=np.float64, sponge_layers=None):
    """Return ``(state, step_fn)``; ``step_fn(state, src_list)`` advances one
    step where ``src_list`` is ``[(comp, (i,j,k), value), ...]`` in order."""
    dtype = np.dtype(dtype).type
    nx, ny, nz = size
    ch = dtype(dt / (float(MU0) * dx))
    ca, cb = e_coeffs(eps_r, sigma, dt, dx, dtype)
    lossy = not np.all(ca == 1)
    st = ExtState(size, dtype)
    damp = None
    if boundary == "absorbing":
        from oracle.ref_fdtd import damping_factor, sigma_avg
        damp = damping_factor(sigma_avg(size, npml if sponge_layers is None
                                        else sponge_layers, dx, dtype), dt, dtype)
    cp = None
    if boundary == "cpml":
        kw = dict(cpml_kw or {})
        cp = {}
        for ax, n in zip("xyz", size):
            t = cpml_coeffs(n, npml, dx, dt, **kw)
            cp[ax] = {k: tuple(np.asarray(a, dtype=dtype) for a in v) for k, v in t.items()}
        for nm in PSI_NAMES:
            st.psi[nm] = np.zeros(size, dtype=dtype)

    def B(ax, kind, idx):
        b, c, kinv = cp[ax][kind]
        shape = [1, 1, 1]
        shape["xyz".index(ax)] = -1
        return (b[idx].reshape(shape), c[idx].reshape(shape), kinv[idx].reshape(shape))

    def step_fn(s, srcs):
        ex, ey, ez, hx, hy, hz = s.f
        if boundary == "periodic":
            from oracle.ref_fdtd import _curl_e_periodic, _curl_h_periodic
            cx, cy, cz = _curl_e_periodic(ex, ey, ez)
            hx = hx - ch * cx
            hy = hy - ch * cy
            hz = hz - ch * cz
            cx, cy, cz = _curl_h_periodic(hx, hy, hz)
            if lossy:
                ex = ca * ex + cb * cx
                ey = ca * ey + cb * cy
                ez = ca * ez + cb * cz
            else:
                ex = ex + cb * cx
                ey = ey + cb * cy
                ez = ez + cb * cz
        else:
            hx, hy, hz = hx.copy(), hy.copy(), hz.copy()
            dzy = ez[:, 1:, :-1] - ez[:, :-1, :-1]
            dyz = ey[:, :-1, 1:] - ey[:, :-1, :-1]
            dxz = ex[:-1, :, 1:] - ex[:-1, :, :-1]
            dzx = ez[1:, :, :-1] - ez[:-1, :, :-1]
            dyx = ey[1:, :-1, :] - ey[:-1, :-1, :]
            dxy = ex[:-1, 1:, :] - ex[:-1, :-1, :]
            if cp is None:
                hx[:, :-1, :-1] = hx[:, :-1, :-1] - ch * (dzy - dyz)
                hy[:-1, :, :-1] = hy[:-1, :, :-1] - ch * (dxz - dzx)
                hz[:-1, :-1, :] = hz[:-1, :-1, :] - ch * (dyx - dxy)
            else:
                P = s.psi
                by, cy_, ky = B("y", "h", slice(0, ny - 1))
                bz, cz_, kz = B("z", "h", slice(0, nz - 1))
                bx, cx_, kx = B("x", "h", slice(0, nx - 1))
                sl = (slice(None), slice(0, ny - 1), slice(0, nz - 1))
                P["hxy"][sl] = by * P["hxy"][sl] + cy_ * dzy
                P["hxz"][sl] = bz * P["hxz"][sl] + cz_ * dyz
                hx[sl] = hx[sl] - ch * ((dzy * ky + P["hxy"][sl]) - (dyz * kz + P["hxz"][sl]))
                sl = (slice(0, nx - 1), slice(None), slice(0, nz - 1))
                P["hyz"][sl] = bz * P["hyz"][sl] + cz_ * dxz
                P["hyx"][sl] = bx * P["hyx"][sl] + cx_ * dzx
                hy[sl] = hy[sl] - ch * ((dxz * kz + P["hyz"][sl]) - (dzx * kx + P["hyx"][sl]))
                sl = (slice(0, nx - 1), slice(0, ny - 1), slice(None))
                P["hzx"][sl] = bx * P["hzx"][sl] + cx_ * dyx
                P["hzy"][sl] = by * P["hzy"][sl] + cy_ * dxy
                hz[sl] = hz[sl] - ch * ((dyx * kx + P["hzx"][sl]) - (dxy * ky + P["hzy"][sl]))
            ex, ey, ez = ex.copy(), ey.copy(), ez.copy()
            a = (slice(None), slice(1, None), slice(1, None))
            b_ = (slice(1, None), slice(None), slice(1, None))
            c_ = (slice(1, None), slice(1, None), slice(None))
            d_hzy = hz[:, 1:, 1:] - hz[:, :-1, 1:]
            d_hyz = hy[:, 1:, 1:] - hy[:, 1:, :-1]
            d_hxz = hx[1:, :, 1:] - hx[1:, :, :-1]
            d_hzx = hz[1:, :, 1:] - hz[:-1, :, 1:]
            d_hyx = hy[1:, 1:, :] - hy[:-1, 1:, :]
            d_hxy = hx[1:, 1:, :] - hx[1:, :-1, :]
            if cp is None:
                cx, cy, cz = d_hzy - d_hyz, d_hxz - d_hzx, d_hyx - d_hxy
            else:
                P = s.psi
                by, cy_, ky = B("y", "e", slice(1, None))
                bz, cz_, kz = B("z", "e", slice(1, None))
                bx, cx_, kx = B("x", "e", slice(1, None))
                P["exy"][a] = by * P["exy"][a] + cy_ * d_hzy
                P["exz"][a] = bz * P["exz"][a] + cz_ * d_hyz
                cx = (d_hzy * ky + P["exy"][a]) - (d_hyz * kz + P["exz"][a])
                P["eyz"][b_] = bz * P["eyz"][b_] + cz_ * d_hxz
                P["eyx"][b_] = bx * P["eyx"][b_] + cx_ * d_hzx
                cy = (d_hxz * kz + P["eyz"][b_]) - (d_hzx * kx + P["eyx"][b_])
                P["ezx"][c_] = bx * P["ezx"][c_] + cx_ * d_hyx
                P["ezy"][c_] = by * P["ezy"][c_] + cy_ * d_hxy
                cz = (d_hyx * kx + P["ezx"][c_]) - (d_hxy * ky + P["ezy"][c_])
            if lossy:
                ex[a] = ca[a] * ex[a] + cb[a] * cx
                ey[b_] = ca[b_] * ey[b_] + cb[b_] * cy
                ez[c_] = ca[c_] * ez[c_] + cb[c_] * cz
            else:
                ex[a] = ex[a] + cb[a] * cx
                ey[b_] = ey[b_] + cb[b_] * cy
                ez[c_] = ez[c_] + cb[c_] * cz
        if damp is not None:
            ex, ey, ez = ex * damp, ey * damp, ez * damp
        for comp, pos, val in srcs:
            [ex, ey, ez][comp][pos] += dtype(val)
        s.f = [ex, ey, ez, hx, hy, hz]
        return s

    return st, step_fn


def run(size, dx, dt, boundary, eps_r, sources, n_steps, sigma=None, npml=0,
        cpml_kw=None, dtype=np.float64, observe=None):
    """Run ``n_steps``; ``sources`` is ``[(comp, (i,j,k), table[n_steps])]``.

    ``observe(n, state)`` is called after every step.  Returns the state."""
    st, fn = make_stepper(size, dx, dt, boundary, eps_r, sigma, npml, cpml_kw, dtype)
    for n in range(n_steps):
        fn(st, [(c, p, tab[n]) for c, p, tab in sources])
        if observe is not None:
            observe(n, st)
    return st
