"""Torch-autograd float64 restatement of the forward step (CPU).

TEST INFRASTRUCTURE ONLY: the gradient oracle for the hand-written adjoint.
Reverse-mode autodiff through this restatement gives the exact discrete
gradient ``dJ/d eps_r`` that ``jax.grad`` of the reference's scan would give
(src/neurosim/em/fdtd3d.py:280-393; the reference's only gradient mechanism
is generic ``jax.grad``, src/neurosim/optimize.py:84).  Same scheme as
``oracle/ext_fdtd.py`` (which it is checked against), reference layout
``(nx, ny, nz)``; small grids only (it stores every step).
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import ext_fdtd as X
from oracle import ref_fdtd as R


def forward(size, dx, dt, boundary, eps_r: torch.Tensor, sources, n_steps, npml=0,
            sigma=None, monitor=None, cpml_kw=None):
    """Run ``n_steps`` and return ``(J, final_fields)``.

    ``eps_r``: float64 tensor (may require grad).  ``sources``:
    ``[(comp, (i,j,k), values[n_steps])]``.  ``monitor``: boolean array; the
    objective is ``J = sum_n sum_monitor Ez^{n+1}**2``.
    """
    dt64 = torch.float64
    nx, ny, nz = size
    ch = dt / (R.MU0 * dx)
    ce = dt / (R.EPS0 * dx)
    if sigma is not None and np.any(np.asarray(sigma) != 0):
        sig = torch.as_tensor(np.asarray(sigma, dtype=np.float64))
        x = sig * dt / (2.0 * R.EPS0 * eps_r)
        ca = (1.0 - x) / (1.0 + x)
        cb = (ce / eps_r) / (1.0 + x)
        lossy = True
    else:
        ca = None
        cb = ce / eps_r
        lossy = False
    damp = None
    if boundary == "absorbing":
        damp = torch.as_tensor(R.damping_factor(R.sigma_avg(size, npml, dx, np.float64), dt,
                                                np.float64))
    cp = None
    if boundary == "cpml":
        cp = {}
        for ax, n in zip("xyz", size):
            t = X.cpml_coeffs(n, npml, dx, dt, **(cpml_kw or {}))
            cp[ax] = {k: tuple(torch.as_tensor(np.asarray(a)) for a in v) for k, v in t.items()}
    mon = torch.as_tensor(np.asarray(monitor, dtype=bool)) if monitor is not None else None

    def B(ax, kind, idx):
        b, c, k = cp[ax][kind]
        shape = [1, 1, 1]
        shape["xyz".index(ax)] = -1
        return b[idx].reshape(shape), c[idx].reshape(shape), k[idx].reshape(shape)

    z = lambda: torch.zeros(size, dtype=dt64)  # noqa: E731
    ex, ey, ez, hx, hy, hz = (z() for _ in range(6))
    psi = {nm: z() for nm in X.PSI_NAMES} if cp is not None else None
    J = torch.zeros((), dtype=dt64)
    per = boundary == "periodic"
    for n in range(n_steps):
        if per:
            cx = (torch.roll(ez, -1, 1) - ez) - (torch.roll(ey, -1, 2) - ey)
            cy = (torch.roll(ex, -1, 2) - ex) - (torch.roll(ez, -1, 0) - ez)
            cz = (torch.roll(ey, -1, 0) - ey) - (torch.roll(ex, -1, 1) - ex)
            hx, hy, hz = hx - ch * cx, hy - ch * cy, hz - ch * cz
            cx = (hz - torch.roll(hz, 1, 1)) - (hy - torch.roll(hy, 1, 2))
            cy = (hx - torch.roll(hx, 1, 2)) - (hz - torch.roll(hz, 1, 0))
            cz = (hy - torch.roll(hy, 1, 0)) - (hx - torch.roll(hx, 1, 1))
            if lossy:
                ex, ey, ez = ca * ex + cb * cx, ca * ey + cb * cy, ca * ez + cb * cz
            else:
                ex, ey, ez = ex + cb * cx, ey + cb * cy, ez + cb * cz
        else:
            dzy = ez[:, 1:, :-1] - ez[:, :-1, :-1]
            dyz = ey[:, :-1, 1:] - ey[:, :-1, :-1]
            dxz = ex[:-1, :, 1:] - ex[:-1, :, :-1]
            dzx = ez[1:, :, :-1] - ez[:-1, :, :-1]
            dyx = ey[1:, :-1, :] - ey[:-1, :-1, :]
            dxy = ex[:-1, 1:, :] - ex[:-1, :-1, :]
            s1 = (slice(None), slice(0, ny - 1), slice(0, nz - 1))
            s2 = (slice(0, nx - 1), slice(None), slice(0, nz - 1))
            s3 = (slice(0, nx - 1), slice(0, ny - 1), slice(None))
            if cp is not None:
                by, cy_, ky = B("y", "h", slice(0, ny - 1))
                bz, cz_, kz = B("z", "h", slice(0, nz - 1))
                bx, cx_, kx = B("x", "h", slice(0, nx - 1))
                P = psi
                p1 = by * P["hxy"][s1] + cy_ * dzy
                p2 = bz * P["hxz"][s1] + cz_ * dyz
                u1 = (dzy * ky + p1) - (dyz * kz + p2)
                p3 = bz * P["hyz"][s2] + cz_ * dxz
                p4 = bx * P["hyx"][s2] + cx_ * dzx
                u2 = (dxz * kz + p3) - (dzx * kx + p4)
                p5 = bx * P["hzx"][s3] + cx_ * dyx
                p6 = by * P["hzy"][s3] + cy_ * dxy
                u3 = (dyx * kx + p5) - (dxy * ky + p6)
                for nm, sl, v in (("hxy", s1, p1), ("hxz", s1, p2), ("hyz", s2, p3),
                                  ("hyx", s2, p4), ("hzx", s3, p5), ("hzy", s3, p6)):
                    t = P[nm].clone()
                    t[sl] = v
                    P[nm] = t
            else:
                u1, u2, u3 = dzy - dyz, dxz - dzx, dyx - dxy
            hx = hx.clone(); hx[s1] = hx[s1] - ch * u1  # noqa: E702
            hy = hy.clone(); hy[s2] = hy[s2] - ch * u2  # noqa: E702
            hz = hz.clone(); hz[s3] = hz[s3] - ch * u3  # noqa: E702
            a = (slice(None), slice(1, None), slice(1, None))
            b_ = (slice(1, None), slice(None), slice(1, None))
            c_ = (slice(1, None), slice(1, None), slice(None))
            d_hzy = hz[:, 1:, 1:] - hz[:, :-1, 1:]
            d_hyz = hy[:, 1:, 1:] - hy[:, 1:, :-1]
            d_hxz = hx[1:, :, 1:] - hx[1:, :, :-1]
            d_hzx = hz[1:, :, 1:] - hz[:-1, :, 1:]
            d_hyx = hy[1:, 1:, :] - hy[:-1, 1:, :]
            d_hxy = hx[1:, 1:, :] - hx[1:, :-1, :]
            if cp is not None:
                by, cy_, ky = B("y", "e", slice(1, None))
                bz, cz_, kz = B("z", "e", slice(1, None))
                bx, cx_, kx = B("x", "e", slice(1, None))
                P = psi
                q1 = by * P["exy"][a] + cy_ * d_hzy
                q2 = bz * P["exz"][a] + cz_ * d_hyz
                w1 = (d_hzy * ky + q1) - (d_hyz * kz + q2)
                q3 = bz * P["eyz"][b_] + cz_ * d_hxz
                q4 = bx * P["eyx"][b_] + cx_ * d_hzx
                w2 = (d_hxz * kz + q3) - (d_hzx * kx + q4)
                q5 = bx * P["ezx"][c_] + cx_ * d_hyx
                q6 = by * P["ezy"][c_] + cy_ * d_hxy
                w3 = (d_hyx * kx + q5) - (d_hxy * ky + q6)
                for nm, sl, v in (("exy", a, q1), ("exz", a, q2), ("eyz", b_, q3),
                                  ("eyx", b_, q4), ("ezx", c_, q5), ("ezy", c_, q6)):
                    t = P[nm].clone()
                    t[sl] = v
                    P[nm] = t
            else:
                w1, w2, w3 = d_hzy - d_hyz, d_hxz - d_hzx, d_hyx - d_hxy
            ex2, ey2, ez2 = ex.clone(), ey.clone(), ez.clone()
            if lossy:
                ex2[a] = ca[a] * ex[a] + cb[a] * w1
                ey2[b_] = ca[b_] * ey[b_] + cb[b_] * w2
                ez2[c_] = ca[c_] * ez[c_] + cb[c_] * w3
            else:
                ex2[a] = ex[a] + cb[a] * w1
                ey2[b_] = ey[b_] + cb[b_] * w2
                ez2[c_] = ez[c_] + cb[c_] * w3
            ex, ey, ez = ex2, ey2, ez2
        if damp is not None:
            ex, ey, ez = ex * damp, ey * damp, ez * damp
        if sources:
            ex, ey, ez = ex.clone(), ey.clone(), ez.clone()
            comps = [ex, ey, ez]
            for comp, pos, vals in sources:
                comps[comp][pos] = comps[comp][pos] + float(vals[n])
            ex, ey, ez = comps
        if mon is not None:
            J = J + (ez[mon] ** 2).sum()
    return J, [ex, ey, ez, hx, hy, hz]


def gradient(size, dx, dt, boundary, eps_r: np.ndarray, sources, n_steps, npml=0, sigma=None,
             monitor=None, cpml_kw=None):
    """``(J, dJ/d eps_r)`` by reverse-mode autodiff, float64."""
    e = torch.tensor(np.asarray(eps_r, dtype=np.float64), requires_grad=True)
    J, _ = forward(size, dx, dt, boundary, e, sources, n_steps, npml, sigma, monitor, cpml_kw)
    J.backward()
    return float(J.detach()), e.grad.numpy()
