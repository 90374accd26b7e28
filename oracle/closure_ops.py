"""CPU oracle for the closure's float32 far-field operator.

TEST INFRASTRUCTURE ONLY (the checker; the product never imports it).

No reference counterpart: the reference has no closure.  The operator is a
linear 27-tap stencil of the six Yee fields (normalised by per-component
scales) with a plain and an (eps_r - 1)-weighted weight set:

    y[co](i) = sum over taps t and fields ci of
               (w[t, ci, 0, co] + (eps(i) - 1) w[t, ci, 1, co]) f_ci(i + off(t)) / scale[ci]

with off(t) = (dx, dy, dz) in {-1, 0, 1}^3, t = ((dz+1)*3 + dy+1)*3 + dx+1,
and zero outside the grid.  Arrays are in the internal (z, y, x) order.
``stencil`` evaluates it array-at-a-time in float64, ``stencil_bruteforce``
cell by cell; the CPU tests pin the first against the second.
"""

from __future__ import annotations

import numpy as np


def _offsets():
    return [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]


def stencil(fields, eps, scale, w):
    """fields: six (nz, ny, nx) arrays; eps (nz, ny, nx); scale (6,);
    w (27, 6, 2, 6).  Returns (nz, ny, nx, 6) float64."""
    f = np.stack([np.asarray(fields[c], np.float64) / float(scale[c]) for c in range(6)])
    nz, ny, nx = f.shape[1:]
    fp = np.pad(f, ((0, 0), (1, 1), (1, 1), (1, 1)))
    w = np.asarray(w, np.float64)
    a = np.zeros((nz, ny, nx, 6))
    b = np.zeros((nz, ny, nx, 6))
    for t, (dx, dy, dz) in enumerate(_offsets()):
        sh = fp[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        a += np.einsum("czyx,co->zyxo", sh, w[t, :, 0, :])
        b += np.einsum("czyx,co->zyxo", sh, w[t, :, 1, :])
    e = np.asarray(eps, np.float64)[..., None] - 1.0
    return a + e * b


def stencil_bruteforce(fields, eps, scale, w):
    """The same operator, one cell and one tap at a time (small grids)."""
    nz, ny, nx = np.asarray(fields[0]).shape
    out = np.zeros((nz, ny, nx, 6))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                e = float(eps[z, y, x]) - 1.0
                for t, (dx, dy, dz) in enumerate(_offsets()):
                    sz, sy, sx = z + dz, y + dy, x + dx
                    if not (0 <= sz < nz and 0 <= sy < ny and 0 <= sx < nx):
                        continue
                    for ci in range(6):
                        v = float(fields[ci][sz, sy, sx]) / float(scale[ci])
                        for co in range(6):
                            out[z, y, x, co] += (w[t, ci, 0, co] + e * w[t, ci, 1, co]) * v
    return out
