"""Fixture cases: small FDTD runs of the oracle whose outputs are committed
as tests/golden/fdtd3d_golden.npz (written by make_fixtures.py).

Each case records what the reference API takes (fdtd3d.py:41-130, 169-209)
plus the extension keywords, and what is stored of the float32 result: the
whole final state for the bit-exact cases (walls, periodic) and, for the
damped ones, the z = nz/2 plane of every saved snapshot plus the float64
norm of every component of every snapshot (``last_only``: of the final
snapshot only).  Snapshots follow the reference's subsampling: the fields
after steps 1, 1 + s, 1 + 2s, ... (fdtd3d.py:395-404), so ``save_every``
is chosen to include the last step."""

import math

import numpy as np

C0 = 299792458.0


def _sphere(size, centre, r):
    i, j, k = np.indices(size)
    return (i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2 <= r * r


def _box(size, lo, hi):
    m = np.zeros(size, bool)
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
    return m


CASES = {
    # reference semantics: sponge, three polarisations, two overlapping media
    "sponge24": dict(
        size=(24, 24, 24), dx=0.01, boundary="absorbing", pml=4,
        sources=[(4e9, (12, 12, 12), 1.0, "z"), (6e9, (8, 15, 10), 0.5, "y"),
                 (5e9, (16, 9, 14), 2.0, "x")],
        materials=[("sphere", (12, 12, 12), 5, 3.0), ("box", (14, 4, 6), (22, 12, 20), 6.0)],
        n_steps=60, save_every=20, store="planes"),
    # periodic rolls, bit-exact on the GPU
    "periodic16": dict(
        size=(16, 16, 16), dx=0.01, boundary="periodic", pml=0,
        sources=[(5e9, (8, 8, 8), 1.0, "z"), (7e9, (0, 15, 3), 1.5, "x")],
        materials=[("box", (2, 2, 2), (9, 6, 12), 2.5)],
        n_steps=40, save_every=13, store="final"),
    # undamped walls on a grid that is not a multiple of the tiles
    "walls20": dict(
        size=(20, 14, 10), dx=0.01, boundary="walls", pml=0,
        sources=[(3e9, (5, 7, 4), 1.0, "z"), (9e9, (19, 0, 9), 0.7, "y")],
        materials=[("sphere", (10, 7, 5), 4, 5.0)],
        n_steps=30, save_every=29, store="final"),
    # extension: CPML with a lossy slab and pulse sources (oracle/ext_fdtd.py)
    "cpml24": dict(
        size=(24, 22, 20), dx=1.0, boundary="cpml", pml=5,
        sources=[(2, (12, 11, 10), "gaussian_derivative", dict(t0=18, spread=5)),
                 (0, (7, 15, 8), "gaussian_sine", dict(t0=20, spread=7, wavelength=9,
                                                        amplitude=0.5))],
        materials=[("sphere", (12, 11, 10), 4, 4.0)], sigma_slab=(6, 9, 0.02),
        n_steps=70, save_every=69, store="planes", last_only=True),
}


def eps_of(case):
    size = case["size"]
    eps = np.ones(size, np.float32)
    for m in case["materials"]:
        mask = _sphere(size, m[1], m[2]) if m[0] == "sphere" else _box(size, m[1], m[2])
        eps[mask] = m[-1]
    return eps


def masks_of(case):
    """(mask, eps_r) pairs in insertion order (reference-style materials)."""
    size = case["size"]
    out = []
    for m in case["materials"]:
        mask = _sphere(size, m[1], m[2]) if m[0] == "sphere" else _box(size, m[1], m[2])
        out.append((mask, float(m[-1])))
    return out


def sigma_of(case):
    """S/m conductivity of the lossy z slab (normalised sigma_n eps0 c / dx)."""
    if "sigma_slab" not in case:
        return None
    z0, z1, sn = case["sigma_slab"]
    eps0 = 8.8541878128e-12
    s = np.zeros(case["size"])
    s[:, :, z0:z1] = sn * eps0 * C0 / case["dx"]
    return s


def dt_cpml(case):
    return 0.5 / math.sqrt(3.0) * case["dx"] / C0
