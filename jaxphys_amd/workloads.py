"""Synthetic workloads of the FDTD3D benchmarks (fixed seeds).

Each builder returns an ``EMGrid3D`` ready to ``simulate``/``build_solver``
plus the step count.  Units are normalised: dx = 1 (``resolution=1.0``) and
``dt = S dx / C0`` with the Courant number S = 0.5/sqrt(3), which map onto
the reference's SI API exactly because ch and ce depend on S only
(src/neurosim/em/fdtd3d.py:223-224).  Normalised conductivity sigma_n (in
units of eps0 c / dx) becomes ``sigma_n * EPS0 * C0 / dx`` S/m.

Pulse definitions (time in steps n):
  gaussian_derivative  -((n-t0)/s) exp(-((n-t0)/s)^2/2)
  gaussian_sine        sin(2 pi S n / wavelength) exp(-((n-t0)/s)^2/2)
  gaussian             exp(-((n-t0)/s)^2/2)
"""

from __future__ import annotations

import math

import numpy as np
import torch

from jaxphys_amd.em.constants import C0, EPS0
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D

COURANT = 0.5 / math.sqrt(3.0)
SEEDS = {"c2_128": 1234, "c3_512": 5678, "c4_160": 42}


def dt_for(dx: float = 1.0, courant: float = COURANT) -> float:
    return courant * dx / C0


def c1_sphere(n: int = 64, reference_compat: bool = False):
    """64^3, 500 steps, 10-cell CPML (m=3, R0=1e-6), eps_r=4 sphere of radius
    12 at the centre, z-dipole 8 cells off-centre along +x with a Gaussian
    derivative pulse (t0=60, spread=20).  ``reference_compat`` swaps in the
    reference's sponge and sine source (wavelength 20 cells)."""
    size = (n, n, n)
    c = n // 2
    if reference_compat:
        g = EMGrid3D(size, 1.0, "absorbing", 10)
    else:
        g = EMGrid3D(size, 1.0, "cpml", 10, pml_R0=1e-6, pml_grading=3.0)
    I, J, K = np.indices(size)
    g.add_material(DielectricRegion(mask=(I - c) ** 2 + (J - c) ** 2 + (K - c) ** 2 <= 144,
                                    epsilon_r=4.0))
    if reference_compat:
        g.add_source(PointSource3D(C0 / 20.0, (c + 8, c, c)))
    else:
        g.add_source(PointSource3D(0.0, (c + 8, c, c), waveform="gaussian_derivative",
                                   t0=60, spread=20))
    return g, 500


def c2_boxes(n: int = 128, seed: int = SEEDS["c2_128"]):
    """128^3, 1000 steps, 12-cell CPML; 40 seeded boxes (sides U{4..32},
    eps_r U[1,12], later boxes overwrite earlier ones); one z-dipole at a
    seeded interior point, Gaussian-modulated sine (wavelength 20 cells,
    spread 30, t0 90)."""
    rng = np.random.default_rng(seed)
    size = (n, n, n)
    g = EMGrid3D(size, 1.0, "cpml", 12)
    for _ in range(40):
        s = rng.integers(4, 33, 3)
        o = [int(rng.integers(0, n - s[a] + 1)) for a in range(3)]
        m = np.zeros(size, bool)
        m[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]] = True
        g.add_material(DielectricRegion(mask=m, epsilon_r=float(rng.uniform(1.0, 12.0))))
    p = tuple(int(v) for v in rng.integers(24, n - 24, 3))
    g.add_source(PointSource3D(0.0, p, waveform="gaussian_sine", wavelength=20.0, t0=90.0,
                               spread=30.0))
    return g, 1000


def gaussian_random_eps(n: int, seed: int, corr: float = 8.0, device=None) -> torch.Tensor:
    """eps_r in {1, 9}: N(0,1) float32 white noise (seeded) filtered by an
    isotropic Gaussian of standard deviation ``corr`` cells (FFT, periodic)
    and thresholded at 0.  Returned in (nx, ny, nz) index order."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    noise = torch.randn((n, n, n), generator=gen, dtype=torch.float32)
    dev = device if device is not None else noise.device
    noise = noise.to(dev)
    k = torch.fft.fftfreq(n, device=dev, dtype=torch.float32) * (2 * math.pi)
    g1 = torch.exp(-0.5 * (k * corr) ** 2)
    spec = torch.fft.rfftn(noise)
    kz = torch.fft.rfftfreq(n, device=dev, dtype=torch.float32) * (2 * math.pi)
    gz = torch.exp(-0.5 * (kz * corr) ** 2)
    spec *= g1[:, None, None] * g1[None, :, None] * gz[None, None, :]
    field = torch.fft.irfftn(spec, s=(n, n, n))
    del spec, noise
    return torch.where(field > 0, torch.tensor(9.0, device=dev), torch.tensor(1.0, device=dev))


def c3_grf(n: int = 512, seed: int = SEEDS["c3_512"], device=None):
    """512^3, 500 steps, 16-cell CPML; Gaussian-random-field eps_r in {1, 9}
    (correlation length 8 cells) plus a conductive slab sigma_n = 0.01 over
    z in [n*400/512, n*420/512); three dipoles (x, y, z) at seeded points
    at least 32 cells (n/16) from every face, Gaussian-modulated sine
    (wavelength 20 cells, spread 30, t0 90)."""
    size = (n, n, n)
    g = EMGrid3D(size, 1.0, "cpml", max(4, n * 16 // 512), device=device)
    eps = gaussian_random_eps(n, seed, device=device)
    g.add_material(DielectricRegion(mask=torch.ones(size, dtype=torch.bool, device=eps.device),
                                    epsilon_r=eps))
    z0, z1 = n * 400 // 512, n * 420 // 512
    slab = torch.zeros(size, dtype=torch.bool, device=eps.device)
    slab[:, :, z0:z1] = True
    g.add_material(DielectricRegion(mask=slab, epsilon_r=None,
                                    sigma=0.01 * EPS0 * C0 / 1.0))
    rng = np.random.default_rng(seed)
    margin = max(4, n // 16)
    for pol in "xyz":
        p = tuple(int(v) for v in rng.integers(margin, n - margin, 3))
        g.add_source(PointSource3D(0.0, p, polarization=pol, waveform="gaussian_sine",
                                   wavelength=20.0, t0=90.0, spread=30.0))
    return g, 500


def bytes_per_cell_step(lossy: bool) -> int:
    """Algorithmic float32 bytes per cell per full step in the two-pass
    design: H pass reads 6 fields and writes 3 (36 B); E pass reads 6 fields
    and Cb (and Ca when lossy) and writes 3 (40 / 44 B)."""
    return 76 + (4 if lossy else 0)


def psi_bytes_per_step(size, npml: int) -> int:
    """CPML: 12 psi arrays, each on the 2*npml-thick slab normal to its
    derivative axis, read and written once per step (8 B per psi cell)."""
    nx, ny, nz = size
    tp = 2 * npml
    return 8 * 4 * (nz * tp * nx + tp * ny * nx + nz * ny * tp)


def c4_design(n: int = 160, device=None, seed: int = SEEDS["c4_160"]):
    """Adjoint workload: n^3 (160), 12-cell CPML, vacuum background, central
    design region [n/5, 4n/5)^3 (96^3 at 160) with eps_r U[1,4] (seeded), a
    z-dipole 20 cells below the region (Gaussian pulse), and the objective's
    probes: Ez on a 32x32 patch on the plane 20 cells above the region.
    Returns (grid, design box (x0,y0,z0,nx,ny,nz), probes, eps_box)."""
    size = (n, n, n)
    g = EMGrid3D(size, 1.0, "cpml", max(4, 12 * n // 160), device=device)
    lo, hi = n // 5, 4 * n // 5
    box = (lo, lo, lo, hi - lo, hi - lo, hi - lo)
    c = n // 2
    npml = max(4, 12 * n // 160)
    zsrc = lo - 20 if lo - 20 >= npml else (lo + npml) // 2
    g.add_source(PointSource3D(0.0, (c, c, zsrc),
                               waveform="gaussian_sine", wavelength=20.0, t0=90.0, spread=30.0))
    zmon = min(hi + 19, n - 14)
    half = min(16, (hi - lo) // 6)
    probes = [("ez", (i, j, zmon)) for i in range(c - half, c + half)
              for j in range(c - half, c + half)]
    rng = np.random.default_rng(seed)
    eps_box = torch.tensor(rng.uniform(1.0, 4.0, (hi - lo,) * 3), dtype=torch.float32,
                           device=device)
    return g, box, probes, eps_box
