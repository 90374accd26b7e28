"""Device state of a 3D Yee run on the GPU and its calls into libyee_hip.so.

``YeeSolver`` owns, as torch tensors on one HIP device, the six field
arrays, the per-cell E coefficients, the boundary tables, the CPML psi
buffer and the plan; it advances them with the fused HIP kernels.  Every
per-cell array is stored x fastest, torch shape ``(nz, ny, nx)``; the
reference's ``F[i, j, k]`` is ``F_internal[k, j, i]``
(src/neurosim/em/fdtd3d.py:212-217, 240).

Coefficients (reference file:line):
* ``ch = dt/(MU0*dx)``, ``ce = dt/(EPS0*dx)``            fdtd3d.py:223-224
* ``cb = ce/eps_r`` as a float32 division                  fdtd3d.py:336-354
* sponge: three 1-D cubic profiles, ``sigma_max=3.2/dx``   fdtd3d.py:139-167
* CPML and lossy Ca/Cb are extensions (Roden & Gedney 2000; Taflove &
  Hagness 3rd ed. ch. 7 and eq. 3.31), see ``cpml_tables``.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from jaxphys_amd import _native
from jaxphys_amd.em.constants import EPS0, ETA0, MU0

FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


def sponge_profile_1d(n: int, n_pml: int, dx: float) -> np.ndarray:
    """fdtd3d.py:149-157: ``s[i] = s[n-1-i] = sigma_max*((n_pml-i)/n_pml)**3``
    written in increasing ``i`` (later writes win when the slabs overlap),
    stored as float32."""
    s = np.zeros(n, dtype=np.float32)
    if n_pml == 0:
        return s
    sigma_max = 0.8 * 4.0 / (dx * 1.0)
    for i in range(n_pml):
        val = sigma_max * ((n_pml - i) / n_pml) ** 3
        s[i] = val
        s[n - 1 - i] = val
    return s


def cpml_tables(n: int, npml: int, dx: float, dt: float, grading: float = 3.0,
                R0: float = 1e-6, kappa_max: float = 1.0, alpha_max: float = 0.0) -> np.ndarray:
    """CPML recursive-convolution coefficients along one axis.

    Returns float32 ``[bE, cE, kinvE, bH, cH, kinvH]`` (each length ``n``).
    E-type derivatives sit at integer positions ``i``, H-type at ``i+1/2``.
    Depth ``d`` is measured from the inner interface (``npml`` on the low
    side, ``n-1-npml`` on the high side), normalised by ``npml`` and clipped
    to [0, 1]; only indices in the slabs ``[0, npml)`` and ``[n-npml, n)``
    are absorbing.  ``sigma = sigma_max d^m``, ``kappa = 1+(kappa_max-1)d^m``,
    ``alpha = alpha_max (1-d)``, ``sigma_max = -(m+1) ln R0 / (2 eta0 npml dx)``,
    ``b = exp(-(sigma/kappa+alpha) dt/eps0)``,
    ``c = sigma (b-1) / (sigma kappa + kappa^2 alpha)``.
    """
    out = np.zeros((6, n), dtype=np.float64)
    sigma_max = -(grading + 1.0) * math.log(R0) / (2.0 * ETA0 * npml * dx)
    idx = np.arange(n)
    slab = (idx < npml) | (idx >= n - npml)
    for kind, half in ((0, 0.0), (1, 0.5)):
        x = idx + half
        d = np.where(idx < npml, (npml - x) / npml,
                     np.where(idx >= n - npml, (x - (n - 1 - npml)) / npml, 0.0))
        d = np.clip(d, 0.0, 1.0)
        sig = sigma_max * d**grading
        kap = 1.0 + (kappa_max - 1.0) * d**grading
        alp = np.where(slab, alpha_max * (1.0 - d), 0.0)
        b = np.exp(-(sig / kap + alp) * dt / EPS0)
        den = sig * kap + kap * kap * alp
        c = np.where(sig > 0, sig * (b - 1.0) / np.where(den > 0, den, 1.0), 0.0)
        out[3 * kind + 0] = b
        out[3 * kind + 1] = c
        out[3 * kind + 2] = 1.0 / kap
    return out.astype(np.float32)


class YeeSolver:
    """GPU state and stepping of one Yee grid (internal layout (nz, ny, nx))."""

    def __init__(self, size, dx: float, dt: float, boundary: str, pml_layers: int,
                 eps_r: torch.Tensor, sigma: torch.Tensor | None = None,
                 cpml: dict | None = None, device=None, fused: bool = True,
                 palette: bool = True):
        self.device = torch.device(device) if device is not None else _native.require_device()
        _native.lib()
        self.nx, self.ny, self.nz = (int(s) for s in size)
        self.dx = float(dx)
        self.dt = float(dt)
        self.boundary_name = boundary
        self.pml_layers = int(pml_layers)
        self.mu0 = float(MU0)
        nx, ny, nz = self.nx, self.ny, self.nz
        shape = (nz, ny, nx)
        dev = self.device
        eps = eps_r.to(device=dev, dtype=torch.float32).contiguous()
        assert tuple(eps.shape) == shape
        ch = np.float32(self.dt / (float(MU0) * self.dx))
        self.ce = self.dt / (float(EPS0) * self.dx)
        self.sigma = None
        if sigma is not None:
            sig = sigma.to(device=dev, dtype=torch.float64).contiguous()
            if bool((sig != 0).any()):
                self.sigma = sig
        self.eps_r = torch.empty_like(eps)
        self.cb = torch.empty_like(eps)
        self.ca = torch.empty_like(eps) if self.sigma is not None else None
        self.plan = None
        self.use_palette = bool(palette)
        self.mat = self.palette = None
        self.set_eps(eps)
        desc = _native.YeeGridDesc()
        desc.nx, desc.ny, desc.nz = nx, ny, nz
        desc.ch = float(ch)
        desc.cb = _native.ptr(self.cb)
        desc.ca = _native.ptr(self.ca) if self.ca is not None else None
        self.sponge = None
        self.cpml = None
        if boundary == "periodic":
            desc.boundary = _native.BOUNDARY_PERIODIC
        elif boundary == "absorbing":
            desc.boundary = _native.BOUNDARY_SPONGE
            prof = np.concatenate([sponge_profile_1d(n, self.pml_layers, self.dx)
                                   for n in (nx, ny, nz)])
            self.sponge = torch.from_numpy(prof).to(dev)
            desc.sponge_sigma = _native.ptr(self.sponge)
            desc.sponge_dt = float(np.float32(self.dt))
            desc.sponge_eps0 = float(np.float32(EPS0))
        elif boundary == "cpml":
            desc.boundary = _native.BOUNDARY_CPML
            kw = dict(cpml or {})
            tabs = np.concatenate([cpml_tables(n, self.pml_layers, self.dx, self.dt, **kw).ravel()
                                   for n in (nx, ny, nz)])
            self.cpml = torch.from_numpy(tabs).to(dev)
            desc.npml = self.pml_layers
            desc.cpml = _native.ptr(self.cpml)
        else:  # any other string: the reference's non-periodic, undamped update
            desc.boundary = _native.BOUNDARY_WALLS
        self.plan = _native.Plan(desc, keep=[self.cb, self.ca, self.sponge, self.cpml])
        self._update_palette()
        npsi = self.plan.psi_floats()
        # two state buffers: the fused kernel reads one and writes the other
        self.fused = bool(fused)
        nbuf = 2 if self.fused else 1
        self._bufs = [[torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(6)]
                      for _ in range(nbuf)]
        self._psis = [torch.zeros(npsi, dtype=torch.float32, device=dev) if npsi else None
                      for _ in range(nbuf)]
        self._cur = 0
        self.n_sources = 0
        self.table_steps = 0

    def set_eps(self, eps: torch.Tensor) -> None:
        """Recompute the E coefficients in place (the plan keeps pointing at
        the same tensors).  Lossless cells: ``cb = f32(ce) / eps_r``, a
        float32 division per cell (fdtd3d.py:336); lossy cells (extension,
        Taflove & Hagness eq. 3.31): ``x = sigma dt / (2 eps0 eps_r)``,
        ``ca = (1-x)/(1+x)``, ``cb = (ce/eps_r)/(1+x)`` in float64, rounded."""
        eps = eps.to(device=self.device, dtype=torch.float32)
        self.eps_r.copy_(eps)
        ce = self.ce
        torch.div(torch.full_like(eps, float(np.float32(ce))), eps, out=self.cb)
        if self.sigma is not None:
            e64 = eps.to(torch.float64)
            x = self.sigma * self.dt / (2.0 * EPS0 * e64)
            lossy = self.sigma != 0
            self.ca.copy_(torch.where(lossy, (1.0 - x) / (1.0 + x), torch.ones_like(x)))
            self.cb.copy_(torch.where(lossy, ((ce / e64) / (1.0 + x)).to(torch.float32), self.cb))
        if self.plan is not None:
            self._update_palette()

    def _update_palette(self) -> None:
        """When the grid holds at most 256 distinct (cb, ca) pairs (bit
        patterns), give the plan a uint8 material id per cell and the table
        of those exact values: the fused and two-pass E updates then read one
        byte per cell instead of four or eight."""
        self.mat = self.palette = None
        if self.use_palette:
            bits = self.cb.view(torch.int32).reshape(-1)
            if self.ca is not None:
                key = torch.stack([bits, self.ca.view(torch.int32).reshape(-1)], dim=1)
                uniq, inv = torch.unique(key, dim=0, return_inverse=True)
            else:
                uniq, inv = torch.unique(bits, return_inverse=True)
                uniq = uniq.view(-1, 1)
            m = uniq.shape[0]
            if m <= 256:
                pal = torch.zeros(512, dtype=torch.int32, device=self.device)
                pal[:m] = uniq[:, 0]
                if self.ca is not None:
                    pal[256:256 + m] = uniq[:, 1]
                self.palette = pal.view(torch.float32)
                self.mat = inv.to(torch.uint8).view(self.cb.shape).contiguous()
                del inv, uniq
        self.plan.set_palette(self.mat, self.palette)

    def d_coeffs_d_eps(self, region=None):
        """(dCb/deps_r, dCa/deps_r) per cell (float64) for the adjoint chain
        rule; ``region`` is an index tuple into the internal layout."""
        eps = self.eps_r.to(torch.float64)
        cb = self.cb.to(torch.float64)
        if region is not None:
            eps, cb = eps[region], cb[region]
        if self.sigma is None:
            return -cb / eps, None
        sig = self.sigma if region is None else self.sigma[region]
        x = sig * self.dt / (2.0 * EPS0 * eps)
        dcb = -cb / (eps * (1.0 + x))
        dca = 2.0 * x / (eps * (1.0 + x) ** 2)
        return dcb, dca

    # -- sources --------------------------------------------------------
    def set_sources(self, comps, positions, table: np.ndarray) -> None:
        """``comps[s]`` in 0..2 (Ex, Ey, Ez), ``positions[s] = (i, j, k)``,
        ``table`` float32 of shape ``(n_steps, n_sources)``."""
        table = np.ascontiguousarray(table, dtype=np.float32)
        if table.ndim != 2 or table.shape[1] != len(comps):
            raise ValueError("source table must be (n_steps, n_sources)")
        self.plan.set_sources(np.asarray(comps, dtype=np.int32),
                              np.asarray(positions, dtype=np.int32).reshape(-1, 3),
                              table, table.shape[0])
        self.n_sources = len(comps)
        self.table_steps = table.shape[0]
        self.source_table = table
        self.source_comps = list(comps)
        self.source_positions = [tuple(p) for p in positions]

    # -- stepping -------------------------------------------------------
    @property
    def fields(self):
        """The six current field tensors (Ex, Ey, Ez, Hx, Hy, Hz)."""
        return self._bufs[self._cur]

    @property
    def psi(self):
        """The current CPML psi buffer (None unless CPML)."""
        return self._psis[self._cur]

    # Fused step only when the slab-free interior (the box of the fused
    # interior launch) holds at least this share of the cells: its CPML shell
    # launches cost about twice the two-pass kernels per cell (on MI355X:
    # 512^3 with 16-cell slabs, 77 % interior: fused 1.78 ms vs 2.31 ms;
    # 256^3 with 16-cell slabs, 57 %: fused 0.349 ms vs 0.369 ms;
    # 160^3 with 12-cell slabs, 42 %: fused 0.123 ms vs 0.075 ms;
    # 128^3 with 8-cell slabs, 38 %: fused 0.086 ms vs 0.043 ms).
    FUSED_MIN_INTERIOR = 0.55

    def interior_fraction(self) -> float:
        """Share of the cells in the fused step's slab-free interior box."""
        if self.cpml is None:
            return 1.0
        p = self.pml_layers
        xa, xb = (p + 1 + 31) // 32 * 32, (self.nx - p) // 32 * 32
        n = (max(0, xb - xa) * max(0, self.ny - p - (p + 1)) * max(0, self.nz - 2 * p))
        return n / float(self.nx * self.ny * self.nz)

    def run(self, n0: int, nsteps: int, stream=None) -> None:
        """Steps n0 .. n0+nsteps-1 with the faster path for this grid: one
        fused step per launch set, or the H and E half-step kernels (the same
        results bit for bit)."""
        if self.fused and self.interior_fraction() >= self.FUSED_MIN_INTERIOR:
            self.run_fused(n0, nsteps, stream)
        else:
            self.run_two_pass(n0, nsteps, stream)

    def run_fused(self, n0: int, nsteps: int, stream=None) -> None:
        """Steps with the fused kernels (ping-pong between the two buffers)."""
        if nsteps <= 0:
            return
        if not self.fused:
            raise RuntimeError("solver built with fused=False has one state buffer")
        a, b = self._cur, 1 - self._cur
        _native.run_fused(self.plan, self._bufs[a], self._bufs[b], self._psis[a],
                          self._psis[b], n0, nsteps, stream)
        self._cur = a if nsteps % 2 == 0 else b

    def run_two_pass(self, n0: int, nsteps: int, stream=None) -> None:
        """Steps with the separate H and E kernels (same results bit for bit)."""
        if nsteps > 0:
            self.plan.run(self.fields, self.psi, n0, nsteps, stream)

    def step_fused(self, n: int, stream=None) -> None:
        a, b = self._cur, 1 - self._cur
        _native.step_fused(self.plan, self._bufs[a], self._bufs[b], self._psis[a],
                           self._psis[b], n, stream)
        self._cur = b

    def step_h(self, stream=None) -> None:
        self.plan.step_h(self.fields, self.psi, stream)

    def step_e(self, n: int, stream=None) -> None:
        self.plan.step_e(self.fields, self.psi, n, stream)

    def reset(self) -> None:
        for buf, ps in zip(self._bufs, self._psis):
            for f in buf:
                f.zero_()
            if ps is not None:
                ps.zero_()
        self._cur = 0

    def energy_fp64(self):
        """float64 (sum eps_r |E|^2, sum |H|^2) as Python floats."""
        w = _native.energy_fp64((self.nx, self.ny, self.nz), self.fields, self.eps_r)
        a, b = w.cpu().tolist()
        return a, b

    def em_energy_joules(self) -> float:
        """Discrete field energy 0.5 eps0 sum eps_r|E|^2 dx^3 + 0.5 mu0 sum|H|^2 dx^3."""
        we, wh = self.energy_fp64()
        v = self.dx**3
        return 0.5 * EPS0 * we * v + 0.5 * MU0 * wh * v

    def reference_view(self, t: torch.Tensor) -> torch.Tensor:
        """(nz, ny, nx) storage as the reference's (nx, ny, nz) index order."""
        return t.permute(2, 1, 0)

    def close(self) -> None:
        self.plan.close()
