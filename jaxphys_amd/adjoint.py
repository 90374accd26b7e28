"""d(objective)/d(eps_r) on the GPU with the hand-written adjoint.

No reference counterpart: the reference would differentiate
``EMGrid3D.simulate`` with generic ``jax.grad`` (src/neurosim/optimize.py:84,
reverse mode through ``lax.scan`` at src/neurosim/em/fdtd3d.py:393), which
keeps residuals of every step (O(n_steps * N^3) memory).  Here:

The adjoint sweep needs, of the forward run, only the probe values and,
per step, d = E^{n+1} - Ca E^n on the design box (plus E^n for lossy
plans), which the forward E kernel records from its registers.  When that
box history fits in a quarter of the free device memory (``record_box``,
automatic), the forward run records it for all steps and the backward pass
is a single adjoint sweep.  Otherwise, checkpointing:

1. forward run on the GPU, saving the full state (6 fields + CPML psi) every
   ``checkpoint_every`` steps and the probe values after every step;
2. for each checkpoint segment, last first: restore the checkpoint, recompute
   the segment forward while recording the design box for every step, then
   run the adjoint steps of the segment in reverse (``yee_adjoint_run``),
   which accumulate ``sum_n g d`` and ``sum_n g E^n`` per design cell.  The
   last segment's box is recorded during the forward run itself (the sweep
   starts there, so it needs no recomputation);

and in both cases the chain rule to eps_r:
``dJ/deps = acc_cb/Cb * dCb/deps + acc_ca * dCa/deps``.

Objective: ``J = sum_n sum_q w_q F_q(after step n)^2`` over probes q
(component, cell), e.g. the time-integrated |Ez|^2 on a monitor patch.
``FDTDObjective`` wraps it as a ``torch.autograd.Function`` of the design
permittivities.
"""

from __future__ import annotations

import numpy as np
import torch

from jaxphys_amd import _native
from jaxphys_amd.exceptions import ConfigurationError

_COMP = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}


def probe_terms(probes, weights, size):
    """Objective terms of ``probes`` (``[(component, (i, j, k)), ...]``, E
    components only) with ``weights`` (None: all 1) on a grid of ``size``
    (nx, ny, nz): probes at the same (component, cell) merge into one term
    with the summed weight (the adjoint adds one term per component and
    cell).  Returns (keys [(c, i, j, k)], weights, probe_box, probe_map):
    the box (x0, y0, z0, nx, ny, nz) bounding the probe cells and the map
    over it, int32 (3, nz, ny, nx) holding each term's index at its
    component and cell and -1 elsewhere, when the box holds at most
    max(2^20, 8 * terms) cells; else (.., None, None)."""
    nx, ny, nz = size
    probes = list(probes)
    w_in = np.ones(len(probes)) if weights is None else np.asarray(weights, np.float64)
    if w_in.shape != (len(probes),):
        raise ConfigurationError("one weight per probe")
    merged: dict[tuple[int, int, int, int], float] = {}
    for (comp, (i, j, k)), wq in zip(probes, w_in):
        c = _COMP[comp] if isinstance(comp, str) else int(comp)
        if c not in (0, 1, 2):
            # the psi_H adjoint update is deferred by one step and reads
            # lamH unmodified; objective terms are therefore E-only
            raise ConfigurationError("adjoint objectives are sums over E components")
        if not (0 <= i < nx and 0 <= j < ny and 0 <= k < nz):
            raise ConfigurationError(f"probe {(i, j, k)} outside the grid")
        key = (c, int(i), int(j), int(k))
        merged[key] = merged.get(key, 0.0) + float(wq)
    keys = list(merged)
    wts = [merged[key] for key in keys]
    if not keys:
        return keys, wts, None, None
    pi = np.array([[i, j, k] for _, i, j, k in keys])
    lo, hi = pi.min(0), pi.max(0) + 1
    ext = hi - lo
    if int(np.prod(ext)) > max(1 << 20, 8 * len(keys)):
        return keys, wts, None, None
    pmap = np.full((3, ext[2], ext[1], ext[0]), -1, dtype=np.int32)
    for q, (c, i, j, k) in enumerate(keys):
        pmap[c, k - lo[2], j - lo[1], i - lo[0]] = q
    box = (int(lo[0]), int(lo[1]), int(lo[2]), int(ext[0]), int(ext[1]), int(ext[2]))
    return keys, wts, box, pmap


class AdjointProblem:
    """Objective and gradient w.r.t. eps_r on a design box of an EMGrid3D.

    Args:
        grid: the EMGrid3D (materials give the background eps_r; sources are
            fixed).  Must not be periodic.
        n_steps, dt: the run.
        design_box: ``(x0, y0, z0, nx, ny, nz)`` in the reference's (i, j, k)
            index order.  Must lie outside absorbing slabs, at i, j, k >= 1,
            and contain no source.
        probes: ``[(component, (i, j, k)), ...]`` with component in
            ex, ey, ez, and optional ``weights``.  H components are
            rejected: the adjoint of the H-part psi update is applied one
            step late, so an objective term cannot be added to the H
            adjoint at the step it belongs to.
        checkpoint_every: forward checkpoint interval K (when the box
            history is not recorded).
        record_box: record E on the design box for every step during the
            forward run (no recomputation); None = when it fits in a quarter
            of the free device memory.
    """

    def __init__(self, grid, n_steps: int, dt: float, design_box, probes, weights=None,
                 checkpoint_every: int = 100, record_box: bool | None = None):
        self.grid = grid
        self.n_steps = int(n_steps)
        self.K = max(1, int(checkpoint_every))
        self.dt = float(dt)
        # the design eps changes every evaluation: per-cell arrays only
        sol = grid.build_solver(self.dt, self.n_steps, palette=False)
        self.solver = sol
        dev = sol.device
        nx, ny, nz = sol.nx, sol.ny, sol.nz
        x0, y0, z0, bx, by, bz = (int(v) for v in design_box)
        self.box_ref = (x0, y0, z0, bx, by, bz)
        if bx <= 0 or by <= 0 or bz <= 0 or x0 < 1 or y0 < 1 or z0 < 1 or \
                x0 + bx > nx or y0 + by > ny or z0 + bz > nz:
            raise ConfigurationError(f"design box {design_box} must lie inside the grid at "
                                     "indices >= 1")
        if sol.boundary_name == "periodic":
            raise ConfigurationError("the adjoint is provided for non-periodic grids")
        p = sol.pml_layers
        if sol.boundary_name == "cpml":
            if x0 < p or y0 < p or z0 < p or x0 + bx > nx - p or y0 + by > ny - p or \
                    z0 + bz > nz - p:
                raise ConfigurationError("design box overlaps the CPML slabs")
        if sol.boundary_name == "absorbing" and sol.sponge is not None:
            prof = sol.sponge.cpu().numpy()
            sx, sy, sz = prof[:nx], prof[nx:nx + ny], prof[nx + ny:]
            if (sx[x0:x0 + bx].any() or sy[y0:y0 + by].any() or sz[z0:z0 + bz].any()):
                raise ConfigurationError("design box overlaps the sponge layers")
        for (i, j, k) in sol.source_positions:
            if x0 <= i < x0 + bx and y0 <= j < y0 + by and z0 <= k < z0 + bz:
                raise ConfigurationError("design box contains a source")
        self.region = (slice(z0, z0 + bz), slice(y0, y0 + by), slice(x0, x0 + bx))
        self.box_int = (x0, y0, z0, bx, by, bz)
        self.bcells = bx * by * bz
        keys, wts, self.probe_box, pmap = probe_terms(probes, weights, (nx, ny, nz))
        self.nprobe = len(keys)
        self.probe_comp = torch.tensor([c for c, _, _, _ in keys], dtype=torch.int32, device=dev)
        self.probe_idx = torch.tensor([(k * ny + j) * nx + i for _, i, j, k in keys],
                                      dtype=torch.int64, device=dev)
        self.w = torch.tensor(wts, dtype=torch.float32, device=dev)
        self.probe_map = None if pmap is None else torch.from_numpy(pmap.reshape(-1)).to(dev)
        self.background = sol.eps_r.clone()
        self.probe_hist = torch.zeros((self.n_steps, self.nprobe), dtype=torch.float32, device=dev)
        self.checkpoints = []
        self.box_hist = None
        # one step's box record: d = E^{n+1} - Ca E^n, and E^n for lossy plans
        self.rec_floats = 3 * self.bcells * (2 if sol.ca is not None else 1)
        need = self.n_steps * self.rec_floats * 4
        if record_box is None:
            free, _ = torch.cuda.mem_get_info(dev)
            record_box = need <= free // 4
        self.record_box = bool(record_box)
        self.seg_buf = None
        self._recorded = False
        # number of forward() calls so far: backward() uses what the latest
        # forward left (probe history, checkpoints or box history, eps), so
        # a caller holding an older forward can check it is still current
        self.forward_count = 0

    # ------------------------------------------------------------------
    def set_design(self, eps_box: torch.Tensor) -> None:
        """``eps_box``: (bx, by, bz) in the reference index order."""
        x0, y0, z0, bx, by, bz = self.box_ref
        if tuple(eps_box.shape) != (bx, by, bz):
            raise ConfigurationError(f"design eps shape {tuple(eps_box.shape)} != {(bx, by, bz)}")
        eps = self.background.clone()
        eps[self.region] = eps_box.detach().to(eps.device, torch.float32).permute(2, 1, 0)
        self.solver.set_eps(eps)

    def forward(self, eps_box: torch.Tensor | None = None) -> float:
        """Forward run with checkpoints; returns J (float64 on the host)."""
        sol = self.solver
        self.forward_count += 1
        if eps_box is not None:
            self.set_design(eps_box)
        sol.reset()
        self.checkpoints = []
        N, K = self.n_steps, self.K
        if self.record_box:
            if self.box_hist is None or self.box_hist.shape[0] != N:
                self.box_hist = torch.empty((N, self.rec_floats), dtype=torch.float32,
                                            device=sol.device)
            _native.run_record(sol.plan, sol.fields, sol.psi, 0, N, self.probe_comp,
                               self.probe_idx, self.probe_hist, box=self.box_int,
                               box_out=self.box_hist, probe_box=self.probe_box,
                               probe_map=self.probe_map)
            self._recorded = True
            ph = self.probe_hist.to(torch.float64)
            return float((self.w.to(torch.float64) * ph * ph).sum())
        self._recorded = False
        if self.seg_buf is None or self.seg_buf.shape[0] < K:
            self.seg_buf = torch.empty((K, self.rec_floats), dtype=torch.float32,
                                       device=sol.device)
        starts = list(range(0, N, K))
        for q, s0 in enumerate(starts):
            self.checkpoints.append(([f.clone() for f in sol.fields],
                                     sol.psi.clone() if sol.psi is not None else None))
            ns = min(K, N - s0)
            last = q == len(starts) - 1
            # the last segment is recorded on the way (the backward pass
            # starts there and needs no recomputation of it)
            _native.run_record(sol.plan, sol.fields, sol.psi, s0, ns, self.probe_comp,
                               self.probe_idx, self.probe_hist[s0:s0 + ns],
                               box=self.box_int if last else None,
                               box_out=self.seg_buf if last else None,
                               probe_box=self.probe_box, probe_map=self.probe_map)
        ph = self.probe_hist.to(torch.float64)
        return float((self.w.to(torch.float64) * ph * ph).sum())

    def backward(self) -> torch.Tensor:
        """dJ/d eps_r on the design box (float64, reference index order)."""
        sol = self.solver
        dev = sol.device
        N, K = self.n_steps, self.K
        if not self.checkpoints and not self._recorded:
            raise RuntimeError("backward() needs a forward() first")
        lam = [torch.zeros_like(f) for f in sol.fields]
        lam_psi = torch.zeros_like(sol.psi) if sol.psi is not None else None
        acc_cb = torch.zeros(self.bcells, dtype=torch.float32, device=dev)
        acc_ca = torch.zeros(self.bcells, dtype=torch.float32, device=dev) \
            if sol.ca is not None else None
        # adjoint of the final E: 2 w F(after step N-1)
        _native.probe_scatter_add(lam, self.probe_comp, self.probe_idx, self.w,
                                  self.probe_hist[N - 1], 2.0)

        def sweep(n_lo, n_hi, box_hist):
            _native.adjoint_run(sol.plan, lam, lam_psi, n_lo, n_hi, self.probe_comp,
                                self.probe_idx, self.w, self.probe_hist, 2.0, box=self.box_int,
                                box_hist=box_hist, acc_cb=acc_cb, acc_ca=acc_ca,
                                probe_box=self.probe_box, probe_map=self.probe_map)

        if self._recorded:
            sweep(0, N, self.box_hist)
            return self._chain_rule(acc_cb, acc_ca)
        for seg in range(len(self.checkpoints) - 1, -1, -1):
            s0, s1 = seg * K, min(N, seg * K + K)
            if seg < len(self.checkpoints) - 1:
                # restore the checkpoint and recompute the segment, recording
                # the design box (the last one was recorded by forward())
                fields, psi = self.checkpoints[seg]
                for f, c in zip(sol.fields, fields):
                    f.copy_(c)
                if psi is not None:
                    sol.psi.copy_(psi)
                _native.run_record(sol.plan, sol.fields, sol.psi, s0, s1 - s0, box=self.box_int,
                                   box_out=self.seg_buf)
            sweep(s0, s1, self.seg_buf)
        return self._chain_rule(acc_cb, acc_ca)

    def _chain_rule(self, acc_cb, acc_ca) -> torch.Tensor:
        sol = self.solver
        x0, y0, z0, bx, by, bz = self.box_ref
        shape = (bz, by, bx)
        dcb, dca = sol.d_coeffs_d_eps(self.region)
        cb = sol.cb[self.region].to(torch.float64)
        g = acc_cb.view(shape).to(torch.float64) / cb * dcb
        if acc_ca is not None:
            g = g + acc_ca.view(shape).to(torch.float64) * dca
        return g.permute(2, 1, 0).contiguous()

    def value_and_grad(self, eps_box: torch.Tensor):
        J = self.forward(eps_box)
        return J, self.backward()


class FDTDObjective(torch.autograd.Function):
    """``J = FDTDObjective.apply(eps_box, problem)``; ``J.backward()`` runs
    the checkpointed adjoint and returns dJ/d eps_box.

    The adjoint sweep uses the state of the problem's latest forward run, so
    the backward of an ``apply`` must come before the next ``apply`` on the
    same problem (one problem per objective term otherwise); a backward
    whose forward is no longer the latest raises ``RuntimeError``."""

    @staticmethod
    def forward(ctx, eps_box, problem):
        J = problem.forward(eps_box)
        ctx.problem = problem
        ctx.count = problem.forward_count
        ctx.dtype = eps_box.dtype
        ctx.device = eps_box.device
        return torch.tensor(J, dtype=torch.float64, device=eps_box.device)

    @staticmethod
    def backward(ctx, gJ):
        if ctx.problem.forward_count != ctx.count:
            raise RuntimeError(
                "FDTDObjective.backward: the AdjointProblem ran another forward pass after "
                "this one; its adjoint would be taken at the wrong design (use one "
                "AdjointProblem per objective term, or call backward before the next apply)")
        g = ctx.problem.backward().to(device=ctx.device)
        return (gJ.to(g.dtype) * g).to(ctx.dtype), None
