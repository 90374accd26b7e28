"""Learned dispersion closure: a coarse (2x) Yee grid plus a small 3D conv net
that reproduces fine-grid fields.  No reference counterpart (the reference
has no neural-network code); the FDTD solver itself generates the training
data and stays the fallback and the judge of the error.

Network (``ClosureNet``): channels 7 -> 32 -> 32 -> 6 (E, H normalised and
eps_r in; a field correction out), 3x3x3 zero-padded convolutions, exact-erf
GELU after the first two layers.  Weights and activations are bfloat16 with
float32 accumulation on the matrix cores (libconv3d_hip.so); float32 master
weights are updated by Adam.  Forward and backward run on the project's own
kernels: ``conv3d_fwd`` (bias + GELU fused; GELU derivative fused into the
input-gradient pass), ``conv3d_flip_weights`` and ``conv3d_wgrad``.

Grids: the fine grid has dx = 1, the coarse one dx = 2, both with the same
Courant number, so one coarse step spans two fine steps.  Staggered
restriction (Yee positions, fine units): coarse Ex at (2I+1, 2J, 2K) is the
mean of fine Ex[2I, 2J, 2K] and Ex[2I+1, 2J, 2K] (likewise Ey, Ez along
their own axes); coarse Hx at (2I, 2J+1, 2K+1) is the mean of the four fine
Hx[2I, 2J+{0,1}, 2K+{0,1}] (likewise Hy, Hz).  In time, coarse E^m is fine
E^{2m}; coarse H^{m-1/2} lives at fine time 2m-1, the mean of fine
H^{2m-3/2} and H^{2m-1/2}.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from jaxphys_amd import _conv
from jaxphys_amd.em.constants import C0
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D

CIN, HID, COUT = 8, 32, 6          # padded input channels (7 real), hidden, output
KP1, KP2 = _conv.kpad_of(CIN), _conv.kpad_of(HID)
OUT_STRIDE = 8                      # float32 output channel stride


def _bf16(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16)


class ClosureNet:
    """3-layer 3D conv net on the BF16 MFMA kernels (see module docstring)."""

    def __init__(self, device, seed: int = 0, out_scale: float = 1e-2):
        self.device = torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)

        def init(cout, cin_real, cin_pad, scale):
            w = torch.zeros(cout, 27, cin_pad)
            w[:, :, :cin_real] = torch.randn(cout, 27, cin_real, generator=g) * scale
            return w.to(self.device)

        self.params = {
            "w1": init(HID, 7, CIN, math.sqrt(2.0 / (27 * 7))),
            "b1": torch.zeros(HID, device=self.device),
            "w2": init(HID, HID, HID, math.sqrt(2.0 / (27 * HID))),
            "b2": torch.zeros(HID, device=self.device),
            "w3": init(COUT, HID, HID, out_scale * math.sqrt(1.0 / (27 * HID))),
            "b3": torch.zeros(COUT, device=self.device),
        }
        self.adam = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self.params.items()}
        self.step_count = 0
        # normalisation (set from data): fields / scale, (eps - shift) / escale
        self.scale = torch.ones(6, device=self.device)
        self.eps_shift, self.eps_scale = 1.0, 1.0
        self._packed = None
        self._ws = {}

    # -- weights in the kernel layout -----------------------------------
    @staticmethod
    def _pack_w(w: torch.Tensor, cout_pad: int, kpad: int) -> torch.Tensor:
        cout, _, cin_pad = w.shape
        out = torch.zeros(cout_pad, kpad, dtype=torch.bfloat16, device=w.device)
        out[:cout, :27 * cin_pad] = w.reshape(cout, 27 * cin_pad).to(torch.bfloat16)
        return out

    def pack(self):
        p = self.params
        self._packed = {
            "w1": self._pack_w(p["w1"], 32, KP1),
            "w2": self._pack_w(p["w2"], 32, KP2),
            "w3": self._pack_w(p["w3"], 16, KP2),
        }
        return self._packed

    def _buf(self, name, shape, dtype):
        b = self._ws.get(name)
        if b is None or b.shape != torch.Size(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device=self.device)
            self._ws[name] = b
        return b

    # -- forward / backward ----------------------------------------------
    def forward(self, x: torch.Tensor, shape_xyz, train: bool = False) -> torch.Tensor:
        """``x``: bf16 (npos, 8) channels last over (nz, ny, nx); returns the
        float32 correction (npos, 8) (channels 0..5 real)."""
        if self._packed is None:
            self.pack()
        nx, ny, nz = shape_xyz
        n = nx * ny * nz
        P, W = self.params, self._packed
        d1 = _conv.desc(shape_xyz, CIN, HID, 32, HID)
        d2 = _conv.desc(shape_xyz, HID, HID, 32, HID)
        d3 = _conv.desc(shape_xyz, HID, COUT, 16, OUT_STRIDE)
        a1 = self._buf("a1", (n, HID), torch.bfloat16)
        a2 = self._buf("a2", (n, HID), torch.bfloat16)
        z1 = self._buf("z1", (n, HID), torch.bfloat16) if train else None
        z2 = self._buf("z2", (n, HID), torch.bfloat16) if train else None
        y = self._buf("y", (n, OUT_STRIDE), torch.float32)
        _conv.fwd(d1, x, W["w1"], P["b1"], _conv.EPI_GELU, a1, z_out=z1)
        _conv.fwd(d2, a1, W["w2"], P["b2"], _conv.EPI_GELU, a2, z_out=z2)
        _conv.fwd(d3, a2, W["w3"], P["b3"], _conv.EPI_LINEAR_F32, y)
        if train:
            self._saved = (x, a1, a2, z1, z2, shape_xyz)
        return y

    def backward(self, dy: torch.Tensor) -> dict:
        """``dy``: float32 (npos, 8) gradient of the loss w.r.t. the output.
        Returns float32 gradients shaped like ``params``."""
        x, a1, a2, z1, z2, shape_xyz = self._saved
        nx, ny, nz = shape_xyz
        n = nx * ny * nz
        W = self._packed
        dyb = dy.to(torch.bfloat16).contiguous()
        d1 = _conv.desc(shape_xyz, CIN, HID, 32, HID)
        d2 = _conv.desc(shape_xyz, HID, HID, 32, HID)
        d3 = _conv.desc(shape_xyz, HID, COUT, 16, OUT_STRIDE)
        work = self._buf("work", (max(_conv.wgrad_work_floats(d) for d in (d1, d2, d3)),),
                         torch.float32)
        grads = {}
        # layer 3
        dw3 = self._buf("dw3", (16, KP2), torch.float32)
        db3 = self._buf("db3", (COUT,), torch.float32)
        _conv.wgrad(d3, a2, dyb, OUT_STRIDE, work, dw3, db3)
        w3t = self._buf("w3t", (32, KP1), torch.bfloat16)
        _conv.flip_weights(d3, W["w3"], OUT_STRIDE, w3t)
        dz2 = self._buf("dz2", (n, HID), torch.bfloat16)
        _conv.fwd(_conv.desc(shape_xyz, OUT_STRIDE, HID, 32, HID), dyb, w3t, None,
                  _conv.EPI_GELU_BWD, dz2, aux=z2)
        # layer 2
        dw2 = self._buf("dw2", (32, KP2), torch.float32)
        db2 = self._buf("db2", (HID,), torch.float32)
        _conv.wgrad(d2, a1, dz2, HID, work, dw2, db2)
        w2t = self._buf("w2t", (32, KP2), torch.bfloat16)
        _conv.flip_weights(d2, W["w2"], HID, w2t)
        dz1 = self._buf("dz1", (n, HID), torch.bfloat16)
        _conv.fwd(_conv.desc(shape_xyz, HID, HID, 32, HID), dz2, w2t, None, _conv.EPI_GELU_BWD,
                  dz1, aux=z1)
        # layer 1
        dw1 = self._buf("dw1", (32, KP1), torch.float32)
        db1 = self._buf("db1", (HID,), torch.float32)
        _conv.wgrad(d1, x, dz1, HID, work, dw1, db1)
        grads["w1"] = dw1[:HID, :27 * CIN].reshape(HID, 27, CIN).clone()
        grads["w1"][:, :, 7:] = 0
        grads["b1"] = db1.clone()
        grads["w2"] = dw2[:HID, :27 * HID].reshape(HID, 27, HID).clone()
        grads["b2"] = db2.clone()
        grads["w3"] = dw3[:COUT, :27 * HID].reshape(COUT, 27, HID).clone()
        grads["b3"] = db3.clone()
        return grads

    def adam_step(self, grads: dict, lr: float, betas=(0.9, 0.999), eps=1e-8) -> None:
        self.step_count += 1
        b1, b2 = betas
        for k, g in grads.items():
            m, v = self.adam[k]
            m.mul_(b1).add_(g, alpha=1 - b1)
            v.mul_(b2).addcmul_(g, g, value=1 - b2)
            mh = m / (1 - b1 ** self.step_count)
            vh = v / (1 - b2 ** self.step_count)
            self.params[k].sub_(lr * mh / (vh.sqrt() + eps))
        self._packed = None

    # -- fields <-> network tensors --------------------------------------
    def pack_fields(self, solver) -> torch.Tensor:
        n = solver.nx * solver.ny * solver.nz
        x = self._buf("xin", (n, CIN), torch.bfloat16)
        _conv.pack_fields((solver.nx, solver.ny, solver.nz), solver.fields, solver.eps_r,
                          self.scale, self.eps_shift, self.eps_scale, x)
        return x

    def correct(self, solver, gain: float = 1.0) -> None:
        """fields += net(fields, eps) (denormalised)."""
        x = self.pack_fields(solver)
        y = self.forward(x, (solver.nx, solver.ny, solver.nz))
        _conv.apply_correction((solver.nx, solver.ny, solver.nz), solver.fields, y, OUT_STRIDE,
                               self.scale, gain)

    def state_dict(self):
        return {"params": {k: v.detach().cpu() for k, v in self.params.items()},
                "scale": self.scale.cpu(), "eps_shift": self.eps_shift,
                "eps_scale": self.eps_scale}


# ---------------------------------------------------------------------------
# staggered restriction and the fine/coarse problem family
# ---------------------------------------------------------------------------

def restrict_e(e: list) -> list:
    """Fine E (internal (nz, ny, nx)) -> coarse E (module docstring)."""
    ex, ey, ez = e
    return [0.5 * (ex[::2, ::2, 0::2] + ex[::2, ::2, 1::2]),
            0.5 * (ey[::2, 0::2, ::2] + ey[::2, 1::2, ::2]),
            0.5 * (ez[0::2, ::2, ::2] + ez[1::2, ::2, ::2])]


def restrict_h(h: list) -> list:
    hx, hy, hz = h
    return [0.25 * (hx[0::2, 0::2, ::2] + hx[0::2, 1::2, ::2] + hx[1::2, 0::2, ::2]
                    + hx[1::2, 1::2, ::2]),
            0.25 * (hy[0::2, ::2, 0::2] + hy[0::2, ::2, 1::2] + hy[1::2, ::2, 0::2]
                    + hy[1::2, ::2, 1::2]),
            0.25 * (hz[::2, 0::2, 0::2] + hz[::2, 0::2, 1::2] + hz[::2, 1::2, 0::2]
                    + hz[::2, 1::2, 1::2])]


COURANT = 0.5 / math.sqrt(3.0)


class ClosureProblem:
    """One seeded fine/coarse pair: fine n^3 (dx=1) and coarse (n/2)^3 (dx=2)
    grids with random dielectric boxes (eps_r in [1, emax]), a CPML of
    ``npml`` fine cells and a z-directed Gaussian-pulse dipole.

    With the default Gaussian pulse the E increments do not sum to zero: the
    source leaves a static near field behind, whose grid representation
    differs between the two resolutions by a fixed ratio, and at late times
    that remnant dominates the coarse-vs-fine difference.
    ``waveform="gaussian_derivative"`` (zero-sum increments) removes it and
    leaves the dispersive error of the travelling pulse."""

    def __init__(self, seed: int, n_fine: int = 256, npml_fine: int = 16, emax: float = 6.0,
                 nbox: int = 24, t0: float = 60.0, spread: float = 15.0,
                 waveform: str = "gaussian", device=None):
        self.seed = seed
        self.nf = n_fine
        self.nc = n_fine // 2
        rng = np.random.default_rng(seed)
        nc = self.nc
        # boxes on the coarse lattice so the fine and coarse media agree exactly
        eps_c = np.ones((nc, nc, nc), np.float32)
        lo, hi = npml_fine // 2 + 2, nc - npml_fine // 2 - 2
        for _ in range(nbox):
            s = rng.integers(3, max(4, nc // 5), 3)
            o = [int(rng.integers(lo, max(lo + 1, hi - s[a]))) for a in range(3)]
            eps_c[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]] = rng.uniform(1.0, emax)
        self.eps_c = eps_c
        self.eps_f = np.repeat(np.repeat(np.repeat(eps_c, 2, 0), 2, 1), 2, 2)
        c = nc // 2
        self.src_c = tuple(int(v) for v in rng.integers(c - nc // 8, c + nc // 8, 3))
        self.t0, self.spread, self.waveform = t0, spread, waveform
        self.npml_f, self.npml_c = npml_fine, npml_fine // 2
        self.device = device

    def fine_grid(self) -> EMGrid3D:
        n = self.nf
        g = EMGrid3D((n, n, n), 1.0, "cpml", self.npml_f, device=self.device)
        g.add_material(DielectricRegion(mask=np.ones((n, n, n), bool), epsilon_r=self.eps_f))
        # a z dipole centred on the coarse Ez node (2I, 2J, 2K+1): two fine
        # Ez cells, half amplitude each; time in fine steps
        I, J, K = self.src_c
        for kk in (2 * K, 2 * K + 1):
            g.add_source(PointSource3D(0.0, (2 * I, 2 * J, kk), 0.5, "z", waveform=self.waveform,
                                       t0=2 * self.t0, spread=2 * self.spread))
        return g

    def coarse_grid(self) -> EMGrid3D:
        n = self.nc
        g = EMGrid3D((n, n, n), 2.0, "cpml", self.npml_c, device=self.device)
        g.add_material(DielectricRegion(mask=np.ones((n, n, n), bool), epsilon_r=self.eps_c))
        # same dipole moment per unit time: (dt_c / dt_f) (dx_f / dx_c)^3 = 1/4
        g.add_source(PointSource3D(0.0, self.src_c, 0.25, "z", waveform=self.waveform,
                                   t0=self.t0, spread=self.spread))
        return g

    def dt_fine(self) -> float:
        return COURANT * 1.0 / C0

    def dt_coarse(self) -> float:
        return COURANT * 2.0 / C0

    def fine_states(self, n_coarse_steps: int, every: int):
        """Run the fine grid 2*n_coarse_steps steps; return the restricted
        coarse states at coarse step counts 0, every, 2*every, ..., as lists
        of six (nz/2, ny/2, nx/2) tensors, and the fine solver."""
        g = self.fine_grid()
        nf = 2 * n_coarse_steps
        sol = g.build_solver(self.dt_fine(), nf + 2)
        shape = (self.nc,) * 3
        states = [[torch.zeros(shape, device=sol.device) for _ in range(6)]]
        done = 0  # fine steps run so far: the state is (E^done, H^{done-1/2})
        for m in range(every, n_coarse_steps + 1, every):
            sol.run(done, 2 * m - 1 - done)          # -> H^{2m-3/2}
            h_prev = restrict_h(sol.fields[3:])
            sol.run(2 * m - 1, 1)                    # -> E^{2m}, H^{2m-1/2}
            e = restrict_e(sol.fields[:3])
            h = restrict_h(sol.fields[3:])
            states.append([t.contiguous() for t in e] +
                          [(0.5 * (a + b)).contiguous() for a, b in zip(h_prev, h)])
            done = 2 * m
        return states, sol


def rel_err(a: list, b: list, comps=(0, 1, 2), margin: int = 0) -> float:
    """Relative L2 difference over the chosen components (interior only)."""
    num = den = 0.0
    for c in comps:
        x, y = a[c], b[c]
        if margin:
            x = x[margin:-margin, margin:-margin, margin:-margin]
            y = y[margin:-margin, margin:-margin, margin:-margin]
        num += float(((x - y).double() ** 2).sum())
        den += float((y.double() ** 2).sum())
    return math.sqrt(num / max(den, 1e-300))


def coarse_rollout(prob: ClosureProblem, net: ClosureNet | None, n_coarse: int, every: int,
                   init=None, gain: float = 1.0):
    """Coarse run of n_coarse steps with the correction every ``every`` steps
    (net=None: plain coarse FDTD).  Returns the final coarse solver."""
    sol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + 1)
    if init is not None:
        for f, v in zip(sol.fields, init):
            f.copy_(v)
    for m in range(0, n_coarse, every):
        sol.run(m, every)
        if net is not None:
            net.correct(sol, gain)
    return sol


# ---------------------------------------------------------------------------
# training and evaluation
# ---------------------------------------------------------------------------

def set_normalisation(net: ClosureNet, states: list) -> None:
    """Per-component RMS over a set of coarse states (6 fields each)."""
    sc = []
    for c in range(6):
        acc = sum(float((s[c].double() ** 2).mean()) for s in states) / len(states)
        sc.append(math.sqrt(acc) if acc > 0 else 1.0)
    net.scale = torch.tensor(sc, dtype=torch.float32, device=net.device)
    net.eps_shift, net.eps_scale = 3.5, 2.5


def _interior_crop(n: int, margin: int, crop: int, rng) -> int:
    lo, hi = margin, n - margin - crop
    return int(rng.integers(lo, hi + 1)) if hi > lo else max(0, (n - crop) // 2)


def train_closure(device, seeds, n_fine: int = 256, n_coarse: int = 400, every: int = 10,
                  epochs: int = 2, crop: int = 64, lr: float = 1e-3, net: ClosureNet | None = None,
                  log=None, seed: int = 0, **problem_kw):
    """Train on fine-run/coarse pairs generated by the FDTD solver.

    A sample: restricted fine state s_m at coarse step m, the coarse solver run
    ``every`` steps from it (psi restarted at zero), target s_{m+every} minus
    that state.  Loss: mean squared normalised error on a random interior
    crop (a 3-cell border of the crop, where zero padding differs from the
    full domain, is excluded)."""
    rng = np.random.default_rng(seed)
    net = net or ClosureNet(device, seed=seed)
    history = []
    t_start = time.time()
    for ep in range(epochs):
        for s_i, sd in enumerate(seeds):
            prob = ClosureProblem(sd, n_fine, device=device, **problem_kw)
            states, _ = prob.fine_states(n_coarse, every)
            if ep == 0 and s_i == 0:
                set_normalisation(net, states[1:])
            csol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + every + 1)
            nc = prob.nc
            margin = prob.npml_c + 2
            order = rng.permutation(len(states) - 1)
            for q in order:
                m = int(q) * every
                csol.reset()
                for f, v in zip(csol.fields, states[q]):
                    f.copy_(v)
                csol.run(m, every)
                x = net.pack_fields(csol).view(nc, nc, nc, CIN)
                tgt = torch.stack([(states[q + 1][c] - csol.fields[c]) / net.scale[c]
                                   for c in range(6)], dim=-1)
                cz = _interior_crop(nc, margin, crop, rng)
                cy = _interior_crop(nc, margin, crop, rng)
                cx = _interior_crop(nc, margin, crop, rng)
                sl = (slice(cz, cz + crop), slice(cy, cy + crop), slice(cx, cx + crop))
                xc = x[sl].contiguous().view(-1, CIN)
                tc = tgt[sl]
                y = net.forward(xc, (crop, crop, crop), train=True).view(crop, crop, crop, 8)
                err = torch.zeros_like(y)
                b = 3
                inner = (slice(b, crop - b),) * 3
                err[inner + (slice(0, 6),)] = y[inner + (slice(0, 6),)] - tc[inner]
                n_in = (crop - 2 * b) ** 3 * 6
                loss = float((err ** 2).sum()) / n_in
                net.adam_step(net.backward((2.0 / n_in) * err.view(-1, 8)), lr=lr)
                history.append(loss)
            if log:
                log(f"epoch {ep} seed {sd}: mean loss {np.mean(history[-len(order):]):.4e} "
                    f"({time.time() - t_start:.1f} s)")
            del states
    return net, history


def evaluate_closure(net: ClosureNet, seed: int, n_fine: int = 256, n_coarse: int = 400,
                     every: int = 10, gain: float = 1.0, **problem_kw):
    """Fine reference vs coarse with / without the closure on one seed.

    Returns a dict with the interior relative L2 errors of E (coarse and
    coarse + closure, against the restricted fine state at the final time)
    and the wall times of the fine run and of the corrected coarse run."""
    prob = ClosureProblem(seed, n_fine, device=net.device, **problem_kw)
    # fine run (timed on its own: the plain fine FDTD it replaces)
    g = prob.fine_grid()
    sol = g.build_solver(prob.dt_fine(), 2 * n_coarse + 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.run(0, 2 * n_coarse - 1)
    torch.cuda.synchronize()
    t_fine_main = time.perf_counter() - t0
    h_prev = restrict_h(sol.fields[3:])
    t0 = time.perf_counter()
    sol.run(2 * n_coarse - 1, 1)
    torch.cuda.synchronize()
    t_fine = t_fine_main + time.perf_counter() - t0
    ref = restrict_e(sol.fields[:3]) + [0.5 * (a + b) for a, b in
                                         zip(h_prev, restrict_h(sol.fields[3:]))]
    del sol
    out = {"seed": seed, "t_fine_s": t_fine}
    margin = prob.npml_c
    # reference states at every correction time (restricted fine states)
    states, fsol = prob.fine_states(n_coarse, every)
    del fsol
    for label, model in (("coarse", None), ("closure", net)):
        csol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + 1)
        if model is not None:
            model.pack()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for m in range(0, n_coarse, every):
            csol.run(m, every)
            if model is not None:
                model.correct(csol, gain)
        torch.cuda.synchronize()
        out[f"t_{label}_s"] = time.perf_counter() - t0
        out[f"err_{label}"] = rel_err(csol.fields, ref, comps=(0, 1, 2), margin=margin)
        # the same rollout again, untimed, compared at every correction time
        csol.reset()
        num = den = 0.0
        for q, m in enumerate(range(0, n_coarse, every)):
            csol.run(m, every)
            if model is not None:
                model.correct(csol, gain)
            a_, b_ = _err_terms(csol.fields, states[q + 1], margin)
            num += a_
            den += b_
        out[f"err_traj_{label}"] = math.sqrt(num / den) if den > 0 else 0.0
    out["speedup"] = out["t_fine_s"] / out["t_closure_s"]
    return out


def _err_terms(a: list, b: list, margin: int):
    """Squared-difference and squared-reference sums of E on the interior."""
    num = den = 0.0
    for c in (0, 1, 2):
        x, y = a[c], b[c]
        if margin:
            x = x[margin:-margin, margin:-margin, margin:-margin]
            y = y[margin:-margin, margin:-margin, margin:-margin]
        num += float(((x - y).double() ** 2).sum())
        den += float((y.double() ** 2).sum())
    return num, den
