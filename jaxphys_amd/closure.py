"""Learned dispersion closure: a coarse (2x) Yee grid plus a learned
correction every few steps that reproduces fine-grid fields.  No reference
counterpart (the reference has no neural-network code); the FDTD solver
itself generates the training data and stays the fallback and the judge of
the error.  The correction has three parts: in the 15^3 box around each
source, the near-source patch (``NearSourcePatch``: per-interval
corrections computed on a local fine/coarse pair of grids around the
source; or, with ``source_patch=False``, the linear source model below,
fitted to the solver's own fine runs); outside those boxes a float32
linear far-field operator (a 27-tap stencil of the fields, fitted by
least squares, which corrects the coarse grid's dispersion away from the
source; ``fit_far_field``); and a BF16 3D conv net, whose gain validation
seeds choose (``select_net_gain``; zero when it does not lower their error
by more than 1 %).  The net is trained through the corrected rollouts
themselves (``train_rollout``): the trajectory error of a whole corrected
coarse run is back-propagated through the coarse steps with the transposed
Yee step of the adjoint kernels and through the far-field operator's
transpose into every correction the net made.  (Trained instead on
one-interval residuals from restricted fine states, ``train_closure``, it
learns the transient of those restarts and validation switches it off.)

Network (``ClosureNet``): channels 7 -> 32 -> 32 -> 6 (E, H normalised and
eps_r in; a field correction out), 3x3x3 zero-padded convolutions, GELU
(erf to 1.5e-7) after the first two layers.  The eighth input channel (the padding of
the 7 to the kernels' multiple of 8) optionally carries the forcing: the E
increment the coarse grid's point sources injected during the interval just
run, at their cells.  A pulse whose increments do not sum to zero leaves
charge behind, and the fine- and coarse-grid near fields of that charge
differ by a fixed local pattern times the injected amount, which the net can
only place where it knows a source acted.  That part of the correction is
linear in the forcing and as large as the fields themselves next to the
source, far outside the range the bf16 GELU layers resolve, so it is left
to the near-source patch, or to a separate float32 linear source model: in
the 15x15x15 box around each
source cell, correction = K0 f0 + K1 f1 + K2 f0 (eps_s - 1), with f0 the
interval's injected E increment, f1 its first time moment and eps_s the
permittivity at the source cell (the sources add to E itself, so in a
uniform medium the static field they leave per increment does not depend
on eps; the third term lets the fit learn how a dielectric at the source
changes it); K0..K2 (``ks``) are fitted by least squares on the
teacher-forced training intervals (a linear regression, solved exactly,
before the net is trained).  The net
is trained on, and corrects, the cells outside those boxes.  The field
channels enter as asinh(field / RMS), so the near-source values (thousands
of RMS) stay in the range of the rest.  The net's response to the zero
state (eps_r only) is subtracted, so a zero field gets no correction.  Corrections are applied only
``margin`` cells or more inside the faces: the absorbing layers, where the
net is never trained, are left to the solver.  Weights and activations are bfloat16 with
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

import dataclasses
import math
import time

import numpy as np
import torch
import torch.nn.functional as F

from jaxphys_amd import _conv
from jaxphys_amd.em.constants import C0
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D
from jaxphys_amd.exceptions import ConfigurationError

CIN, HID, COUT = 8, 32, 6          # padded input channels (7 real), hidden, output
KP1, KP2 = _conv.kpad_of(CIN), _conv.kpad_of(HID)
OUT_STRIDE = 8                      # float32 output channel stride
STENCIL = 15                        # side of the linear source-model box
NFEAT = 3                           # source features: sum, first moment, sum * (eps - 1)
_OFFS = [(dx, dy, dz) for dz in range(-(STENCIL // 2), STENCIL // 2 + 1)
         for dy in range(-(STENCIL // 2), STENCIL // 2 + 1)
         for dx in range(-(STENCIL // 2), STENCIL // 2 + 1)]
# the linear far-field operator: 27 taps x 6 fields x (plain, eps - 1) -> 6
FAR_TAPS = 27
FAR_FEAT = 2 * 6 * FAR_TAPS
FAR_RIDGE = 3e-3        # ridge, relative to the mean diagonal of the normal matrix
FAR_GAIN = 0.75         # share of the fitted correction applied per interval
FAR_SAMPLES = 60000     # cells sampled per training interval
FAR_GROWTH = 1.1        # largest accepted energy ratio of the stability check
PATCH_HALF = 16         # half side (coarse cells) of the near-source patch's local grids
PATCH_TAIL = 2          # patch intervals after the last one a source injects into


def _bf16(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16)


class ClosureNet:
    """3-layer 3D conv net on the BF16 MFMA kernels (see module docstring)."""

    def __init__(self, device, seed: int = 0, out_scale: float = 1e-2, forcing: bool = True):
        self.device = torch.device(device)
        self.forcing = bool(forcing)
        self.force_scale = 1.0
        self.pack_flags = _conv.PACK_ASINH
        self.n_in = 8 if self.forcing else 7
        g = torch.Generator(device="cpu").manual_seed(seed)

        def init(cout, cin_real, cin_pad, scale):
            w = torch.zeros(cout, 27, cin_pad)
            w[:, :, :cin_real] = torch.randn(cout, 27, cin_real, generator=g) * scale
            return w.to(self.device)

        self.params = {
            "w1": init(HID, self.n_in, CIN, math.sqrt(2.0 / (27 * self.n_in))),
            "b1": torch.zeros(HID, device=self.device),
            "w2": init(HID, HID, HID, math.sqrt(2.0 / (27 * HID))),
            "b2": torch.zeros(HID, device=self.device),
            "w3": init(COUT, HID, HID, out_scale * math.sqrt(1.0 / (27 * HID))),
            "b3": torch.zeros(COUT, device=self.device),
        }
        # the linear source model (physical units; least squares, not Adam)
        self.ks = torch.zeros(NFEAT, COUT, STENCIL ** 3, device=self.device)
        # the linear far-field operator (float32 stencil, least squares; the
        # applied gain is folded into the weights); off until fitted
        self.far_w = torch.zeros(FAR_TAPS * 6 * 2 * 6, device=self.device)
        self.far_gain = 0.0
        # scale of the net's correction (select_net_gain picks it on
        # validation seeds after training)
        self.net_gain = 1.0
        # whether the net also corrects inside the source boxes (on top of
        # the source model) or only outside them
        self.net_in_box = False
        # the source boxes' correction: the near-source patch computed on
        # local fine/coarse grids (True, attach_patch), or the linear source
        # model ks (False)
        self.source_patch = True
        self.patch_half = PATCH_HALF
        self.patch_tail = PATCH_TAIL   # -1: the patch covers every interval
        self.margin = 0
        self.adam = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self.params.items()}
        self.step_count = 0
        # normalisation (set from data): fields / scale, (eps - shift) / escale
        self.scale = torch.ones(6, device=self.device)
        self.eps_shift, self.eps_scale = 1.0, 1.0
        self._packed = None
        self._ws = {}
        self._saved = {}
        self.version = 0  # bumped by every weight update

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
    def forward(self, x: torch.Tensor, shape_xyz, train: bool = False,
                tag: str = "") -> torch.Tensor:
        """``x``: bf16 (npos, 8) channels last over (nz, ny, nx); returns the
        float32 network output (npos, 8) (channels 0..5 real).  ``tag``
        names the set of buffers (and saved activations for ``backward``)
        used, so that two forward passes can be kept at once."""
        if self._packed is None:
            self.pack()
        nx, ny, nz = shape_xyz
        n = nx * ny * nz
        P, W = self.params, self._packed
        d1 = _conv.desc(shape_xyz, CIN, HID, 32, HID)
        d2 = _conv.desc(shape_xyz, HID, HID, 32, HID)
        d3 = _conv.desc(shape_xyz, HID, COUT, 16, OUT_STRIDE)
        a1 = self._buf("a1" + tag, (n, HID), torch.bfloat16)
        a2 = self._buf("a2" + tag, (n, HID), torch.bfloat16)
        z1 = self._buf("z1" + tag, (n, HID), torch.bfloat16) if train else None
        z2 = self._buf("z2" + tag, (n, HID), torch.bfloat16) if train else None
        y = self._buf("y" + tag, (n, OUT_STRIDE), torch.float32)
        _conv.fwd(d1, x, W["w1"], P["b1"], _conv.EPI_GELU, a1, z_out=z1)
        _conv.fwd(d2, a1, W["w2"], P["b2"], _conv.EPI_GELU, a2, z_out=z2)
        _conv.fwd(d3, a2, W["w3"], P["b3"], _conv.EPI_LINEAR_F32, y)
        if train:
            self._saved[tag] = (x, a1, a2, z1, z2, shape_xyz)
        return y

    def backward(self, dy: torch.Tensor, tag: str = "") -> dict:
        """``dy``: float32 (npos, 8) gradient of the loss w.r.t. the output
        of the forward pass ``tag``.  Returns float32 gradients shaped like
        the network's weights and biases."""
        x, a1, a2, z1, z2, shape_xyz = self._saved[tag]
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
        grads["w1"][:, :, self.n_in:] = 0
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
        self.version += 1

    # -- fields <-> network tensors --------------------------------------
    def pack_fields(self, solver, m: int | None = None, every: int | None = None,
                    fields=None) -> torch.Tensor:
        """bf16 network input of the solver's state (or of ``fields``, six
        tensors on the solver's grid).  With the forcing channel, ``m`` and
        ``every`` name the interval just run (steps m .. m+every-1)."""
        n = solver.nx * solver.ny * solver.nz
        x = self._buf("xin", (n, CIN), torch.bfloat16)
        _conv.pack_fields((solver.nx, solver.ny, solver.nz),
                          solver.fields if fields is None else fields, solver.eps_r,
                          self.scale, self.eps_shift, self.eps_scale, x, self.pack_flags)
        if self.forcing:
            if m is None or every is None:
                raise ValueError("a closure net with the forcing channel needs the interval")
            fc = interval_forcing(solver, every)
            if fc is not None:
                q = m // every
                if m % every or q >= fc.feats.shape[0]:
                    raise ValueError(f"interval start {m} is not a multiple of {every} "
                                     "inside the source table")
                x[fc.flat, 7] = (fc.feats[q, :, 0] * (1.0 / self.force_scale)).to(torch.bfloat16)
        return x

    def add_far_field(self, y: torch.Tensor, solver, fields=None) -> None:
        """y[.., :6] += the linear far-field operator of the solver's state
        (or of ``fields``), normalised units, gain included."""
        _conv.stencil_f32((solver.nx, solver.ny, solver.nz),
                          solver.fields if fields is None else fields, solver.eps_r,
                          self.scale, self.far_w, y, OUT_STRIDE)

    def correct(self, solver, gain: float = 1.0, m: int | None = None,
                every: int | None = None, use_net: bool = True, use_far: bool = True) -> None:
        """fields += net(fields, eps[, forcing of steps m .. m+every-1])
        plus the linear far-field operator (denormalised) outside the source
        boxes, plus the near-source patch (or the linear source model)
        inside them.  ``use_net=False`` / ``use_far=False`` leave the net /
        the far-field operator out (for measuring what each part adds).

        The net runs on its box of the grid only (``conv3d_pack_fields_box``
        packs that box), the far-field operator writes its own buffer, and
        one kernel (``conv3d_apply_closure``) forms net - net(0) + far field,
        clears the source boxes and adds the result to the fields: the same
        float operations as ``_correct_unfused``, without the full-grid
        correction tensor and the passes that build it."""
        if self.forcing and not self.source_patch:
            self._correct_unfused(solver, gain, m, every, use_net, use_far)
            return
        shape = (solver.nx, solver.ny, solver.nz)
        yb = y0b = box6 = None
        if use_net and self.net_gain > 0:
            yb, y0b, box6 = self._net_box(solver, m, every)
        s = None
        if use_far and self.far_gain > 0:
            # on the corrected cells only (``margin`` or more from the faces)
            mg = self.margin
            inner = (mg, mg, mg) + tuple(max(1, n - 2 * mg) for n in shape)
            s = self._buf("far_s", (solver.nx * solver.ny * solver.nz, OUT_STRIDE), torch.float32)
            _conv.stencil_f32_set(shape, solver.fields, solver.eps_r, self.scale, self.far_w, s,
                                  OUT_STRIDE, inner if min(n - 2 * mg for n in shape) > 0 else None)
        mask, mask_all, patch = None, True, None
        if self.forcing:
            box = self._box_terms(solver, shape, (0, 0, 0), every)[0]
            if not self.net_in_box or s is not None:
                mask = self._box_mask(solver, shape, every)
                mask_all = not self.net_in_box
            patch = getattr(solver, "_closure_patch", None)
            if patch is None and box.numel() > 0:
                raise RuntimeError("this closure corrects the source boxes with the "
                                   "near-source patch: call closure.attach_patch(solver, "
                                   "grid, n_steps, every) first")
        _conv.apply_closure(shape, solver.fields, s, OUT_STRIDE, yb, y0b, OUT_STRIDE, box6,
                            mask, mask_all, self.net_gain if yb is not None else 1.0,
                            self.scale, gain, self.margin)
        if patch is not None:
            patch.apply(solver, m // every, gain)

    def _box_mask(self, solver, shape_xyz, every) -> torch.Tensor:
        """uint8 per cell: 1 in the source boxes (cached on the solver)."""
        key = (tuple(shape_xyz), every)
        cache = getattr(solver, "_closure_box_mask", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        nx, ny, nz = shape_xyz
        mask = torch.zeros(nx * ny * nz, dtype=torch.uint8, device=self.device)
        mask[self._box_terms(solver, shape_xyz, (0, 0, 0), every)[0]] = 1
        solver._closure_box_mask = (key, mask)
        return mask

    def _net_box(self, solver, m: int | None, every: int | None):
        """The net's output and zero-state response on its box of the grid
        (float32 (npos, 8) each, box order) and the box (x0, y0, z0, bnx,
        bny, bnz)."""
        sl = self._crop(solver)
        box6 = (sl[2].start, sl[1].start, sl[0].start, sl[2].stop - sl[2].start,
                sl[1].stop - sl[1].start, sl[0].stop - sl[0].start)
        shape_c = box6[3:]
        nb = shape_c[0] * shape_c[1] * shape_c[2]
        xc = self._buf("xc", (nb, CIN), torch.bfloat16)
        _conv.pack_fields_box((solver.nx, solver.ny, solver.nz), solver.fields, solver.eps_r,
                              self.scale, self.eps_shift, self.eps_scale, box6, xc,
                              self.pack_flags)
        if self.forcing:
            if m is None or every is None:
                raise ValueError("a closure net with the forcing channel needs the interval")
            fc = interval_forcing(solver, every)
            if fc is not None:
                q = m // every
                if m % every or q >= fc.feats.shape[0]:
                    raise ValueError(f"interval start {m} is not a multiple of {every} "
                                     "inside the source table")
                cflat, vals = self._forcing_in_box(solver, fc, box6)
                xc[cflat, 7] = vals[q]
        yc = self.forward(xc, shape_c)
        y0 = self.zero_response(solver).view(nb, OUT_STRIDE)
        return yc, y0, box6

    def _forcing_in_box(self, solver, fc, box6):
        """(box-order flat indices of the forcing cells inside the box, the
        forcing channel's bf16 values there for every interval: (intervals,
        cells)), cached on the solver: one scatter per correction."""
        key = (id(fc), tuple(box6), self.force_scale)
        cache = getattr(solver, "_closure_fc_box", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        nx, ny = solver.nx, solver.ny
        x0, y0, z0, bx, by, bz = box6
        f = fc.flat
        x, y, z = f % nx - x0, (f // nx) % ny - y0, f // (nx * ny) - z0
        inside = (x >= 0) & (x < bx) & (y >= 0) & (y < by) & (z >= 0) & (z < bz)
        sel = torch.nonzero(inside).flatten()
        cflat = ((z * by + y) * bx + x)[sel]
        vals = (fc.feats[:, sel, 0] * (1.0 / self.force_scale)).to(torch.bfloat16)
        solver._closure_fc_box = (key, (cflat, vals))
        return cflat, vals

    def _correct_unfused(self, solver, gain: float = 1.0, m: int | None = None,
                         every: int | None = None, use_net: bool = True,
                         use_far: bool = True) -> None:
        """``correct`` built from the full-grid correction tensor: the net's
        correction, the far-field operator added into it, the source boxes
        cleared or given the source model, ``conv3d_apply_correction``."""
        if use_net and self.net_gain > 0:
            y = self.net_correction(solver, m, every)
            if self.net_gain != 1.0:
                y.mul_(self.net_gain)
        else:
            y = self._buf("ycorr", (solver.nx * solver.ny * solver.nz, OUT_STRIDE), torch.float32)
            y.zero_()
        shape = (solver.nx, solver.ny, solver.nz)
        box = self._box_terms(solver, shape, (0, 0, 0), every)[0] if self.forcing else None
        if use_far and self.far_gain > 0:
            if self.net_in_box and box is not None:
                # the far-field operator stays out of the source boxes; the
                # net's correction is kept there
                far = self._buf("far_c", tuple(y.shape), torch.float32)
                far.zero_()
                self.add_far_field(far, solver)
                far.view(-1, OUT_STRIDE)[box] = 0.0
                y.add_(far)
            else:
                self.add_far_field(y, solver)
        patch = None
        if self.forcing:
            if not self.net_in_box:
                y.view(-1, OUT_STRIDE)[box] = 0.0
            if self.source_patch:
                patch = getattr(solver, "_closure_patch", None)
                if patch is None and box.numel() > 0:
                    raise RuntimeError("this closure corrects the source boxes with the "
                                       "near-source patch: call closure.attach_patch(solver, "
                                       "grid, n_steps, every) first")
            else:
                self.add_source_model(y, solver, shape, (0, 0, 0), m, every)
        _conv.apply_correction((solver.nx, solver.ny, solver.nz), solver.fields, y, OUT_STRIDE,
                               self.scale, gain, self.margin)
        if patch is not None:
            patch.apply(solver, m // every, gain)

    # the net's receptive field: three 3x3x3 layers
    HALO = 3

    def _crop(self, solver):
        """(z, y, x) slices of the box the net is evaluated on: the cells the
        correction is applied to (``margin`` or more inside the faces) and
        the net's receptive field around them, clamped to the grid."""
        lo = max(0, self.margin - self.HALO)
        return tuple(slice(lo, max(lo, n - lo)) for n in (solver.nz, solver.ny, solver.nx))

    def net_correction(self, solver, m: int | None = None, every: int | None = None):
        """net(state) - net(zero state) on the whole grid, float32 (npos, 8)
        with channels 0-5 real, zero outside the cells ``margin`` or more
        inside the faces (which is all the correction uses).  The net runs on the box of those cells
        plus its receptive field only: every output inside that region then
        depends on inputs inside the box alone, and the convolutions compute
        each output with the same operations whatever the grid's size, so
        the values are those of the net on the whole grid, bit for bit."""
        x = self.pack_fields(solver, m, every)
        sl = self._crop(solver)
        nz, ny, nx = solver.nz, solver.ny, solver.nx
        shape_c = (sl[2].stop - sl[2].start, sl[1].stop - sl[1].start, sl[0].stop - sl[0].start)
        xc = self._buf("xc", (shape_c[0] * shape_c[1] * shape_c[2], CIN), torch.bfloat16)
        xc.view(shape_c[2], shape_c[1], shape_c[0], CIN).copy_(x.view(nz, ny, nx, CIN)[sl])
        yc = self.forward(xc, shape_c).view(shape_c[2], shape_c[1], shape_c[0], OUT_STRIDE)
        y0 = self.zero_response(solver)
        y = self._buf("ycorr", (nx * ny * nz, OUT_STRIDE), torch.float32)
        y.zero_()
        mg = self.margin
        inner = tuple(slice(mg, max(mg, n - mg)) for n in (nz, ny, nx))
        rel = tuple(slice(a.start - c.start, a.stop - c.start) for a, c in zip(inner, sl))
        ch = (slice(0, COUT),)  # the output layer writes the real channels only
        y.view(nz, ny, nx, OUT_STRIDE)[inner + ch] = yc[rel + ch] - y0[rel + ch]
        return y

    def pack_zero(self, solver) -> torch.Tensor:
        """Network input of the zero state of the solver's grid (eps_r only):
        a bf16 (npos, 8) tensor of its own."""
        shape = (solver.nz, solver.ny, solver.nx)
        zeros = [torch.zeros(shape, dtype=torch.float32, device=self.device)] * 6
        x0 = torch.empty(solver.nx * solver.ny * solver.nz, CIN, dtype=torch.bfloat16,
                         device=self.device)
        _conv.pack_fields((solver.nx, solver.ny, solver.nz), zeros, solver.eps_r, self.scale,
                          self.eps_shift, self.eps_scale, x0, self.pack_flags)
        return x0

    def zero_response(self, solver) -> torch.Tensor:
        """The network's output for zero fields and zero forcing on the box
        ``net_correction`` evaluates it on ((z, y, x, 8) float32; it depends
        on eps_r only), cached on the solver until the weights change.
        Every correction subtracts it, so the correction of a zero state is
        exactly zero: no constant is added to the fields at every
        correction."""
        key = (id(self), self.version, self.margin)
        cache = getattr(solver, "_closure_y0", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        sl = self._crop(solver)
        shape_c = (sl[2].stop - sl[2].start, sl[1].stop - sl[1].start, sl[0].stop - sl[0].start)
        x0 = self.pack_zero(solver).view(solver.nz, solver.ny, solver.nx, CIN)[sl].contiguous()
        y0 = self.forward(x0.view(-1, CIN), shape_c, tag="0").clone()
        y0 = y0.view(shape_c[2], shape_c[1], shape_c[0], OUT_STRIDE)
        solver._closure_y0 = (key, y0)
        return y0

    def _box_terms(self, solver, shape_xyz, origin, every):
        """(flat indices into the box of ``shape_xyz`` at ``origin``, slot in
        the source-model box, source-cell number) of the cells of each
        source's box that lie inside it (device tensors, cached)."""
        key = (tuple(shape_xyz), tuple(origin), every)
        cache = getattr(solver, "_closure_boxes", {})
        if key in cache:
            return cache[key]
        fc = interval_forcing(solver, every)
        nx, ny, nz = shape_xyz
        idx, slot, which = [], [], []
        for c_i, (i, j, k) in enumerate(fc.xyz if fc is not None else []):
            for o, (dx, dy, dz) in enumerate(_OFFS):
                x_, y_, z_ = i + dx - origin[0], j + dy - origin[1], k + dz - origin[2]
                if 0 <= x_ < nx and 0 <= y_ < ny and 0 <= z_ < nz:
                    idx.append((z_ * ny + y_) * nx + x_)
                    slot.append(o)
                    which.append(c_i)
        dev = self.device
        out = (torch.tensor(idx, dtype=torch.long, device=dev),
               torch.tensor(slot, dtype=torch.long, device=dev),
               torch.tensor(which, dtype=torch.long, device=dev))
        if len(cache) > 64:
            cache = {}
        cache[key] = out
        solver._closure_boxes = cache
        return out

    def add_source_model(self, y, solver, shape_xyz, origin, m, every) -> None:
        """y[.., :6] += (K0 f0 + K1 f1) / scale in the box around each source
        cell (box-relative indices)."""
        fc = interval_forcing(solver, every)
        if fc is None:
            return
        idx, slot, which = self._box_terms(solver, shape_xyz, origin, every)
        if idx.numel() == 0:
            return
        phi = source_features(fc, m // every)                    # (cells, NFEAT)
        k = self.ks[:, :, slot]                                 # (NFEAT, 6, E)
        contrib = (k * phi[which].T[:, None, :]).sum(0).T / self.scale[None, :]
        y.view(-1, OUT_STRIDE)[:, :COUT].index_add_(0, idx, contrib)

    def state_dict(self):
        return {"params": {k: v.detach().cpu() for k, v in self.params.items()},
                "scale": self.scale.cpu(), "eps_shift": self.eps_shift,
                "eps_scale": self.eps_scale, "forcing": self.forcing,
                "force_scale": self.force_scale, "margin": self.margin,
                "ks": self.ks.cpu(), "far_w": self.far_w.cpu(), "far_gain": self.far_gain,
                "net_gain": self.net_gain}

    # scalar settings stored beside the tensors of a saved closure
    _META = ("eps_shift", "eps_scale", "force_scale", "margin", "far_gain", "net_gain",
             "net_in_box", "source_patch", "patch_half", "patch_tail")

    def save(self, path: str, info: dict | None = None) -> None:
        """Write the trained closure (weights, normalisation, the two linear
        parts and the gains) as a safetensors file: tensors only, plus the
        scalar settings and ``info`` as JSON metadata (nothing executable)."""
        import json

        from safetensors.torch import save_file

        tensors = {f"params.{k}": v.detach().float().cpu().contiguous()
                   for k, v in self.params.items()}
        tensors.update({"scale": self.scale.float().cpu().contiguous(),
                        "ks": self.ks.float().cpu().contiguous(),
                        "far_w": self.far_w.float().cpu().contiguous()})
        meta = {k: float(getattr(self, k)) for k in self._META}
        meta["forcing"] = bool(self.forcing)
        save_file(tensors, path, metadata={"closure": json.dumps(meta),
                                           "info": json.dumps(info or {})})

    @classmethod
    def load(cls, path: str, device) -> "ClosureNet":
        """A closure saved by ``save``, on ``device``, ready to ``correct``."""
        import json

        from safetensors import safe_open

        with safe_open(path, framework="pt") as fh:
            meta = json.loads(fh.metadata()["closure"])
            tensors = {k: fh.get_tensor(k) for k in fh.keys()}
        net = cls(device, forcing=bool(meta["forcing"]))
        for k in net.params:
            t = tensors[f"params.{k}"]
            if t.shape != net.params[k].shape:
                raise ValueError(f"saved closure: {k} has shape {tuple(t.shape)}, "
                                 f"expected {tuple(net.params[k].shape)}")
            net.params[k] = t.to(net.device)
        net.scale = tensors["scale"].to(net.device)
        net.ks = tensors["ks"].to(net.device)
        net.far_w = tensors["far_w"].to(net.device)
        for k in cls._META:
            if k not in meta:
                continue
            v = meta[k]
            setattr(net, k, int(v) if k in ("margin", "patch_half", "patch_tail") else
                    (bool(v) if k in ("net_in_box", "source_patch") else float(v)))
        net.adam = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in net.params.items()}
        net.pack()
        net.version += 1
        return net

    @staticmethod
    def info(path: str) -> dict:
        """The ``info`` metadata a closure file was saved with."""
        import json

        from safetensors import safe_open

        with safe_open(path, framework="pt") as fh:
            return json.loads(fh.metadata().get("info", "{}"))


class _Forcing:
    """Per-interval source features of one solver (see interval_forcing)."""

    def __init__(self, flat, feats, xyz, eps):
        self.flat, self.feats, self.xyz, self.eps = flat, feats, xyz, eps


def interval_forcing(solver, every: int):
    """The point sources' E increments over each interval of ``every`` steps
    (intervals start at multiples of ``every``), per distinct source cell:
    ``flat`` cell indices, ``feats`` float32 (n_intervals, n_cells, 2) =
    (sum of the increments, their first moment sum((n - n_mid)/every * s_n)),
    the cells' (i, j, k) and eps_r there; None without sources.  Cached on
    the solver."""
    cache = getattr(solver, "_closure_forcing", None)
    if cache is not None and cache[0] == every:
        return cache[1]
    tab = np.asarray(getattr(solver, "source_table", np.zeros((0, 0))), np.float64)
    if tab.size == 0:
        out = None
    else:
        nq = -(-tab.shape[0] // every)
        pad = np.zeros((nq * every, tab.shape[1]))
        pad[:tab.shape[0]] = tab
        per = pad.reshape(nq, every, -1)
        w = (np.arange(every) - (every - 1) / 2.0) / every
        f0 = per.sum(axis=1)
        f1 = (per * w[None, :, None]).sum(axis=1)
        cells: dict = {}
        for s_i, (i, j, k) in enumerate(solver.source_positions):
            cells.setdefault((k * solver.ny + j) * solver.nx + i, []).append(s_i)
        flat = sorted(cells)
        feats = np.stack([np.stack([f[:, cells[c]].sum(axis=1) for c in flat], axis=1)
                          for f in (f0, f1)], axis=2)
        xyz = [(c % solver.nx, (c // solver.nx) % solver.ny, c // (solver.nx * solver.ny))
               for c in flat]
        dev = solver.device
        flat_t = torch.tensor(flat, dtype=torch.long, device=dev)
        eps = solver.eps_r.reshape(-1)[flat_t].to(torch.float32)
        out = _Forcing(flat_t, torch.tensor(feats, dtype=torch.float32, device=dev), xyz, eps)
    solver._closure_forcing = (every, out)
    return out


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
        sol = g.build_solver(self.dt_fine(), 2 * n_coarse_steps + 2)
        shape = (self.nc,) * 3
        states = [[torch.zeros(shape, device=sol.device) for _ in range(6)]]
        states += list(restricted_states(sol, n_coarse_steps, every))
        return states, sol


def restricted_states(sol, n_coarse_steps: int, every: int):
    """Advance a fine solver (from step 0) and yield the restricted coarse
    state at coarse step counts every, 2*every, ... <= n_coarse_steps: six
    tensors, E^{2m} restricted and the mean of the restricted H^{2m-3/2} and
    H^{2m-1/2} (the module docstring's staggering in time)."""
    done = 0  # fine steps run so far: the state is (E^done, H^{done-1/2})
    for m in range(every, n_coarse_steps + 1, every):
        sol.run(done, 2 * m - 1 - done)          # -> H^{2m-3/2}
        h_prev = restrict_h(sol.fields[3:])
        sol.run(2 * m - 1, 1)                    # -> E^{2m}, H^{2m-1/2}
        e = restrict_e(sol.fields[:3])
        h = restrict_h(sol.fields[3:])
        yield ([t.contiguous() for t in e] +
               [(0.5 * (a + b)).contiguous() for a, b in zip(h_prev, h)])
        done = 2 * m


# ---------------------------------------------------------------------------
# near-source patch
# ---------------------------------------------------------------------------

_AXIS = {"x": 0, "y": 1, "z": 2}


def _fine_twin(s: PointSource3D, pos_c) -> list:
    """The fine-grid (dx/2, dt/2) sources that inject what the coarse source
    ``s`` injects at coarse position ``pos_c``: two fine cells along the
    driven component's axis (the coarse node sits between them), twice the
    amplitude each (the same dipole moment per unit time: (dt_c / dt_f)
    (dx_f / dx_c)^3 = 1/4 of the two cells' sum), the pulse's centre, width
    and carrier wavelength in fine steps / cells."""
    wf = s.waveform
    if wf is not None and not isinstance(wf, str):
        raise ConfigurationError("the near-source patch needs a named waveform "
                                 "(an array of per-step values has no fine-step samples)")
    # fdtd3d.py:239-244: "z" -> Ez, "x" -> Ex, anything else -> Ey
    a = _AXIS.get(s.polarization, 1)
    out = []
    for d in (0, 1):
        p = [2 * int(v) for v in pos_c]
        p[a] += d
        out.append(dataclasses.replace(
            s, position=tuple(p), amplitude=2.0 * s.amplitude, t0=2.0 * s.t0,
            spread=2.0 * s.spread,
            wavelength=None if s.wavelength is None else 2.0 * s.wavelength))
    return out


def _refine(t: torch.Tensor) -> torch.Tensor:
    """A cell-centred (x, y, z) array on the 2x-refined grid: each value
    repeated 2x2x2."""
    return t.repeat_interleave(2, 0).repeat_interleave(2, 1).repeat_interleave(2, 2)


class NearSourcePatch:
    """Corrections of the (2R+1)^3 box around each point-source cell of a
    coarse grid, one per correction interval, computed on a local
    fine/coarse pair of grids instead of fitted.

    Next to a point source the coarse grid misrepresents the field by as
    much as the field itself (the source's own near field and the static
    remnant of a pulse whose increments do not sum to zero), and how it does
    depends on the permittivity around the source.  For each source cell a
    local coarse grid of side 2 ``half`` around it and its 2x-refined twin
    (the same permittivity repeated 2x2x2, the same source injected as
    ``_fine_twin`` does, CPML of the coarse grid's thickness and twice it)
    run side by side: after every interval the local coarse state in the
    box is replaced by the restricted fine state, and the difference is the
    interval's correction.  The global coarse run adds the same corrections
    to the same cells (``apply``), so in the box it follows the local fine
    run, which for the fields the source itself makes is the fine run.
    Corrections of separate sources add (the field equations are linear);
    each source's local grids hold only the sources at its own cell.  With
    ``tail`` the local runs stop ``tail`` intervals after the last interval
    into which a source injects anything (1e-6 of its peak); later
    intervals get no correction in the box, which the coarse grid then
    carries on its own (on the 256^3 / 128^3 problem tail 2 is as accurate
    as correcting every interval and costs a third of it)."""

    def __init__(self, grid: EMGrid3D, dt: float, n_coarse: int, every: int,
                 half: int = PATCH_HALF, radius: int = STENCIL // 2, device=None,
                 tail: int | None = None, overlap: bool = True):
        from jaxphys_amd import _native
        if grid._boundary != "cpml":
            raise ConfigurationError("the near-source patch needs a cpml grid")
        dev = torch.device(device or grid._device or _native.require_device())
        npml = int(grid._pml_layers)
        n_loc = 2 * half
        if n_loc - 2 * npml < 2 * radius + 1:
            raise ConfigurationError(
                f"near-source patch: local grid {n_loc} with {npml}-cell CPML leaves no room "
                f"for the {2 * radius + 1}-cell box")
        eps, sigma = grid._build_media(dev)           # internal (z, y, x)
        size = grid.size
        cells: dict = {}
        for s in grid._sources:
            cells.setdefault(tuple(int(v) for v in s.position), []).append(s)
        self.radius, self.half, self.every = radius, half, every
        self.boxes = []      # (z, y, x) slices of the global grid per source cell
        self.corr = []       # (n_intervals, 6, 2R+1, 2R+1, 2R+1) per source cell
        cpml = dict(pml_R0=grid._cpml["R0"], pml_grading=grid._cpml["grading"],
                    pml_kappa_max=grid._cpml["kappa_max"], pml_alpha_max=grid._cpml["alpha_max"])
        for pos, srcs in cells.items():
            o = [min(max(p - half, 0), n - n_loc) for p, n in zip(pos, size)]
            if min(o) < 0:
                raise ConfigurationError(f"near-source patch: grid {size} is smaller than "
                                         f"its {n_loc}-cell local grid")
            loc = [p - a for p, a in zip(pos, o)]
            if any(v - radius < npml or v + radius >= n_loc - npml for v in loc):
                raise ConfigurationError(
                    f"near-source patch: source at {pos} is too close to the grid's faces")
            sl = (slice(o[2], o[2] + n_loc), slice(o[1], o[1] + n_loc), slice(o[0], o[0] + n_loc))
            eps_c = eps[sl].permute(2, 1, 0).contiguous()      # (x, y, z)
            sig_c = None if sigma is None else sigma[sl].permute(2, 1, 0).contiguous()
            gc = EMGrid3D((n_loc,) * 3, grid._dx, "cpml", npml, device=dev, **cpml)
            gf = EMGrid3D((2 * n_loc,) * 3, 0.5 * grid._dx, "cpml", 2 * npml, device=dev, **cpml)
            ones = np.ones((n_loc,) * 3, bool)
            gc.add_material(DielectricRegion(mask=ones, epsilon_r=eps_c, sigma=sig_c))
            gf.add_material(DielectricRegion(
                mask=np.ones((2 * n_loc,) * 3, bool), epsilon_r=_refine(eps_c),
                sigma=None if sig_c is None else _refine(sig_c)))
            for s in srcs:
                gc.add_source(dataclasses.replace(s, position=tuple(loc)))
                for t in _fine_twin(s, loc):
                    gf.add_source(t)
            csol = gc.build_solver(dt, n_coarse + every + 1)
            n_run = n_coarse
            if tail is not None:
                # intervals up to the last one the source injects into, then
                # ``tail`` more
                tab = np.abs(np.asarray(csol.source_table, np.float64)[:n_coarse])
                live = np.nonzero(tab.max(axis=1) > 1e-6 * max(float(tab.max()), 1e-300))[0]
                last = int(live[-1]) // every if live.size else -1
                n_run = min(n_coarse, (last + 1 + tail) * every)
            fsol = gf.build_solver(0.5 * dt, 2 * n_run + 2)
            side = 2 * radius + 1
            org = tuple(v - radius for v in loc)                # box origin (x, y, z)
            fshape, cshape = (2 * n_loc,) * 3, (n_loc,) * 3
            n_int = n_run // every
            corr = torch.empty((n_int, 6, side, side, side), device=dev)
            hprev = torch.empty((3, side, side, side), device=dev)
            # the fine run advanced interval by interval: H^{2m-3/2} restricted
            # into hprev, then one more step to (E^{2m}, H^{2m-1/2}); the
            # kernels restrict, average and correct the coarse box in place
            # (the steps of restricted_states, on the box only).  The local
            # coarse run of an interval goes on a second stream beside the
            # fine run: both grids are too small to fill the GPU, and they
            # meet only in patch_correct (which then waits for the coarse
            # run, and which the next coarse run waits for: it corrects the
            # coarse box in place)
            # (overlap=False: both on the current stream, for comparison)
            main = torch.cuda.current_stream(dev)
            aside = _side_stream(dev) if overlap else main
            done = 0
            for q in range(n_int):
                m = (q + 1) * every
                ev = torch.cuda.Event()
                ev.record(main)
                aside.wait_event(ev)
                csol.run(q * every, every, stream=aside)
                fsol.run(done, 2 * m - 1 - done)
                _conv.patch_restrict_h(fsol.fields[3:], fshape, org, side, hprev)
                fsol.run(2 * m - 1, 1)
                ev = torch.cuda.Event()
                ev.record(aside)
                main.wait_event(ev)
                _conv.patch_correct(fsol.fields[:3], fsol.fields[3:], hprev, fshape,
                                    csol.fields, cshape, org, side, corr[q])
                done = 2 * m
            del csol, fsol
            self.corr.append(corr)
            self.boxes.append(tuple(slice(v - radius, v + radius + 1)
                                    for v in (pos[2], pos[1], pos[0])))

    def apply(self, solver, q: int, gain: float = 1.0) -> None:
        """fields += gain * correction of interval q in every source box (one
        launch per box: ``conv3d_patch_add``)."""
        shape = (solver.nx, solver.ny, solver.nz)
        for box, c in zip(self.boxes, self.corr):
            if q < c.shape[0]:
                _conv.patch_add(shape, solver.fields, c[q],
                                (box[2].start, box[1].start, box[0].start),
                                (box[2].stop - box[2].start, box[1].stop - box[1].start,
                                 box[0].stop - box[0].start), gain)


_SIDE_STREAMS: dict = {}


def _side_stream(dev) -> "torch.cuda.Stream":
    """One second stream per device, kept: a stream's first use creates its
    queue, which costs milliseconds (torch hands out a different pooled
    stream at every ``torch.cuda.Stream()``)."""
    key = torch.device(dev).index or 0
    st = _SIDE_STREAMS.get(key)
    if st is None:
        st = _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return st


def attach_patch(solver, grid: EMGrid3D, n_coarse: int, every: int,
                 half: int = PATCH_HALF, tail: int | None = None) -> NearSourcePatch:
    """Compute the near-source patch of ``grid`` for ``n_coarse`` steps of
    ``solver`` (built from that grid) and keep it on the solver, where
    ``ClosureNet.correct`` uses it in place of the linear source model."""
    patch = NearSourcePatch(grid, solver.dt, n_coarse, every, half=half, device=solver.device,
                            tail=tail)
    solver._closure_patch = patch
    return patch


def coarse_solver(prob: "ClosureProblem", net, n_steps: int, every: int, patch=None):
    """The coarse solver of ``prob`` for ``n_steps`` steps; with the
    near-source patch attached when ``net`` corrects its source boxes with
    one (``patch``: one computed before for the same problem)."""
    g = prob.coarse_grid()
    sol = g.build_solver(prob.dt_coarse(), n_steps + every + 1)
    if net is not None and net.forcing and net.source_patch:
        if patch is None:
            attach_patch(sol, g, n_steps, every, half=net.patch_half,
                         tail=None if net.patch_tail < 0 else net.patch_tail)
        else:
            sol._closure_patch = patch
    return sol


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
    sol = coarse_solver(prob, net, n_coarse, every)
    if init is not None:
        for f, v in zip(sol.fields, init):
            f.copy_(v)
    for m in range(0, n_coarse, every):
        sol.run(m, every)
        if net is not None:
            net.correct(sol, gain, m, every)
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


def _crop_start(n: int, margin: int, crop: int, rng, centre: int | None = None,
                border: int = 3) -> int:
    """Start of a crop of ``crop`` cells inside [margin, n - margin); with
    ``centre``, one whose loss region (the crop minus ``border``) holds the
    whole source-model box around it when the crop is large enough, else
    the centre itself."""
    lo, hi = margin, n - margin - crop
    if hi < lo:
        return max(0, (n - crop) // 2)
    if centre is not None:
        for pad in (border + STENCIL // 2, border):
            lo_c, hi_c = max(lo, centre - crop + pad + 1), min(hi, centre - pad)
            if hi_c >= lo_c:
                return int(rng.integers(lo_c, hi_c + 1))
    return int(rng.integers(lo, hi + 1))


def _set_force_scale(net: ClosureNet, solver, every: int) -> None:
    fc = interval_forcing(solver, every)
    peak = float(fc.feats[..., 0].abs().max()) if fc is not None else 0.0
    net.force_scale = peak if peak > 0 else 1.0


def source_features(fc, q: int) -> torch.Tensor:
    """(cells, NFEAT) features of interval q: f0, f1, f0 (eps_s - 1)."""
    f0, f1 = fc.feats[q, :, 0], fc.feats[q, :, 1]
    return torch.stack([f0, f1, f0 * (fc.eps - 1.0)], dim=1)


def _accumulate_source_fit(net, solver, every, q, tc, box, slot, which, origin, crop, b, ls):
    """Normal equations of the linear source model: for every source cell
    whose whole box lies in the crop's loss region, G += phi phi^T and
    B[:, c, o] += phi * target_c(o) (physical units)."""
    fc = interval_forcing(solver, every)
    if fc is None or box.numel() == 0:
        return
    h = STENCIL // 2
    full = [c_i for c_i, (i, j, k) in enumerate(fc.xyz)
            if all(b + h <= v - o < crop - b - h for v, o in zip((i, j, k), origin))]
    if not full:
        return
    t_phys = tc.reshape(-1, COUT)[box].double() * net.scale.double()[None, :]   # (E, 6)
    phi = source_features(fc, q).double()                                      # (cells, F)
    sel = torch.zeros(len(fc.xyz), dtype=torch.bool, device=box.device)
    sel[full] = True
    keep = sel[which]
    for c_i in full:
        ls["G"] += torch.outer(phi[c_i], phi[c_i])
    pw = phi[which[keep]]                                                       # (E', F)
    for f in range(NFEAT):
        ls["B"][f].index_add_(1, slot[keep], (t_phys[keep] * pw[:, f:f + 1]).T)


def _solve_source_fit(net, ls) -> None:
    """K = G^-1 B for every (component, box cell); a tiny ridge keeps G
    invertible when a feature never varies."""
    G = ls["G"]
    if float(G.trace()) <= 0:
        return
    ridge = 1e-9 * float(G.trace()) * torch.eye(NFEAT, dtype=G.dtype, device=G.device)
    K = torch.linalg.solve(G + ridge, ls["B"].reshape(NFEAT, -1))
    net.ks = K.reshape(NFEAT, COUT, STENCIL ** 3).to(torch.float32)


def _init_net(net: ClosureNet, prob: ClosureProblem, states: list, csol, every: int) -> None:
    """Normalisation, forcing scale and correction margin from the first
    training problem."""
    set_normalisation(net, states[1:])
    if net.forcing:
        _set_force_scale(net, csol, every)
    net.margin = prob.npml_c + 4


def _far_cells(prob: ClosureProblem, net: ClosureNet) -> torch.Tensor:
    """Flat indices of the cells the far-field operator is fitted on: the
    corrected region (``net.margin`` from every face) minus the source-model
    box around the source."""
    nc = prob.nc
    m = torch.zeros((nc, nc, nc), dtype=torch.bool, device=net.device)
    a, b = net.margin, nc - net.margin
    m[a:b, a:b, a:b] = True
    i, j, k = prob.src_c
    h = STENCIL // 2
    m[max(0, k - h):k + h + 1, max(0, j - h):j + h + 1, max(0, i - h):i + h + 1] = False
    return torch.nonzero(m.view(-1)).view(-1)


def _far_features(net: ClosureNet, fields, eps_m1: torch.Tensor,
                  sel: torch.Tensor) -> torch.Tensor:
    """(n, FAR_FEAT) float64 features of the far-field operator at the flat
    cells ``sel``: f_ci(cell + off(t)) / scale[ci] for tap t and field ci
    (column t*6 + ci), then the same times eps - 1; zero outside the grid."""
    nz, ny, nx = fields[0].shape
    f = torch.stack([fields[c].double() / float(net.scale[c]) for c in range(6)])
    fp = F.pad(f, (1, 1, 1, 1, 1, 1))
    z, y, x = sel // (ny * nx), (sel // nx) % ny, sel % nx
    cols = [fp[:, z + 1 + (t // 9 - 1), y + 1 + ((t // 3) % 3 - 1), x + 1 + (t % 3 - 1)]
            for t in range(FAR_TAPS)]
    x1 = torch.cat(cols, 0).T
    return torch.cat([x1, x1 * eps_m1.reshape(-1)[sel][:, None]], 1)


def far_kernel_weights(W: torch.Tensor) -> torch.Tensor:
    """Least-squares solution (FAR_FEAT, 6) in the feature order of
    ``_far_features`` (plain/eps, tap, field) -> the flat float32 weights
    of ``conv3d_stencil_f32`` (tap, field, plain/eps, out)."""
    return W.reshape(2, FAR_TAPS, 6, COUT).permute(1, 2, 0, 3).reshape(-1).to(torch.float32)


def _far_energy_ratio(net: ClosureNet, prob: ClosureProblem, every: int, n_steps: int) -> float:
    """E energy after ``n_steps`` coarse steps of ``prob`` with the far-field
    operator alone applied every ``every`` steps, over the same run without
    it: a ratio well above 1 means the operator amplifies some mode."""
    energy = []
    for use in (False, True):
        sol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_steps + 1)
        shape = (sol.nx, sol.ny, sol.nz)
        y = torch.zeros(sol.nx * sol.ny * sol.nz, OUT_STRIDE, device=net.device)
        for m in range(0, n_steps, every):
            sol.run(m, every)
            if use:
                y.zero_()
                net.add_far_field(y, sol)
                _conv.apply_correction(shape, sol.fields, y, OUT_STRIDE, net.scale, 1.0,
                                       net.margin)
        energy.append(sum(float((f.double() ** 2).sum()) for f in sol.fields[:3]))
    return energy[1] / energy[0] if energy[0] > 0 else float("inf")


def fit_far_field(net: ClosureNet, device, seeds, n_fine: int, n_coarse: int, every: int,
                  log=None, seed: int = 0, **problem_kw) -> None:
    """Least-squares fits of the two linear parts on the same teacher-forced
    intervals: the source model (when the net has the forcing channel) and
    the linear far-field operator (float32 27-tap
    stencil of the six normalised fields and of the same times eps - 1) to
    the teacher-forced one-interval residuals, restricted fine state minus
    the coarse state ``every`` steps after starting from the restricted fine
    state, on FAR_SAMPLES random cells per interval outside the source-model
    boxes.  Ridge FAR_RIDGE; gain FAR_GAIN, halved until a run three times
    the training horizon with the operator alone keeps its field energy
    within FAR_GROWTH of the uncorrected run (an unconstrained fit amplifies
    modes the training states do not contain); off if no gain passes.
    Needs the normalisation set (``_init_net``)."""
    rng = np.random.default_rng(seed + 1)
    G = torch.zeros(FAR_FEAT, FAR_FEAT, dtype=torch.float64, device=net.device)
    B = torch.zeros(FAR_FEAT, COUT, dtype=torch.float64, device=net.device)
    ls = {"G": torch.zeros(NFEAT, NFEAT, dtype=torch.float64, device=net.device),
          "B": torch.zeros(NFEAT, COUT, STENCIL ** 3, dtype=torch.float64, device=net.device)}
    t0 = time.time()
    for sd in seeds:
        prob = ClosureProblem(sd, n_fine, device=device, **problem_kw)
        states, _ = prob.fine_states(n_coarse, every)
        csol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + every + 1)
        eps_m1 = csol.eps_r.double() - 1.0
        cand = _far_cells(prob, net)
        nc = prob.nc
        for q in range(len(states) - 1):
            csol.reset()
            for f, v in zip(csol.fields, states[q]):
                f.copy_(v)
            csol.run(q * every, every)
            cur = csol.fields
            pick = torch.from_numpy(rng.integers(0, max(1, cand.numel()), FAR_SAMPLES))
            sel = cand[pick.to(net.device)] if cand.numel() > 0 else cand
            if cand.numel() > 0:
                X = _far_features(net, cur, eps_m1, sel)
                Y = torch.stack([((states[q + 1][c] - cur[c]).double() / float(net.scale[c]))
                                 .reshape(-1)[sel] for c in range(COUT)], 1)
                G += X.T @ X
                B += X.T @ Y
            if net.forcing:
                # the source model, on the same teacher-forced interval
                box, slot, which = net._box_terms(csol, (nc,) * 3, (0, 0, 0), every)
                tgt = torch.stack([(states[q + 1][c] - cur[c]) / net.scale[c]
                                   for c in range(COUT)], dim=-1)
                _accumulate_source_fit(net, csol, every, q, tgt, box, slot, which, (0, 0, 0),
                                       nc, 0, ls)
        del states
    if net.forcing:
        _solve_source_fit(net, ls)
    if float(G.trace()) <= 0:
        # no cell outside the correction margin and the source boxes
        net.far_w.zero_()
        net.far_gain = 0.0
        if log:
            log("far-field operator: no cells to fit on, left off")
        return
    lam = FAR_RIDGE * float(G.trace()) / FAR_FEAT
    W = torch.linalg.solve(G + lam * torch.eye(FAR_FEAT, dtype=G.dtype, device=G.device), B)
    wk = far_kernel_weights(W)
    check = ClosureProblem(seeds[-1], n_fine, device=device, **problem_kw)
    gain = FAR_GAIN
    for _ in range(4):
        net.far_w = (wk * gain).contiguous()
        net.far_gain = gain
        ratio = _far_energy_ratio(net, check, every, 3 * n_coarse)
        if log:
            log(f"far-field operator: gain {gain:g}, energy ratio {ratio:.4f} after "
                f"{3 * n_coarse} steps ({time.time() - t0:.1f} s)")
        if ratio <= FAR_GROWTH:
            return
        gain *= 0.5
    net.far_w.zero_()
    net.far_gain = 0.0


def select_net_gain(net: ClosureNet, device, seeds, n_fine: int, n_coarse: int, every: int,
                    gains=(0.0, 0.25, 0.5, 1.0), log=None, tol: float = 0.01,
                    near_radius: float | None = None, **problem_kw) -> dict:
    """Scale of the net's correction, chosen on validation seeds (neither
    training nor test seeds) by the trajectory error of E of the corrected
    coarse runs against the restricted fine states, summed over the seeds:
    the smallest gain within ``tol`` (relative) of the lowest error, so that
    a difference at the noise level of two seeds does not switch the net
    on.  With ``near_radius`` the error leaves out the cells within that
    many coarse cells of the source (the far-field error, where the net
    corrects; the source's own cells otherwise weigh most).  Sets
    ``net.net_gain``; returns the errors per gain."""
    errs = {g: 0.0 for g in gains}
    for sd in seeds:
        prob = ClosureProblem(sd, n_fine, device=device, **problem_kw)
        states, _ = prob.fine_states(n_coarse, every)
        mask = (None if near_radius is None else
                far_mask(prob.nc, prob.npml_c, prob.src_c, near_radius, net.device))
        patch = None
        for g in gains:
            net.net_gain = g
            net.pack()
            csol = coarse_solver(prob, net, n_coarse, every, patch)
            patch = getattr(csol, "_closure_patch", None)
            num = den = 0.0
            for q, m in enumerate(range(0, n_coarse, every)):
                csol.run(m, every)
                net.correct(csol, 1.0, m, every)
                a_, b_ = _err_terms(csol.fields, states[q + 1], prob.npml_c, mask)
                num, den = num + a_, den + b_
            errs[g] += math.sqrt(num / den) if den > 0 else 0.0
        del states
    best = min(errs.values())
    net.net_gain = min(g for g in gains if errs[g] <= best * (1.0 + tol))
    if log:
        log("net gain on validation seeds " + ", ".join(f"{g:g}: {e / len(seeds):.5f}"
                                                        for g, e in errs.items())
            + f" -> {net.net_gain:g}")
    return errs


def _reinit_net_weights(net: ClosureNet, seed: int) -> None:
    """Fresh network weights (the initialisation of ``ClosureNet``, with the
    output layer at zero so that the net starts with no correction), Adam
    state cleared; the linear parts and the normalisation are kept."""
    fresh = ClosureNet(net.device, seed=seed, forcing=net.forcing)
    fresh.params["w3"].zero_()
    net.params = fresh.params
    net.adam = fresh.adam
    net.step_count = 0
    net._packed = None
    net.version += 1


def far_field_transpose_weights(net: ClosureNet) -> tuple:
    """Weights of the far-field operator's transpose for ``conv3d_stencil_f32``.

    The operator (gain folded in) maps the fields f to
    y_co(i) = sum_{t,ci} (W0[t,ci,co] + e(i) W1[t,ci,co]) f_ci(i + o_t) / s_ci
    with e = eps - 1.  For u = dL/dy its transpose is
    dL/df_ci(j) = sum_{t,co} W0[t,ci,co] / s_ci u_co(j - o_t)
                + sum_{t,co} W1[t,ci,co] / s_ci (e u_co)(j - o_t),
    two stencils of the same kernel with the taps mirrored (t -> 26 - t),
    field and output channels swapped, 1 / s_ci folded into the weights
    and unit scales: applied to u and to e u."""
    w = net.far_w.view(FAR_TAPS, 6, 2, COUT)
    inv = (1.0 / net.scale).view(1, 1, 6)
    out = []
    for pe in (0, 1):
        v = torch.zeros(FAR_TAPS, COUT, 2, 6, device=net.device)
        # v[t', co, 0, ci] = W[t, ci, pe, co] / s_ci,  t' = 26 - t
        v[:, :, 0, :] = (w[:, :, pe, :].permute(0, 2, 1) * inv).flip(0)
        out.append(v.reshape(-1).contiguous())
    return tuple(out)


def _rollout_window(net: ClosureNet, csol, states: list, q0: int, q1: int, every: int,
                    wm: torch.Tensor, cm: torch.Tensor, box, den: float, wc: list,
                    far_t: tuple | None, grads: bool = True):
    """Run intervals q0 .. q1-1 of a corrected coarse rollout (the solver
    holds the state at step q0 * every) and return the loss of those
    intervals (trajectory error against ``states`` after each correction,
    weights ``wm`` per cell and ``wc`` per field, over ``den``, or over the
    window's own reference sum when ``den`` is None) and, with ``grads``,
    its gradient with respect to the net's weights (see
    ``train_rollout``)."""
    from jaxphys_amd import _native

    nc = csol.nx
    shape = (nc, nc, nc)
    npos = nc ** 3
    sc = [float(v) for v in net.scale]
    patch = getattr(csol, "_closure_patch", None)
    net.pack()
    x0 = net.pack_zero(csol)
    y0 = net.forward(x0, shape, train=grads, tag="z").clone()
    grad_state = []
    loss = 0.0
    if den is None:
        # normalised by the window's own reference sum
        den = sum(wc[c] * float((wm * (states[q + 1][c] / sc[c]) ** 2).sum())
                  for q in range(q0, q1) for c in range(6))
    for q in range(q0, q1):
        m = q * every
        csol.run(m, every)
        x = net.pack_fields(csol, m, every)
        tag = f"r{q - q0}"
        xq = net._buf("x" + tag, tuple(x.shape), x.dtype)
        xq.copy_(x)
        y = net.forward(xq, shape, train=grads, tag=tag)
        y.sub_(y0)
        if net.far_gain > 0:
            net.add_far_field(y, csol)
        if box is not None:
            y.view(-1, OUT_STRIDE)[box] = 0.0
        _conv.apply_correction(shape, csol.fields, y, OUT_STRIDE, net.scale, 1.0, net.margin)
        if patch is not None:
            patch.apply(csol, q, 1.0)
        gs = []
        for c in range(6):
            r = csol.fields[c] - states[q + 1][c]
            loss += wc[c] * float((wm * r * r).sum()) / (sc[c] ** 2)
            if grads:
                gs.append((2.0 * wc[c] / (den * sc[c] ** 2)) * wm * r)
        grad_state.append(gs)
    loss /= den
    if not grads:
        return loss, None
    # adjoint of the state after each correction, from the window's last
    # interval to its first; the psi adjoints start at zero (every loss and
    # correction term lies outside the absorbing slabs, so the deferred
    # psi_H-adjoint update of the first step has nothing to carry)
    lam = [torch.zeros(shape, device=net.device) for _ in range(6)]
    lam_psi = None if csol.psi is None else torch.zeros_like(csol.psi)
    dy = torch.zeros(npos, OUT_STRIDE, device=net.device)
    dy_sum = torch.zeros(npos, OUT_STRIDE, device=net.device)
    tmp = torch.zeros(npos, OUT_STRIDE, device=net.device) if far_t is not None else None
    ones = torch.ones(6, device=net.device)
    e = (csol.eps_r - 1.0).view(-1)
    out = None
    for q in range(q1 - 1, q0 - 1, -1):
        for c in range(6):
            lam[c].add_(grad_state[q - q0][c])
            dy[:, c] = (sc[c] * cm) * lam[c].view(-1)
        g = net.backward(dy, tag=f"r{q - q0}")
        out = g if out is None else {k: out[k] + g[k] for k in g}
        dy_sum.add_(dy)
        if far_t is not None:
            # the far-field operator's contribution to the state's adjoint
            u = [dy[:, c].reshape(shape).contiguous() for c in range(6)]
            ue = [(dy[:, c] * e).reshape(shape).contiguous() for c in range(6)]
            tmp.zero_()
            _conv.stencil_f32(shape, u, csol.eps_r, ones, far_t[0], tmp, OUT_STRIDE)
            _conv.stencil_f32(shape, ue, csol.eps_r, ones, far_t[1], tmp, OUT_STRIDE)
            for c in range(6):
                lam[c].view(-1).add_(tmp[:, c])
        if q > q0:
            for _ in range(every):
                _native.adjoint_step(csol.plan, lam, lam_psi)
    g0 = net.backward(dy_sum, tag="z")
    return loss, {k: out[k] - g0[k] for k in out}


def rollout_setup(net: ClosureNet, prob: "ClosureProblem", csol, states: list, every: int,
                  near_radius: float = 3.0, h_weight: float = 1.0):
    """Per-problem tensors of ``_rollout_window``: the loss weights (the
    far-field evaluation cells: ``npml`` cells or more inside the faces
    minus ``near_radius`` around the source), the correction mask (where
    the correction is applied: ``net.margin`` inside the faces, outside the
    source boxes), the source-box cells, the loss normaliser and the field
    weights."""
    nc = prob.nc
    shape = (nc, nc, nc)
    sc = [float(v) for v in net.scale]
    wc = [1.0, 1.0, 1.0, h_weight, h_weight, h_weight]
    wm = torch.zeros(shape, device=net.device)
    a_, b_ = prob.npml_c, nc - prob.npml_c
    wm[a_:b_, a_:b_, a_:b_] = far_mask(nc, prob.npml_c, prob.src_c, near_radius,
                                        net.device).to(torch.float32)
    n_int = len(states) - 1
    den = sum(wc[c] * float((wm * (states[q + 1][c] / sc[c]) ** 2).sum())
              for q in range(n_int) for c in range(6))
    cm = torch.zeros(shape, device=net.device)
    mg = net.margin
    cm[mg:nc - mg, mg:nc - mg, mg:nc - mg] = 1.0
    cm = cm.view(-1)
    box = net._box_terms(csol, shape, (0, 0, 0), every)[0] if net.forcing else None
    if box is not None:
        cm[box] = 0.0
    return wm, cm, box, den, wc


def train_rollout(net: ClosureNet, device, seeds, n_fine: int = 256, n_coarse: int = 400,
                  every: int = 10, epochs: int = 8, lr: float = 6e-5,
                  lr_final: float | None = 6e-6, window: int | None = 10,
                  h_weight: float = 0.0, near_radius: float = 3.0, reinit: bool = True,
                  relative: bool = True, seed: int = 0, val_seeds=None, log=None,
                  **problem_kw) -> list:
    """Train the net through the corrected coarse rollouts themselves
    (backpropagation through time) instead of on one-interval residuals.

    For each training seed the closure (near-source patch, far-field
    operator and the net at gain 1) runs the whole ``n_coarse``-step
    rollout from the zero state, keeping the net's activations of every
    correction.  The loss is the trajectory error of that rollout against
    the restricted fine states after each correction, over the cells the
    far-field error is measured on (the interior ``npml`` cells or more
    from the faces, minus ``near_radius`` cells around the source), of the
    normalised fields (H weighted by ``h_weight``), divided by the
    reference's own sum.  Its gradient flows backwards through the coarse
    steps with the transposed Yee step (``yee_adjoint_step``, the kernels
    of the adjoint), through the far-field operator's transpose (the same
    stencil kernel with mirrored weights) and into the net at every
    correction.  The net's output enters the gradient as an additive
    forcing: its dependence on the state it is computed from is not
    differentiated (exact while that dependence is weak; the net starts
    with a zero output layer).  With ``window`` the rollout is cut into
    windows of that many intervals, each followed by its own update (the
    next window continues from the state the previous one reached).  With
    ``relative`` each window's loss is its relative error (the square root
    of its squared error over its reference sum), so that every seed and
    window weighs alike whatever its error; otherwise the squared relative
    error over the whole rollout.  The linear parts (normalisation, far-field operator, correction margin,
    patch settings) must already be set, e.g. by ``train_closure`` or
    ``ClosureNet.load``; ``reinit`` starts from fresh network weights.
    With ``val_seeds`` the net's gain is chosen on them afterwards
    (``select_net_gain``).  Returns the loss of every update."""
    rng = np.random.default_rng(seed)
    if reinit:
        _reinit_net_weights(net, seed)
    net.net_gain = 1.0
    n_int = n_coarse // every
    window = n_int if not window else int(window)
    per_seed = (n_int + window - 1) // window
    lr_final = lr if lr_final is None else lr_final
    n_updates = epochs * len(seeds) * per_seed
    upd, history = 0, []
    far_t = far_field_transpose_weights(net) if net.far_gain > 0 else None
    t_start = time.time()
    for ep in range(epochs):
        for sd in [seeds[i] for i in rng.permutation(len(seeds))]:
            prob = ClosureProblem(sd, n_fine, device=device, **problem_kw)
            states, _ = prob.fine_states(n_coarse, every)
            csol = coarse_solver(prob, net, n_coarse, every)
            wm, cm, box, den, wc = rollout_setup(net, prob, csol, states, every, near_radius,
                                                 h_weight)
            for q0 in range(0, n_int, window):
                loss, g = _rollout_window(net, csol, states, q0, min(n_int, q0 + window), every,
                                          wm, cm, box, None if relative else den, wc, far_t)
                if relative and loss > 0:
                    # d sqrt(L) = dL / (2 sqrt(L))
                    g = {k: v * (0.5 / math.sqrt(loss)) for k, v in g.items()}
                    loss = math.sqrt(loss)
                frac = upd / max(1, n_updates - 1)
                lr_now = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
                upd += 1
                if math.isfinite(loss) and all(bool(torch.isfinite(v).all()) for v in g.values()):
                    net.adam_step(g, lr=lr_now)
                history.append(loss)
                if not all(bool(torch.isfinite(f).all()) for f in csol.fields[:3]):
                    break
            if log:
                log(f"rollout epoch {ep} seed {sd}: loss {np.mean(history[-per_seed:]):.4e} "
                    f"({time.time() - t_start:.1f} s)")
            del states, csol
    # the per-correction activations are training buffers only
    net._saved = {}
    net._ws = {}
    if val_seeds:
        select_net_gain(net, device, list(val_seeds), n_fine, n_coarse, every, log=log,
                        near_radius=near_radius, **problem_kw)
    return history


def train_closure(device, seeds, n_fine: int = 256, n_coarse: int = 400, every: int = 10,
                  epochs: int = 2, crop: int = 64, lr: float = 1e-3, net: ClosureNet | None = None,
                  log=None, seed: int = 0, onpolicy_from: int | None = None,
                  lr_final: float | None = None, source_crops: float = 0.5,
                  forcing: bool = True, reset_tol: float = 1.0, far_field: bool = True,
                  val_seeds=None, onpolicy_with_net: bool = False, warm_start: bool = False,
                  net_in_box: bool = False, source_patch: bool = True,
                  patch_half: int = PATCH_HALF, patch_tail: int = PATCH_TAIL, **problem_kw):
    """Train on fine-run/coarse pairs generated by the FDTD solver.

    A sample is a coarse state at coarse step m + every and its target, the
    restricted fine state at that step minus it.  Teacher-forced samples (all
    epochs before ``onpolicy_from``) start the coarse solver from the
    restricted fine state at step m (psi restarted at zero) and run it
    ``every`` steps.  On-policy samples (epochs from ``onpolicy_from`` on)
    are the states a corrected coarse run reaches before each correction:
    corrected by the source model and the far-field operator only (the net
    then learns the error that accumulates over many intervals of that
    stable run), or with ``onpolicy_with_net`` by the whole closure.  The
    source model is fitted on teacher-forced samples only.  With
    ``warm_start`` a teacher-forced sample starts one interval earlier and
    applies the linear parts (source model, far-field operator) after that
    interval, so the transient of a restart from a restricted fine state
    has passed when the sample's interval runs.  Loss: mean squared normalised error on an interior
    crop (a 3-cell border, where zero padding differs from the full domain,
    is excluded); a share ``source_crops`` of the crops holds the source.
    Adam, learning rate cosine-annealed from ``lr`` to ``lr_final``.
    With ``far_field`` the linear far-field operator is fitted first
    (``fit_far_field``) and the net learns what remains after it; with
    ``val_seeds`` the net's gain is then chosen on them
    (``select_net_gain``).  With ``net_in_box`` the net also corrects the
    source boxes, learning there what the source model leaves."""
    rng = np.random.default_rng(seed)
    net = net or ClosureNet(device, seed=seed, forcing=forcing)
    net.net_in_box = bool(net_in_box) and net.forcing
    net.source_patch, net.patch_half = bool(source_patch), int(patch_half)
    net.patch_tail = int(patch_tail)
    initialised = False
    if far_field:
        prob0 = ClosureProblem(seeds[0], n_fine, device=device, **problem_kw)
        states0, _ = prob0.fine_states(n_coarse, every)
        _init_net(net, prob0, states0,
                  prob0.coarse_grid().build_solver(prob0.dt_coarse(), n_coarse + every + 1), every)
        del states0
        initialised = True
        fit_far_field(net, device, seeds, n_fine, n_coarse, every, log=log, seed=seed,
                      **problem_kw)
    history, target_var = [], []
    ls = {"G": torch.zeros(NFEAT, NFEAT, dtype=torch.float64, device=net.device),
          "B": torch.zeros(NFEAT, COUT, STENCIL ** 3, dtype=torch.float64, device=net.device)}
    t_start = time.time()
    lr_final = lr if lr_final is None else lr_final
    n_updates = epochs * len(seeds) * (n_coarse // every)
    upd = 0
    b = 3
    for ep in range(epochs):
        onpolicy = onpolicy_from is not None and ep >= onpolicy_from
        for s_i, sd in enumerate(seeds):
            prob = ClosureProblem(sd, n_fine, device=device, **problem_kw)
            states, _ = prob.fine_states(n_coarse, every)
            if onpolicy or warm_start:
                csol = coarse_solver(prob, net, n_coarse, every)
            else:
                csol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + every + 1)
            if not initialised:
                _init_net(net, prob, states, csol, every)
                initialised = True
            nc = prob.nc
            # crops whose loss region (crop minus its b-cell border) starts at
            # the correction margin
            margin = max(0, net.margin - b)
            x0 = net.pack_zero(csol).view(nc, nc, nc, CIN)
            inputs = None
            if onpolicy:
                # the corrected rollout's own states before each correction;
                # a rollout that has run away (not finite, or error above
                # reset_tol) restarts from the fine state of that time
                inputs = []
                net.pack()
                for q_, m in enumerate(range(0, n_coarse, every)):
                    csol.run(m, every)
                    inputs.append([f.clone() for f in csol.fields])
                    net.correct(csol, 1.0, m, every, use_net=onpolicy_with_net)
                    bad = not all(bool(torch.isfinite(f).all()) for f in csol.fields[:3])
                    if bad or rel_err(csol.fields, states[q_ + 1], margin=prob.npml_c) > reset_tol:
                        for f, v in zip(csol.fields, states[q_ + 1]):
                            f.copy_(v)
            order = rng.permutation(len(states) - 1)
            src = prob.src_c  # (i, j, k) = (x, y, z)
            for q in order:
                m = int(q) * every
                if inputs is None:
                    csol.reset()
                    if warm_start and q >= 1:
                        # one interval earlier, corrected by the linear parts,
                        # so the restart transient has passed before the sample
                        for f, v in zip(csol.fields, states[q - 1]):
                            f.copy_(v)
                        csol.run(m - every, every)
                        net.correct(csol, 1.0, m - every, every, use_net=False)
                    else:
                        for f, v in zip(csol.fields, states[q]):
                            f.copy_(v)
                    csol.run(m, every)
                    cur = csol.fields
                else:
                    cur = inputs[q]
                x = net.pack_fields(csol, m, every, fields=cur).view(nc, nc, nc, CIN)
                tgt = torch.stack([(states[q + 1][c] - cur[c]) / net.scale[c]
                                   for c in range(6)], dim=-1)
                far = None
                if net.far_gain > 0:
                    # the net learns what the far-field operator leaves (the
                    # source model, inside the boxes where the operator is
                    # not applied, keeps the full residual)
                    far = net._buf("far", (nc ** 3, OUT_STRIDE), torch.float32)
                    far.zero_()
                    net.add_far_field(far, csol, fields=cur)
                    far = far.view(nc, nc, nc, OUT_STRIDE)[..., :6]
                near = rng.random() < source_crops
                cz = _crop_start(nc, margin, crop, rng, src[2] if near else None, b)
                cy = _crop_start(nc, margin, crop, rng, src[1] if near else None, b)
                cx = _crop_start(nc, margin, crop, rng, src[0] if near else None, b)
                sl = (slice(cz, cz + crop), slice(cy, cy + crop), slice(cx, cx + crop))
                xc = x[sl].contiguous().view(-1, CIN)
                tc = tgt[sl]
                tn = tc if far is None else tc - far[sl]
                # correction = net(state) - net(zero state)
                y = net.forward(xc, (crop, crop, crop), train=True).view(crop, crop, crop, 8)
                y0 = net.forward(x0[sl].contiguous().view(-1, CIN), (crop, crop, crop),
                                 train=True, tag="0").view(crop, crop, crop, 8)
                y = y - y0
                # the net's loss region: the crop minus its b-cell border and
                # minus the source-model boxes (where its output is zeroed)
                w = torch.zeros((crop,) * 3, device=y.device)
                w[(slice(b, crop - b),) * 3] = 1.0
                if net.forcing:
                    box, slot, which = net._box_terms(csol, (crop,) * 3, (cx, cy, cz), every)
                    if net.net_in_box:
                        # inside the boxes the net learns what the source
                        # model leaves (the far-field operator is not
                        # applied there)
                        sm = net._buf("sm", (crop ** 3, OUT_STRIDE), torch.float32)
                        sm.zero_()
                        net.add_source_model(sm, csol, (crop,) * 3, (cx, cy, cz), m, every)
                        tn = tn.clone()
                        tnv = tn.view(-1, 6)
                        if far is not None:
                            tnv[box] += far[sl].reshape(-1, 6)[box]
                        tnv[box] -= sm[box, :6]
                    else:
                        w.view(-1)[box] = 0.0
                        y.view(-1, OUT_STRIDE)[box] = 0.0
                    if not onpolicy and not far_field:
                        _accumulate_source_fit(net, csol, every, q, tc, box, slot, which,
                                               (cx, cy, cz), crop, b, ls)
                err = torch.zeros_like(y)
                err[..., :6] = (y[..., :6] - tn) * w[..., None]
                n_in = float(w.sum()) * 6
                loss = float((err ** 2).sum()) / n_in
                target_var.append(float(((tn * w[..., None]) ** 2).sum()) / n_in)
                frac = upd / max(1, n_updates - 1)
                lr_now = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
                upd += 1
                if not math.isfinite(loss):
                    continue  # a sample from a diverged state: no update
                g_out = (2.0 / n_in) * err.view(-1, 8)
                grads = net.backward(g_out)
                grads0 = net.backward(g_out, tag="0")
                for k_ in grads:
                    grads[k_] = grads[k_] - grads0[k_]
                net.adam_step(grads, lr=lr_now)
                history.append(loss)
            if net.forcing and not far_field and s_i == len(seeds) - 1:
                _solve_source_fit(net, ls)
            if log:
                log(f"epoch {ep}{' on-policy' if onpolicy else ''} seed {sd}: mean loss "
                    f"{np.mean(history[-len(order):]):.4e} (zero correction "
                    f"{np.mean(target_var[-len(order):]):.4e}) ({time.time() - t_start:.1f} s)")
            del states, inputs
    if val_seeds:
        select_net_gain(net, device, list(val_seeds), n_fine, n_coarse, every, log=log,
                        **problem_kw)
    return net, history


def evaluate_closure(net: ClosureNet, seed: int, n_fine: int = 256, n_coarse: int = 400,
                     every: int = 10, gain: float = 1.0, near_radius: float = 3.0,
                     **problem_kw):
    """Fine reference vs coarse with / without the closure on one seed.

    Returns a dict with the interior relative L2 errors of E against the
    restricted fine state, at the final time (``err_*``) and over all
    correction times (``err_traj_*``), each also without the cells within
    ``near_radius`` coarse cells of the source (``err_far_*``,
    ``err_traj_far_*``: the point source's own cells, where both grids hold
    a resolution-dependent singular value), and the wall times of the fine
    run and of the corrected coarse run."""
    prob = ClosureProblem(seed, n_fine, device=net.device, **problem_kw)
    # fine run (timed on its own: the plain fine FDTD it replaces), after two
    # untimed steps that load its kernels
    g = prob.fine_grid()
    sol = g.build_solver(prob.dt_fine(), 2 * n_coarse + 2)
    sol.run(0, 2)
    sol.reset()
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
    far = far_mask(prob.nc, margin, prob.src_c, near_radius, net.device)
    # reference states at every correction time (restricted fine states)
    states, fsol = prob.fine_states(n_coarse, every)
    del fsol
    runs = [("coarse", None, True, True), ("closure", net, True, True)]
    if net.forcing:
        runs.append(("source_only", net, False, False))
    if net.far_gain > 0:
        runs.append(("far_field", net, False, True))
    if net.net_gain == 0:
        # what the net would do at full gain (validation switched it off)
        runs.append(("net_gain1", net, True, True))
    saved_gain = net.net_gain
    patch = None
    # untimed: one interval of the corrected run (and one patch computation),
    # which loads the kernels and library routines they use, so the timed
    # runs below measure the method itself (the patch is computed again there)
    wg = prob.coarse_grid()
    wsol = wg.build_solver(prob.dt_coarse(), n_coarse + every + 1)
    if net.forcing and net.source_patch:
        attach_patch(wsol, wg, n_coarse, every, half=net.patch_half,
                     tail=None if net.patch_tail < 0 else net.patch_tail)
    net.pack()
    wsol.run(0, every)
    net.correct(wsol, gain, 0, every)
    del wsol, wg
    for label, model, use_net, use_far in runs:
        net.net_gain = 1.0 if label == "net_gain1" else saved_gain
        grid = prob.coarse_grid()
        csol = grid.build_solver(prob.dt_coarse(), n_coarse + every + 1)
        if model is not None:
            model.pack()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if model is not None and model.forcing and model.source_patch:
            # the patch is part of the closure's cost: computed inside the
            # first timed run, reused by the others
            if patch is None:
                patch = attach_patch(csol, grid, n_coarse, every, half=model.patch_half,
                                     tail=None if model.patch_tail < 0 else model.patch_tail)
            else:
                csol._closure_patch = patch
        for m in range(0, n_coarse, every):
            csol.run(m, every)
            if model is not None:
                model.correct(csol, gain, m, every, use_net, use_far)
        torch.cuda.synchronize()
        out[f"t_{label}_s"] = time.perf_counter() - t0
        a_, b_ = _err_terms(csol.fields, ref, margin)
        out[f"err_{label}"] = math.sqrt(a_ / b_) if b_ > 0 else 0.0
        a_, b_ = _err_terms(csol.fields, ref, margin, far)
        out[f"err_far_{label}"] = math.sqrt(a_ / b_) if b_ > 0 else 0.0
        # the same rollout again, untimed, compared at every correction time
        csol.reset()
        num = den = num_f = den_f = 0.0
        for q, m in enumerate(range(0, n_coarse, every)):
            csol.run(m, every)
            if model is not None:
                model.correct(csol, gain, m, every, use_net, use_far)
            a_, b_ = _err_terms(csol.fields, states[q + 1], margin)
            num, den = num + a_, den + b_
            a_, b_ = _err_terms(csol.fields, states[q + 1], margin, far)
            num_f, den_f = num_f + a_, den_f + b_
        out[f"err_traj_{label}"] = math.sqrt(num / den) if den > 0 else 0.0
        out[f"err_traj_far_{label}"] = math.sqrt(num_f / den_f) if den_f > 0 else 0.0
    net.net_gain = saved_gain
    out["speedup"] = out["t_fine_s"] / out["t_closure_s"]
    return out


def far_mask(n: int, margin: int, src_xyz, radius: float, device) -> torch.Tensor:
    """float mask over the interior box [margin, n - margin)^3 (internal
    (z, y, x) order): 0 within ``radius`` cells of the source node, else 1."""
    m = n - 2 * margin
    ax = torch.arange(margin, n - margin, device=device, dtype=torch.float32)
    i, j, k = src_xyz
    d2 = ((ax - k)[:, None, None] ** 2 + (ax - j)[None, :, None] ** 2
          + (ax - i)[None, None, :] ** 2)
    out = (d2 > radius * radius).to(torch.float64)
    assert out.shape == (m, m, m)
    return out


def _err_terms(a: list, b: list, margin: int, mask: torch.Tensor | None = None):
    """Squared-difference and squared-reference sums of E on the interior
    (optionally weighted by ``mask`` over that interior)."""
    num = den = 0.0
    for c in (0, 1, 2):
        x, y = a[c], b[c]
        if margin:
            x = x[margin:-margin, margin:-margin, margin:-margin]
            y = y[margin:-margin, margin:-margin, margin:-margin]
        d2 = (x - y).double() ** 2
        r2 = y.double() ** 2
        if mask is not None:
            d2, r2 = d2 * mask, r2 * mask
        num += float(d2.sum())
        den += float(r2.sum())
    return num, den
