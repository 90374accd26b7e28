"""GPU adjoint: dJ/d eps_r from libyee_hip.so's transposed step vs the
float64 torch-autograd oracle (exact discrete gradient of the same scheme)
and vs central finite differences of the GPU forward."""

import math

import numpy as np
import pytest
import torch

from jaxphys_amd import DielectricRegion, EMGrid3D, PointSource3D
from jaxphys_amd.adjoint import AdjointProblem, FDTDObjective
from jaxphys_amd.em.constants import C0, EPS0
from oracle import autograd_fdtd as A
from oracle import ext_fdtd as X

pytestmark = pytest.mark.gpu

DT = 0.5 / math.sqrt(3.0) / C0


def _setup(boundary, lossy, size=(22, 20, 18), npml=4, n=70, seed=3):
    rng = np.random.default_rng(seed)
    eps = (1.0 + 2.0 * rng.random(size)).astype(np.float32)
    sigma = None
    if lossy:
        sigma = np.zeros(size)
        sigma[:, :, 6:9] = 0.03 * EPS0 * C0
    box = (8, 7, 6, 6, 5, 4)
    tab = X.waveform_table("gaussian_derivative", n, t0=15, spread=5)
    src = (2, (11, 10, 3))
    mon = [(2, (i, j, 13)) for i in range(8, 13) for j in range(7, 11)]
    return eps, sigma, box, tab, src, mon


def _gpu_grid(size, boundary, npml, eps, sigma, tab, src):
    g = EMGrid3D(size=size, resolution=1.0, boundary=boundary, pml_layers=npml)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    comp, pos = src
    g.add_source(PointSource3D(0.0, pos, 1.0, "xyz"[comp], waveform=tab))
    return g


@pytest.mark.parametrize("boundary,lossy,size", [
    ("cpml", False, (22, 20, 18)), ("cpml", True, (22, 20, 18)),
    ("absorbing", False, (22, 20, 18)), ("walls", False, (22, 20, 18)),
    # wide enough for the slab-free interior launches
    ("cpml", True, (76, 26, 22))])
def test_adjoint_matches_autograd_oracle(gpu, boundary, lossy, size):
    npml, n = 4, 70
    eps, sigma, box, tab, src, mon = _setup(boundary, lossy, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    print(f"[adjoint] {boundary} lossy={lossy} size={size}", flush=True)
    g = _gpu_grid(size, boundary, npml, eps, sigma, tab, src)
    prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=16)
    eps_box = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu)
    J, grad = prob.value_and_grad(eps_box)
    mask = np.zeros(size, bool)
    for _, p in mon:
        mask[p] = True
    J_ref, g_ref = A.gradient(size, 1.0, DT, boundary, eps.astype(np.float64),
                              [(src[0], src[1], tab)], n, npml, sigma, mask)
    g_ref_box = g_ref[x0:x0 + bx, y0:y0 + by, z0:z0 + bz]
    got = grad.cpu().numpy()
    rel = np.linalg.norm(got - g_ref_box) / np.linalg.norm(g_ref_box)
    print(f"[adjoint] J {J:.6e} vs {J_ref:.6e}; grad rel-L2 {rel:.2e}", flush=True)
    assert J == pytest.approx(J_ref, rel=1e-4)
    assert rel < 2e-3


def test_adjoint_checkpoint_interval_invariance(gpu):
    size, npml, n = (22, 20, 18), 4, 50
    eps, sigma, box, tab, src, mon = _setup("cpml", False, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    eps_box = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu)
    grads = []
    for K, rec in ((7, False), (50, False), (100, False), (7, True)):
        g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
        prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=K,
                              record_box=rec)
        grads.append(prob.value_and_grad(eps_box)[1])
    # recomputation reproduces the forward bit for bit, so neither K nor
    # recording the box history in the forward run changes anything
    for q in range(1, 4):
        assert torch.equal(grads[0], grads[q]), q


@pytest.mark.parametrize("rec", [True, False])
def test_probe_map_equals_probe_scatter(gpu, rec):
    """The objective terms added by the E-adjoint kernel through the cell ->
    probe map equal those of the per-step scatter launch bit for bit, on
    probes of all three E components; a probe listed twice is one term with
    the summed weight."""
    size, npml, n = (22, 20, 18), 4, 40
    eps, sigma, box, tab, src, mon = _setup("cpml", True, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    eps_box = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu)
    probes = [("ex", (9, 8, 13)), ("ey", (10, 9, 13)), ("ez", (10, 9, 13)),
              ("ez", (12, 10, 14)), ("ex", (3, 15, 12))]
    weights = [1.0, 0.5, 2.0, 1.5, 0.25]
    res = []
    for mapped in (True, False):
        g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
        prob = AdjointProblem(g, n, DT, box, probes, weights, checkpoint_every=16,
                              record_box=rec)
        assert prob.probe_map is not None
        if not mapped:
            prob.probe_box, prob.probe_map = None, None
        res.append(prob.value_and_grad(eps_box))
    assert res[0][0] == res[1][0]
    assert torch.equal(res[0][1], res[1][1])
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    dup = AdjointProblem(g, n, DT, box, probes + [probes[3]], weights + [0.5],
                         checkpoint_every=16, record_box=rec)
    assert dup.nprobe == len(probes)
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    one = AdjointProblem(g, n, DT, box, probes, weights[:3] + [2.0, 0.25],
                         checkpoint_every=16, record_box=rec)
    Jd, gd = dup.value_and_grad(eps_box)
    Jo, go = one.value_and_grad(eps_box)
    assert Jd == Jo
    assert torch.equal(gd, go)


@pytest.mark.parametrize("nsrc", [1, 10])
def test_run_record_equals_steps_and_gathers(gpu, nsrc):
    """yee_run_record's box record of every step is d = E^{n+1} - Ca E^n
    followed by E^n (a lossy plan), and its probe values the fields after
    the step, bit for bit against plain steps with gathers before and after
    (d in NumPy float32: the same two operations, no contraction): when the
    E kernel records (every source fused into it, probes mapped) and with
    the record launches around the step (probes without the map, or more
    sources than the E kernel fuses)."""
    from jaxphys_amd import _native

    size, npml, n = (22, 20, 18), 4, 30
    eps, sigma, box, tab, src, mon = _setup("cpml", True, size, npml, n)
    g = EMGrid3D(size=size, resolution=1.0, boundary="cpml", pml_layers=npml)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    rng = np.random.default_rng(7)
    for q in range(nsrc):  # outside the design box (z 6..9)
        pos = (int(rng.integers(2, 20)), int(rng.integers(2, 18)), int(rng.integers(2, 5)))
        g.add_source(PointSource3D(0.0, pos, 1.0, "xyz"[q % 3], waveform=tab))
    probes = [("ex", (9, 8, 13)), ("ey", (10, 9, 13)), ("ez", (10, 9, 13)), ("ez", (3, 15, 12))]
    prob = AdjointProblem(g, n, DT, box, probes, checkpoint_every=n)
    sol = prob.solver
    assert sol.ca is not None and prob.rec_floats == 6 * prob.bcells
    nx, ny, nz = sol.nx, sol.ny, sol.nz
    x0, y0, z0, bx, by, bz = prob.box_int
    ca = sol.ca[z0:z0 + bz, y0:y0 + by, x0:x0 + bx].reshape(-1).cpu().numpy()
    ca3 = np.concatenate([ca, ca, ca])
    bc3 = 3 * prob.bcells
    ref_box = np.empty((n, 2 * bc3), np.float32)
    ref_probe = torch.empty((n, prob.nprobe), device=gpu)
    before = torch.empty(bc3, device=gpu)
    after = torch.empty(bc3, device=gpu)
    sol.reset()
    for s in range(n):
        _native.box_gather(sol.fields, (nx, ny, nz), prob.box_int, before)
        sol.run_two_pass(s, 1)
        _native.box_gather(sol.fields, (nx, ny, nz), prob.box_int, after)
        _native.probe_gather(sol.fields, prob.probe_comp, prob.probe_idx, ref_probe[s])
        e0, e1 = before.cpu().numpy(), after.cpu().numpy()
        ref_box[s, :bc3] = e1 - ca3 * e0
        ref_box[s, bc3:] = e0
    for mapped in (True, False):
        out_box = torch.full((n, 2 * bc3), float("nan"), device=gpu)
        out_probe = torch.full((n, prob.nprobe), float("nan"), device=gpu)
        sol.reset()
        _native.run_record(sol.plan, sol.fields, sol.psi, 0, n, prob.probe_comp,
                           prob.probe_idx, out_probe, box=prob.box_int, box_out=out_box,
                           probe_box=prob.probe_box if mapped else None,
                           probe_map=prob.probe_map if mapped else None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_box.cpu().numpy(), ref_box, err_msg=str(mapped))
        assert torch.equal(out_probe, ref_probe), mapped


def test_adjoint_step_equals_one_step_of_adjoint_run(gpu):
    """yee_adjoint_step with a YeeDesignBox (d and E^n of the step's record,
    a lossy plan) equals yee_adjoint_run over that single step, bit for bit,
    on a random adjoint state (fields, psi adjoints and both accumulators)."""
    from jaxphys_amd import _native

    size, npml, n = (22, 20, 18), 4, 20
    eps, sigma, box, tab, src, mon = _setup("cpml", True, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=n,
                          record_box=True)
    prob.forward(torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu))
    sol = prob.solver
    gen = torch.Generator().manual_seed(0)
    lam0 = [torch.randn(tuple(f.shape), generator=gen).to(gpu) for f in sol.fields]
    lpsi0 = torch.randn(tuple(sol.psi.shape), generator=gen).to(gpu)
    m = n - 1
    rec = prob.box_hist[m]
    out = []
    for mode in ("run", "step"):
        lam = [t.clone() for t in lam0]
        lpsi = lpsi0.clone()
        acc_cb = torch.zeros(prob.bcells, device=gpu)
        acc_ca = torch.zeros(prob.bcells, device=gpu)
        if mode == "run":
            _native.adjoint_run(sol.plan, lam, lpsi, m, m + 1, None, None, None, None, 0.0,
                                box=prob.box_int, box_hist=prob.box_hist[m:], acc_cb=acc_cb,
                                acc_ca=acc_ca)
        else:
            b = _native.YeeDesignBox()
            b.x0, b.y0, b.z0, b.nx, b.ny, b.nz = prob.box_int
            b.d = rec.data_ptr()
            b.e_cur = rec[3 * prob.bcells:].data_ptr()
            b.acc_cb = acc_cb.data_ptr()
            b.acc_ca = acc_ca.data_ptr()
            _native.adjoint_step(sol.plan, lam, lpsi, box=b)
        torch.cuda.synchronize()
        out.append(lam + [lpsi, acc_cb, acc_ca])
    assert float(out[0][-2].abs().max()) > 0
    for a, b_ in zip(*out):
        assert torch.equal(a, b_)


def test_autograd_function_and_finite_differences(gpu):
    size, npml, n = (22, 20, 18), 4, 60
    eps, sigma, box, tab, src, mon = _setup("cpml", False, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=20)
    e = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu, requires_grad=True)
    J = FDTDObjective.apply(e, prob)
    J.backward()
    grad = e.grad.double()
    rng = np.random.default_rng(0)
    h = 1e-2
    for _ in range(4):
        c = tuple(int(rng.integers(0, s)) for s in (bx, by, bz))
        ep = e.detach().clone(); ep[c] += h  # noqa: E702
        em = e.detach().clone(); em[c] -= h  # noqa: E702
        fd = (prob.forward(ep) - prob.forward(em)) / (2 * h)
        assert float(grad[c]) == pytest.approx(fd, rel=2e-2, abs=1e-3 * float(grad.abs().max())), \
            (c, fd)


def test_design_permittivity_increases_probe_energy(gpu):
    """A few Adam steps of the inverse-design driver on the adjoint gradient
    raise the probe objective, and the iterate stays within its bounds."""
    from jaxphys_amd.optimize import design_permittivity

    size, npml, n = (22, 20, 18), 4, 60
    eps, sigma, box, tab, src, mon = _setup("cpml", False, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=20)
    e0 = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu)
    J0 = prob.forward(e0)
    res = design_permittivity(prob, e0, maximize=True, learning_rate=0.1, max_iterations=5,
                              eps_min=1.0, eps_max=4.0)
    print(f"[design] J {J0:.4e} -> {res.fun:.4e}", flush=True)
    assert res.fun > J0
    assert float(res.x.min()) >= 1.0 and float(res.x.max()) <= 4.0


def test_objective_backward_after_another_forward_raises(gpu):
    """FDTDObjective's backward uses the problem's latest forward run; a
    second apply on the same problem before the first backward would give
    the gradient at the wrong design, so it raises instead."""
    size, npml, n = (22, 20, 18), 4, 30
    eps, sigma, box, tab, src, mon = _setup("cpml", False, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    g = _gpu_grid(size, "cpml", npml, eps, sigma, tab, src)
    prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=8)
    e1 = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu,
                      dtype=torch.float64).requires_grad_(True)
    e2 = (e1.detach() + 0.5).requires_grad_(True)
    J = FDTDObjective.apply(e1, prob) + FDTDObjective.apply(e2, prob)
    with pytest.raises(RuntimeError, match="another forward"):
        J.backward()
    # one apply, then backward: the gradient of the problem's own forward
    J1 = FDTDObjective.apply(e1, prob)
    J1.backward()
    _, ref = prob.value_and_grad(e1.detach())
    assert torch.equal(e1.grad.to(ref.dtype).cpu(), ref.cpu())


@pytest.mark.parametrize("boundary,lossy,size", [
    ("cpml", False, (22, 20, 18)), ("cpml", True, (76, 26, 22)),
    ("absorbing", False, (22, 20, 18)), ("walls", True, (22, 20, 18))])
def test_pipelined_adjoint_kernels_equal_plain(gpu, boundary, lossy, size, monkeypatch):
    """The software-pipelined adjoint bodies (default) and the plain ones
    (JAXPHYS_AMD_ADJ_PIPE=0) perform the same arithmetic in the same order:
    the same J and gradient bit for bit."""
    npml, n = 4, 60
    eps, sigma, box, tab, src, mon = _setup(boundary, lossy, size, npml, n)
    x0, y0, z0, bx, by, bz = box
    out = []
    for pipe in ("1", "0"):
        monkeypatch.setenv("JAXPHYS_AMD_ADJ_PIPE", pipe)
        g = _gpu_grid(size, boundary, npml, eps, sigma, tab, src)
        prob = AdjointProblem(g, n, DT, box, [("ez", p) for _, p in mon], checkpoint_every=13)
        eps_box = torch.tensor(eps[x0:x0 + bx, y0:y0 + by, z0:z0 + bz], device=gpu)
        out.append(prob.value_and_grad(eps_box))
    (Ja, ga), (Jb, gb) = out
    assert Ja == Jb
    assert torch.equal(ga, gb), float((ga - gb).abs().max())
    assert float(ga.abs().max()) > 0


@pytest.mark.parametrize("boundary,lossy,size", [
    ("cpml", True, (22, 20, 18)), ("cpml", False, (76, 26, 22)),
    ("cpml", False, (22, 20, 18)), ("cpml", True, (76, 26, 22)), ("walls", True, (22, 20, 18)),
    ("walls", False, (22, 20, 18)), ("absorbing", False, (22, 20, 18))])
def test_adjoint_run_is_the_transpose_of_the_forward_run(gpu, boundary, lossy, size):
    """Dot-product test of the transposed step: for random fields x_0 (psi
    zero) and a random adjoint state lam_N (psi adjoints zero),
    <lam_N, S^N x_0> = <(S^T)^N lam_N, x_0> over the six fields, where S is
    one forward step (the psi components pair with zeros on both sides).
    lam_N has random E components and zero H components: an adjoint sweep
    starts there (objectives are sums over E), and yee_adjoint_run's first
    step applies the psi_H-adjoint update deferred from step N, which is
    zero only then (with a random H adjoint at step N the CPML cases differ
    by 7e-4 relative).  Float32 runs of 30 steps, products summed in
    float64."""
    from jaxphys_amd import _native

    npml, n = 4, 30
    eps, sigma, *_ = _setup(boundary, lossy, size, npml, n)
    g = EMGrid3D(size=size, resolution=1.0, boundary=boundary, pml_layers=npml)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    sol = g.build_solver(DT, n + 1)
    gen = torch.Generator().manual_seed(5)
    x0 = [torch.randn(tuple(f.shape), generator=gen).to(gpu) for f in sol.fields]
    lam_n = [torch.randn(tuple(f.shape), generator=gen).to(gpu) for f in sol.fields[:3]]
    lam_n += [torch.zeros_like(f) for f in sol.fields[3:]]
    sol.reset()
    for f, v in zip(sol.fields, x0):
        f.copy_(v)
    sol.run(0, n)
    lhs = sum(float((a.double() * b.double()).sum()) for a, b in zip(lam_n, sol.fields))
    lam = [t.clone() for t in lam_n]
    lpsi = torch.zeros_like(sol.psi) if sol.psi is not None else None
    _native.adjoint_run(sol.plan, lam, lpsi, 0, n, None, None, None, None, 0.0)
    torch.cuda.synchronize()
    rhs = sum(float((a.double() * b.double()).sum()) for a, b in zip(lam, x0))
    scale = math.sqrt(sum(float((a.double() ** 2).sum()) for a in lam_n) *
                      sum(float((b.double() ** 2).sum()) for b in x0))
    print(f"[adjoint] {boundary} lossy={lossy}: <lam, S x> {lhs:.9e}, <S^T lam, x> {rhs:.9e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)
