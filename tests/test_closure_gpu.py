"""The learned closure end to end at a small size: fine/coarse pairs from the
FDTD solver, training on the BF16-MFMA kernels, and the gain over the plain
coarse grid on held-out seeds (fine FDTD is the judge)."""

import numpy as np
import pytest
import torch

from jaxphys_amd import closure as CL

pytestmark = pytest.mark.gpu


def test_restricted_fine_start_matches_coarse_start(gpu):
    """Before any step both grids are zero; after the first coarse interval
    the restricted fine state and the coarse state agree to first order
    (same pulse, same medium): the relative difference is well below 1."""
    prob = CL.ClosureProblem(seed=5, n_fine=64, npml_fine=8, device=gpu)
    states, _ = prob.fine_states(40, 10)
    sol = CL.coarse_rollout(prob, None, 40, 10)
    err = CL.rel_err(sol.fields, states[-1], margin=prob.npml_c)
    print(f"[closure] plain coarse vs restricted fine after 40 coarse steps: {err:.3f}", flush=True)
    assert 0.0 < err < 0.8


def test_closure_training_beats_plain_coarse(gpu):
    kw = dict(npml_fine=8)
    net, hist = CL.train_closure(gpu, seeds=list(range(6)), n_fine=64, n_coarse=100, every=10,
                                 epochs=3, crop=32, lr=2e-3, **kw)
    # the least-squares source model was fitted (the crops held whole boxes)
    assert float(net.ks.abs().max()) > 0
    assert np.isfinite(hist).all()
    assert np.mean(hist[-20:]) < np.mean(hist[:20])
    res = [CL.evaluate_closure(net, s, n_fine=64, n_coarse=100, every=10, **kw)
           for s in (1000, 1001)]
    for r in res:
        print(f"[closure] seed {r['seed']}: coarse {r['err_coarse']:.4f} "
              f"closure {r['err_closure']:.4f}", flush=True)
    assert np.mean([r["err_closure"] for r in res]) < np.mean([r["err_coarse"] for r in res])


def test_far_field_operator_reduces_held_out_residual(gpu):
    """The least-squares far-field operator, fitted on four seeds, lowers
    the one-interval residual (restricted fine state minus the coarse state
    run from the restricted fine state) of a held-out seed away from the
    source, and passes its own long-run stability check."""
    n_fine, n_coarse, every, kw = 64, 100, 10, dict(npml_fine=8)
    net = CL.ClosureNet(gpu, seed=0)
    prob0 = CL.ClosureProblem(0, n_fine, device=gpu, **kw)
    states0, _ = prob0.fine_states(n_coarse, every)
    CL._init_net(net, prob0, states0,
                 prob0.coarse_grid().build_solver(prob0.dt_coarse(), n_coarse + every + 1), every)
    CL.fit_far_field(net, gpu, [0, 1, 2, 3], n_fine, n_coarse, every, **kw)
    assert net.far_gain > 0
    prob = CL.ClosureProblem(1000, n_fine, device=gpu, **kw)
    states, _ = prob.fine_states(n_coarse, every)
    csol = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + every + 1)
    cells = CL._far_cells(prob, net)
    n = csol.nx * csol.ny * csol.nz
    y = torch.zeros(n, CL.OUT_STRIDE, device=gpu)
    before = after = 0.0
    for q in range(len(states) - 1):
        csol.reset()
        for f, v in zip(csol.fields, states[q]):
            f.copy_(v)
        csol.run(q * every, every)
        y.zero_()
        net.add_far_field(y, csol)
        for c in range(6):
            r = ((states[q + 1][c] - csol.fields[c]) / net.scale[c]).reshape(-1)[cells].double()
            before += float((r ** 2).sum())
            after += float(((r - y[cells, c].double()) ** 2).sum())
    print(f"[closure] held-out far residual with the operator: {after / before:.3f} of without",
          flush=True)
    assert after < 0.97 * before


def test_near_source_patch_reproduces_fine_run_in_source_box(gpu):
    """When the global grids are no larger than the patch's local grids
    (coarse 32^3 = 2 x half 16, the same CPML), the local pair is the global
    pair: the coarse run corrected by the patch alone then equals the
    restricted fine run in the source box after every interval (to the
    rounding of one addition), while the plain coarse run does not."""
    n_fine, n_coarse, every, kw = 64, 100, 10, dict(npml_fine=8)
    net = CL.ClosureNet(gpu, seed=0)
    net.far_gain, net.net_gain = 0.0, 0.0
    prob = CL.ClosureProblem(3, n_fine, device=gpu, **kw)
    states, _ = prob.fine_states(n_coarse, every)
    sol = CL.coarse_solver(prob, net, n_coarse, every)
    plain = prob.coarse_grid().build_solver(prob.dt_coarse(), n_coarse + every + 1)
    patch = sol._closure_patch
    assert len(patch.boxes) == 1 and patch.corr[0].shape[0] == n_coarse // every
    box = patch.boxes[0]
    worst = worst_plain = 0.0
    for q, m in enumerate(range(0, n_coarse, every)):
        sol.run(m, every)
        net.correct(sol, 1.0, m, every, use_net=False, use_far=False)
        plain.run(m, every)
        for c in range(6):
            ref = states[q + 1][c][box]
            scale = float(ref.abs().max()) + 1e-30
            worst = max(worst, float((sol.fields[c][box] - ref).abs().max()) / scale)
            worst_plain = max(worst_plain, float((plain.fields[c][box] - ref).abs().max()) / scale)
    print(f"[closure] patch box vs restricted fine: {worst:.2e} (plain coarse {worst_plain:.2e})",
          flush=True)
    assert worst <= 1e-5
    assert worst_plain > 0.1


def test_near_source_patch_overlapped_equals_sequential(gpu):
    """The patch's local coarse runs go on a second stream beside the fine
    runs; the corrections must be those of running everything in order on
    one stream, bit for bit, for every interval."""
    prob = CL.ClosureProblem(5, 96, npml_fine=16, device=gpu)
    g = prob.coarse_grid()
    dt = prob.dt_coarse()
    a = CL.NearSourcePatch(g, dt, 120, 10, half=16, device=gpu, tail=2)
    b = CL.NearSourcePatch(g, dt, 120, 10, half=16, device=gpu, tail=2, overlap=False)
    assert len(a.corr) == len(b.corr) == 1 and a.corr[0].shape[0] >= 2
    assert torch.equal(a.corr[0], b.corr[0])
    assert float(a.corr[0].abs().max()) > 0.0


def test_near_source_patch_beats_plain_coarse_far_from_source(gpu):
    """With the near-source patch (and nothing else) the coarse run's field
    error away from the source (more than 3 cells) on held-out seeds is
    well below the plain coarse run's, at a size where the local grids are
    smaller than the domain (coarse 48^3, local 32^3)."""
    n_fine, n_coarse, every, kw = 96, 120, 10, dict(npml_fine=8)
    net = CL.ClosureNet(gpu, seed=0)
    net.far_gain, net.net_gain = 0.0, 0.0
    for seed in (1000, 1005):
        prob = CL.ClosureProblem(seed, n_fine, device=gpu, **kw)
        states, _ = prob.fine_states(n_coarse, every)
        far = CL.far_mask(prob.nc, prob.npml_c, prob.src_c, 3.0, gpu)
        errs = {}
        for label, model in (("plain", None), ("patch", net)):
            sol = CL.coarse_rollout(prob, model, n_coarse, every)
            a, b = CL._err_terms(sol.fields, states[-1], prob.npml_c, far)
            errs[label] = (a / b) ** 0.5
        print(f"[closure] seed {seed} far-field error: patch {errs['patch']:.4f}, "
              f"plain {errs['plain']:.4f}", flush=True)
        assert errs["patch"] < 0.5 * errs["plain"]


def test_patch_kernels_match_torch_restriction(gpu):
    """conv3d_patch_restrict_h / conv3d_patch_correct give the same bits as
    the torch restriction (restrict_e, restrict_h, the mean of two H
    restrictions) on the box, and touch no coarse cell outside it."""
    from jaxphys_amd import _conv

    g = torch.Generator(device="cpu").manual_seed(4)
    nzf, nyf, nxf = 40, 36, 32
    side, org = 7, (5, 4, 3)                                  # origin (x, y, z)

    def rnd(*shape):
        return torch.randn(*shape, generator=g).to(gpu)

    e = [rnd(nzf, nyf, nxf) for _ in range(3)]
    h0 = [rnd(nzf, nyf, nxf) for _ in range(3)]
    h1 = [rnd(nzf, nyf, nxf) for _ in range(3)]
    coarse = [rnd(nzf // 2, nyf // 2, nxf // 2) for _ in range(6)]
    before = [c.clone() for c in coarse]
    hprev = torch.empty(3, side, side, side, device=gpu)
    corr = torch.empty(6, side, side, side, device=gpu)
    fshape, cshape = (nxf, nyf, nzf), (nxf // 2, nyf // 2, nzf // 2)
    _conv.patch_restrict_h(h0, fshape, org, side, hprev)
    box = (slice(org[2], org[2] + side), slice(org[1], org[1] + side),
           slice(org[0], org[0] + side))
    rh0 = CL.restrict_h(h0)
    for c in range(3):
        assert torch.equal(hprev[c], rh0[c][box])
    _conv.patch_correct(e, h1, hprev, fshape, coarse, cshape, org, side, corr)
    st = CL.restrict_e(e) + [0.5 * (a + b) for a, b in zip(rh0, CL.restrict_h(h1))]
    for c in range(6):
        d = st[c][box] - before[c][box]
        assert torch.equal(corr[c], d)
        want = before[c].clone()
        want[box] = before[c][box] + d
        assert torch.equal(coarse[c], want)


def test_near_source_patch_general_source(gpu):
    """The patch for any point source of a coarse CPML grid: an x-polarised
    gaussian_sine source in a random lossy dielectric, on a coarse grid no
    larger than the local grids.  The fine grid is built the way the patch
    builds its twin (permittivity and conductivity repeated 2x2x2, the two
    fine sources of _fine_twin), so the patched coarse run must equal its
    restriction in the source box after every interval."""
    from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D

    n, every, n_coarse = 32, 10, 80
    rng = np.random.default_rng(11)
    eps = np.ones((n, n, n), np.float32)
    eps[10:20, 12:25, 8:18] = 3.0
    eps[16:28, 5:14, 14:30] = 1.7
    sigma = np.zeros((n, n, n))
    sigma[18:26, 15:24, 10:20] = 5e-4   # sigma dt / (2 eps0) about 0.05
    src = PointSource3D(2e9, (15, 16, 17), 0.3, "x", waveform="gaussian_sine", t0=20, spread=6,
                        wavelength=8.0)
    coarse = EMGrid3D((n, n, n), 2.0, "cpml", 4, device=gpu)
    coarse.add_material(DielectricRegion(mask=np.ones((n,) * 3, bool), epsilon_r=eps,
                                         sigma=sigma))
    coarse.add_source(src)
    fine = EMGrid3D((2 * n,) * 3, 1.0, "cpml", 8, device=gpu)
    rep = lambda a: np.repeat(np.repeat(np.repeat(a, 2, 0), 2, 1), 2, 2)
    fine.add_material(DielectricRegion(mask=np.ones((2 * n,) * 3, bool), epsilon_r=rep(eps),
                                       sigma=rep(sigma)))
    for t in CL._fine_twin(src, src.position):
        fine.add_source(t)
    dt = CL.COURANT * 2.0 / CL.C0
    fsol = fine.build_solver(0.5 * dt, 2 * n_coarse + 2)
    states = list(CL.restricted_states(fsol, n_coarse, every))
    net = CL.ClosureNet(gpu, seed=0)
    net.far_gain, net.net_gain, net.patch_tail = 0.0, 0.0, -1
    sol = coarse.build_solver(dt, n_coarse + every + 1)
    patch = CL.attach_patch(sol, coarse, n_coarse, every, half=16)
    box = patch.boxes[0]
    worst = 0.0
    for q, m in enumerate(range(0, n_coarse, every)):
        sol.run(m, every)
        net.correct(sol, 1.0, m, every, use_net=False, use_far=False)
        for c in range(6):
            ref = states[q][c][box]
            scale = float(ref.abs().max()) + 1e-30
            worst = max(worst, float((sol.fields[c][box] - ref).abs().max()) / scale)
    print(f"[closure] general source: patch box vs restricted fine {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_far_field_transpose_on_the_kernel(gpu):
    """conv3d_stencil_f32 with the transpose weights is the adjoint of the
    far-field operator on the GPU: <T f, u> = <f, T^T u> over the six
    channels, to float32 rounding."""
    torch.manual_seed(4)
    n = 24
    shape = (n, n, n)
    net = CL.ClosureNet(gpu, seed=0)
    net.far_w = torch.randn(27 * 6 * 2 * 6, device=gpu)
    net.scale = torch.tensor([0.5, 1.5, 2.0, 3e-3, 4e-3, 1e-3], device=gpu)
    f = [torch.randn(shape, device=gpu) * float(net.scale[c]) for c in range(6)]
    eps = 1.0 + 3.0 * torch.rand(shape, device=gpu)
    u = torch.randn(n ** 3, CL.OUT_STRIDE, device=gpu)
    y = torch.zeros(n ** 3, CL.OUT_STRIDE, device=gpu)
    from jaxphys_amd import _conv
    _conv.stencil_f32(shape, f, eps, net.scale, net.far_w, y, CL.OUT_STRIDE)
    lhs = float((y[:, :6].double() * u[:, :6].double()).sum())
    v0, v1 = CL.far_field_transpose_weights(net)
    ones = torch.ones(6, device=gpu)
    g = torch.zeros(n ** 3, CL.OUT_STRIDE, device=gpu)
    e = (eps - 1.0).reshape(-1)
    _conv.stencil_f32(shape, [u[:, c].reshape(shape).contiguous() for c in range(6)], eps, ones,
                      v0, g, CL.OUT_STRIDE)
    _conv.stencil_f32(shape, [(u[:, c] * e).reshape(shape).contiguous() for c in range(6)], eps,
                      ones, v1, g, CL.OUT_STRIDE)
    rhs = sum(float((g[:, c].double() * f[c].reshape(-1).double()).sum()) for c in range(6))
    print(f"[closure] transpose identity: {lhs:.6e} vs {rhs:.6e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_rollout_gradient_matches_finite_differences(gpu):
    """The gradient of the rollout loss (back-propagated through the coarse
    Yee steps with the transposed step, through the far-field operator's
    transpose and into the net) against central differences of the loss
    along the gradient's direction in the output-layer weights, on a
    64^3 / 32^3 problem with the committed closure's linear parts and a
    small random output layer (where the net's own dependence on the state,
    which the gradient leaves out, is negligible).  Without the far-field
    transpose the gradient is off by about 1 %; with it, 0.1 %."""
    net = CL.ClosureNet.load(CL.__file__.replace("closure.py", "data/closure_c5.safetensors"),
                             gpu)
    CL._reinit_net_weights(net, 5)
    g_ = torch.Generator(device="cpu").manual_seed(7)
    w3 = torch.randn(net.params["w3"].shape, generator=g_) * 2e-6
    w3[:, :, 32:] = 0
    net.params["w3"] = w3.to(gpu)
    net.net_gain = 1.0
    every, n_coarse = 10, 200
    prob = CL.ClosureProblem(3, 128, npml_fine=16, device=gpu)
    states, _ = prob.fine_states(n_coarse, every)
    csol = CL.coarse_solver(prob, net, n_coarse, every)
    wm, cm, box, den, wc = CL.rollout_setup(net, prob, csol, states, every)
    assert float(cm.sum()) > 0
    far_t = CL.far_field_transpose_weights(net)
    Q = n_coarse // every
    csol.reset()
    loss, g = CL._rollout_window(net, csol, states, 0, Q, every, wm, cm, box, den, wc, far_t)
    base = {k: v.clone() for k, v in net.params.items()}
    D = g["w3"] / g["w3"].norm()
    h = 0.1 * float(base["w3"].norm())
    ls = []
    for s in (1, -1):
        net.params = {k: v.clone() for k, v in base.items()}
        net.params["w3"] = base["w3"] + s * h * D
        net._packed = None
        net.version += 1
        csol.reset()
        ls.append(CL._rollout_window(net, csol, states, 0, Q, every, wm, cm, box, den, wc, None,
                                     grads=False)[0])
    fd = (ls[0] - ls[1]) / (2 * h)
    an = float((g["w3"] * D).sum())
    print(f"[closure] rollout loss {loss:.4e}: directional derivative {an:.6e}, central "
          f"difference {fd:.6e}", flush=True)
    assert abs(fd - an) <= 3e-3 * abs(an)


def test_net_on_the_corrected_box_equals_the_whole_grid(gpu):
    """The closure evaluates its net only on the cells it corrects plus the
    net's receptive field; there the correction equals the net evaluated
    on the whole grid bit for bit, and it is zero elsewhere."""
    net = CL.ClosureNet.load(CL.__file__.replace("closure.py", "data/closure_c5.safetensors"),
                             gpu)
    prob = CL.ClosureProblem(7, 96, npml_fine=16, device=gpu)
    sol = prob.coarse_grid().build_solver(prob.dt_coarse(), 40)
    g_ = torch.Generator(device="cpu").manual_seed(3)
    for c, f in enumerate(sol.fields):
        f.copy_(torch.randn(f.shape, generator=g_) * float(net.scale[c]))
    got = net.net_correction(sol, 0, 10).clone()
    n = sol.nx
    x = net.pack_fields(sol, 0, 10).clone()
    full = net.forward(x, (n, n, n)).clone()
    zero = net.forward(net.pack_zero(sol), (n, n, n)).clone()
    want = (full - zero).view(n, n, n, CL.OUT_STRIDE)
    mg = net.margin
    inner = (slice(mg, n - mg),) * 3
    # channels 0-5 (the output layer does not write its padding channels)
    g4 = got.view(n, n, n, CL.OUT_STRIDE)[..., :CL.COUT]
    want = want[..., :CL.COUT]
    assert mg > 0 and n - 2 * mg > 4
    assert torch.equal(g4[inner], want[inner])
    assert float(g4[inner].abs().max()) > 0.0
    mask = torch.ones(n, n, n, dtype=torch.bool, device=gpu)
    mask[inner] = False
    assert float(g4[mask].abs().max()) == 0.0


def test_train_rollout_runs_and_updates_the_net(gpu):
    """train_rollout on a 128^3 / 64^3 problem: finite losses, the output
    layer leaves its zero start, the linear parts stay as loaded, and the
    validation step sets a gain from its candidates."""
    net = CL.ClosureNet.load(CL.__file__.replace("closure.py", "data/closure_c5.safetensors"),
                             gpu)
    far_w = net.far_w.clone()
    hist = CL.train_rollout(net, gpu, seeds=[3, 4], n_fine=128, n_coarse=100, every=10,
                            epochs=1, window=5, lr=6e-5, val_seeds=[905], npml_fine=16)
    print(f"[closure] rollout losses {[round(h, 5) for h in hist]}, gain {net.net_gain}",
          flush=True)
    assert len(hist) == 4 and all(np.isfinite(hist))
    assert float(net.params["w3"].abs().max()) > 0.0
    assert torch.equal(net.far_w, far_w)
    assert net.net_gain in (0.0, 0.25, 0.5, 1.0)


@pytest.mark.parametrize("net_in_box,net_gain,use_far", [(False, 1.0, True), (True, 0.5, True),
                                                         (False, 1.0, False)])
def test_fused_correction_equals_the_unfused_one(gpu, net_in_box, net_gain, use_far):
    """``correct`` (the net on its box, the far-field operator into its own
    buffer, one apply kernel) changes the fields exactly as
    ``_correct_unfused`` (the full-grid correction tensor built by separate
    passes), interval after interval of a corrected rollout with the
    near-source patch."""
    net = CL.ClosureNet.load(CL.__file__.replace("closure.py", "data/closure_c5.safetensors"),
                             gpu)
    assert net.forcing and net.source_patch
    net.net_in_box, net.net_gain = net_in_box, net_gain
    prob = CL.ClosureProblem(7, 96, npml_fine=16, device=gpu)
    every, n_coarse = 10, 40
    a = CL.coarse_solver(prob, net, n_coarse, every)
    b = CL.coarse_solver(prob, net, n_coarse, every, patch=a._closure_patch)
    for m in range(0, n_coarse, every):
        a.run(m, every)
        b.run(m, every)
        net.correct(a, 1.0, m, every, True, use_far)
        net._correct_unfused(b, 1.0, m, every, True, use_far)
        for fa, fb in zip(a.fields, b.fields):
            assert torch.equal(fa, fb), m
    assert float(a.fields[2].abs().max()) > 0.0


def test_patch_add_equals_torch_add(gpu):
    """``conv3d_patch_add`` (one launch for the six fields of a box) gives
    the bits of torch's ``add_(corr, alpha=gain)`` on each field's box."""
    from jaxphys_amd import _conv
    g_ = torch.Generator(device="cpu").manual_seed(11)
    n = (20, 18, 16)
    fields = [torch.randn(n[2], n[1], n[0], generator=g_).to(gpu) for _ in range(6)]
    ref = [f.clone() for f in fields]
    corr = torch.randn(6, 5, 6, 7, generator=g_).to(gpu)
    org, gain = (3, 4, 5), 0.7
    _conv.patch_add(n, fields, corr, org, (7, 6, 5), gain)
    box = (slice(5, 10), slice(4, 10), slice(3, 10))
    for k in range(6):
        ref[k][box].add_(corr[k], alpha=gain)
    for f, r in zip(fields, ref):
        assert torch.equal(f, r)
