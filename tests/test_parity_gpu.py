"""GPU parity: libyee_hip.so through the EMGrid3D API vs the CPU oracle.

Bar: float32 results within relative L2 1e-5 of the oracle's float32
restatement (the tolerance named for this path).  Where no transcendental
function is evaluated per cell (periodic and undamped walls) the kernels use
the reference's exact float32 operation order, so the comparison there is
bit-for-bit.  The sponge evaluates expf per cell on the GPU (numpy's exp on
the CPU), which may differ in the last ulp.
"""

import math

import numpy as np
import pytest
import torch

from jaxphys_amd import DielectricRegion, EMGrid3D, PointSource3D
from oracle import ext_fdtd as X
from oracle import ref_fdtd as R

pytestmark = pytest.mark.gpu

NAMES = ["ex", "ey", "ez", "hx", "hy", "hz"]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _gpu_history(size, dx, boundary, pml, srcs, mats, t_span=None, dt=None, save_every=20,
                 n_steps=None):
    g = EMGrid3D(size=size, resolution=dx, boundary=boundary, pml_layers=pml)
    for s in srcs:
        g.add_source(PointSource3D(s.frequency, s.position, s.amplitude, s.polarization))
    for m in mats:
        g.add_material(DielectricRegion(mask=m.mask, epsilon_r=m.epsilon_r))
    kw = {}
    if t_span is not None:
        kw["t_span"] = t_span
    h = g.simulate(dt=dt, save_every=save_every, n_steps=n_steps, **kw)
    return h, g


# the setups of the reference's tests (tests/test_fdtd3d.py:16-139)
REF_SETUPS = [
    ("basic", (20, 20, 20), 0.01, "absorbing", 8, [R.Src(3e9, (10, 10, 10))], [], (0, 5e-10), 10),
    ("excite", (20, 20, 20), 0.01, "absorbing", 8, [R.Src(3e9, (10, 10, 10), 1.0)], [], (0, 3e-10), 5),
    ("six", (16, 16, 16), 0.01, "absorbing", 8, [R.Src(5e9, (8, 8, 8))], [], (0, 2e-10), 5),
    ("polx", (16, 16, 16), 0.01, "absorbing", 8, [R.Src(3e9, (8, 8, 8), 1.0, "x")], [], (0, 2e-10), 5),
    ("poly", (16, 16, 16), 0.01, "absorbing", 8, [R.Src(3e9, (8, 8, 8), 1.0, "y")], [], (0, 2e-10), 5),
    ("diel", (16, 16, 16), 0.01, "absorbing", 8, [R.Src(5e9, (8, 8, 8))], "slab10", (0, 1e-9), 20),
    ("periodic", (16, 16, 16), 0.01, "periodic", 0, [R.Src(5e9, (8, 8, 8))], [], (0, 2e-10), 5),
    ("coords", (16, 16, 16), 0.02, "absorbing", 8, [R.Src(1e9, (8, 8, 8))], [], (0, 1e-10), 5),
    ("dipole40", (40, 40, 40), 0.01, "absorbing", 6, [R.Src(3e9, (20, 20, 20))], "slab25", (0, 3e-9), 10),
]


@pytest.mark.parametrize("setup", REF_SETUPS, ids=[s[0] for s in REF_SETUPS])
def test_reference_setups_match_oracle(gpu, setup):
    name, size, dx, boundary, pml, srcs, mats, t_span, se = setup
    if mats == "slab10":
        m = np.zeros(size, bool)
        m[10:] = True
        mats = [R.Mat(m, 4.0)]
    elif mats == "slab25":
        m = np.zeros(size, bool)
        m[25:] = True
        mats = [R.Mat(m, 4.0)]
    print(f"[parity] {name} size={size} boundary={boundary}", flush=True)
    h, _ = _gpu_history(size, dx, boundary, pml, srcs, mats, t_span=t_span, save_every=se)
    dt = R.default_dt(dx, np.float32)
    t_ref, ref = R.simulate(size, dx, boundary, pml, srcs, mats, t_span=t_span, dt=dt,
                            save_every=se, dtype=np.float32)
    np.testing.assert_array_equal(h.t.cpu().numpy(), t_ref)
    for c in NAMES:
        got = getattr(h, c).cpu().numpy()
        assert got.shape == ref[c].shape
        if boundary == "periodic":
            np.testing.assert_array_equal(got, ref[c])
        else:
            # sponge: GPU expf vs NumPy exp may differ in the last ulp
            for q in range(got.shape[0]):
                assert rel_l2(got[q], ref[c][q]) <= 1e-5, (c, q)


@pytest.mark.parametrize("boundary", ["walls", "periodic"])
def test_bitwise_without_transcendentals(gpu, boundary):
    size = (70, 9, 13)  # not multiples of the 64x4 block
    rng = np.random.default_rng(7)
    mats = [R.Mat(rng.random(size) > 0.5, 3.0), R.Mat(rng.random(size) > 0.7, 7.5)]
    srcs = [R.Src(4e9, (3, 4, 5)), R.Src(6e9, (69, 0, 12), 0.5, "x"),
            R.Src(6e9, (0, 8, 0), 2.0, "y")]
    dt = R.default_dt(0.01, np.float32)
    h, _ = _gpu_history(size, 0.01, boundary, 0, srcs, mats, dt=dt, save_every=7, n_steps=40)
    _, ref = R.simulate(size, 0.01, boundary, 0, srcs, mats, dt=dt, save_every=7, n_steps=40,
                        dtype=np.float32)
    for c in NAMES:
        np.testing.assert_array_equal(getattr(h, c).cpu().numpy(), ref[c])


def test_many_sources_same_cell_order(gpu):
    """11 sources (more than the 8 fused into the E kernel), repeated cells
    accumulate in insertion order (fdtd3d.py:236-244)."""
    size = (24, 20, 18)
    srcs = [R.Src(2e9 + 1e8 * q, (5 + q % 4, 6, 7 + q % 3), 1.0 + 0.1 * q, "xyz"[q % 3])
            for q in range(11)]
    dt = R.default_dt(0.01, np.float32)
    h, _ = _gpu_history(size, 0.01, "absorbing", 4, srcs, [], dt=dt, save_every=5, n_steps=30)
    _, ref = R.simulate(size, 0.01, "absorbing", 4, srcs, [], dt=dt, save_every=5, n_steps=30,
                        dtype=np.float32)
    for c in NAMES:
        got = getattr(h, c).cpu().numpy()
        for q in range(got.shape[0]):
            assert rel_l2(got[q], ref[c][q]) <= 1e-5


def test_c1_reference_compat_64cubed_500_steps(gpu):
    """64^3, 500 steps, sponge 10 layers, eps_r=4 sphere r=12, z-dipole 8
    cells off-centre, dx=1, Courant 0.5/sqrt(3), the reference's sine
    source (wavelength 20 cells).  rel-L2 <= 1e-5 on all six fields."""
    n = 64
    size = (n, n, n)
    I, J, K = np.indices(size)
    sphere = (I - 32) ** 2 + (J - 32) ** 2 + (K - 32) ** 2 <= 144
    mats = [R.Mat(sphere, 4.0)]
    srcs = [R.Src(R.C0 / 20.0, (40, 32, 32))]
    dt = 0.5 / math.sqrt(3.0) * 1.0 / R.C0
    # save_every=499 keeps the states after steps 1 and 500 (fdtd3d.py:396-404)
    h, g = _gpu_history(size, 1.0, "absorbing", 10, srcs, mats, dt=dt, save_every=499,
                        n_steps=500)
    _, ref = R.simulate(size, 1.0, "absorbing", 10, srcs, mats, dt=dt, save_every=499,
                        n_steps=500, dtype=np.float32)
    sol = g.last_solver
    worst = 0.0
    for q, c in enumerate(NAMES):
        final = sol.reference_view(sol.fields[q]).cpu().numpy()
        np.testing.assert_array_equal(final, getattr(h, c)[-1].cpu().numpy())
        worst = max(worst, rel_l2(final, ref[c][-1]))
    print(f"[parity] C1 64^3x500 worst rel-L2 {worst:.3e}", flush=True)
    assert worst <= 1e-5


def _ext_gpu(size, dx, dt, boundary, npml, eps, sigma, srcs_spec, n_steps):
    g = EMGrid3D(size=size, resolution=dx, boundary=boundary, pml_layers=npml)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    for comp, pos, kind, kw in srcs_spec:
        g.add_source(PointSource3D(0.0, pos, kw.get("amplitude", 1.0), "xyz"[comp], waveform=kind,
                                   t0=kw.get("t0", 0.0), spread=kw.get("spread", 1.0),
                                   wavelength=kw.get("wavelength")))
    g.simulate(dt=dt, save_every=n_steps + 1, n_steps=n_steps)
    sol = g.last_solver
    return [sol.reference_view(f).cpu().numpy() for f in sol.fields]


def _ext_ref(size, dx, dt, boundary, npml, eps, sigma, srcs_spec, n_steps):
    S = R.C0 * dt / dx
    srcs = []
    for comp, pos, kind, kw in srcs_spec:
        tab = X.waveform_table(kind, n_steps, amplitude=kw.get("amplitude", 1.0),
                               t0=kw.get("t0", 0.0), spread=kw.get("spread", 1.0),
                               wavelength=kw.get("wavelength"), courant=S)
        srcs.append((comp, pos, tab))
    st = X.run(size, dx, dt, boundary, eps, srcs, n_steps, sigma=sigma, npml=npml,
               dtype=np.float32)
    return st.f


def _boxes(size, rng, nbox, lo, hi, emin, emax):
    eps = np.ones(size, np.float32)
    for _ in range(nbox):
        s = rng.integers(lo, hi + 1, 3)
        o = [rng.integers(0, max(1, size[a] - s[a])) for a in range(3)]
        eps[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]] = rng.uniform(emin, emax)
    return eps


@pytest.mark.parametrize("lossy", [False, True])
def test_cpml_matches_ext_oracle(gpu, lossy):
    size = (30, 26, 22)
    dx = 1.0
    dt = 0.5 / math.sqrt(3.0) / R.C0
    rng = np.random.default_rng(11)
    eps = _boxes(size, rng, 8, 3, 10, 1.0, 6.0)
    sigma = None
    if lossy:
        sigma = np.zeros(size)
        sigma[:, :, 12:16] = 0.02 * R.EPS0 * R.C0 / dx
    srcs = [(2, (15, 13, 11), "gaussian_derivative", dict(t0=20, spread=6)),
            (0, (8, 20, 9), "gaussian_sine", dict(t0=25, spread=8, wavelength=10, amplitude=0.5)),
            (1, (22, 7, 14), "gaussian", dict(t0=15, spread=5, amplitude=2.0))]
    n_steps = 90
    got = _ext_gpu(size, dx, dt, "cpml", 6, eps, sigma, srcs, n_steps)
    ref = _ext_ref(size, dx, dt, "cpml", 6, eps, sigma, srcs, n_steps)
    worst = max(rel_l2(a, b) for a, b in zip(got, ref))
    print(f"[parity] cpml lossy={lossy} worst rel-L2 {worst:.3e}", flush=True)
    assert worst <= 1e-6


def test_energy_fp64_matches_numpy(gpu):
    size = (33, 20, 17)
    rng = np.random.default_rng(5)
    g = EMGrid3D(size=size, resolution=0.01, boundary="walls", pml_layers=0)
    eps = (1 + 3 * rng.random(size)).astype(np.float32)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps))
    g.add_source(PointSource3D(3e9, (10, 10, 10)))
    sol = g.build_solver(1e-12, 4)
    vals = [rng.standard_normal(size).astype(np.float32) for _ in range(6)]
    for q in range(6):
        sol.reference_view(sol.fields[q]).copy_(torch.from_numpy(vals[q]))
    we, wh = sol.energy_fp64()
    e64 = [v.astype(np.float64) for v in vals]
    np.testing.assert_allclose(we, sum((eps.astype(np.float64) * v * v).sum() for v in e64[:3]),
                               rtol=1e-12)
    np.testing.assert_allclose(wh, sum((v * v).sum() for v in e64[3:]), rtol=1e-12)


def test_copy_probe(gpu):
    from jaxphys_amd import _native
    src = torch.randn(1_000_003, device=gpu)
    dst = torch.zeros_like(src)
    _native.copy_probe(src, dst)
    assert torch.equal(src, dst)


@pytest.mark.parametrize("variant,xcd,rows", [(0, True, 4), (1, True, 4), (2, True, 4),
                                               (3, None, 4), (3, False, 8), (3, True, 16),
                                               (4, None, 4), (4, None, 16), (4, 2, 4),
                                               (3, 2, 8), (3, 3, 4), (3, 4, 8), (3, 6, 4)])
@pytest.mark.parametrize("boundary,lossy,size", [
    ("absorbing", False, (70, 37, 29)), ("periodic", False, (70, 37, 29)),
    ("walls", True, (70, 37, 29)), ("cpml", False, (70, 37, 29)), ("cpml", True, (70, 37, 29)),
    # wide enough for a slab-free CPML interior (split launches of variant 2)
    ("cpml", True, (200, 41, 30)), ("cpml", False, (140, 40, 33))])
def test_fused_step_bitwise_equals_two_pass(gpu, boundary, lossy, size, variant, xcd, rows):
    """One fused step vs the H and E half-step kernels: the same arithmetic
    in the same order, so every field must agree bit for bit."""
    rng = np.random.default_rng(21)
    eps = _boxes(size, rng, 10, 3, 14, 1.0, 9.0)
    sigma = None
    if lossy:
        sigma = np.zeros(size)
        sigma[:, 10:20, :] = 0.03 * R.EPS0 * R.C0
    g = EMGrid3D(size=size, resolution=1.0, boundary=boundary, pml_layers=6)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    nx, ny, nz = size
    for q, pos in enumerate([(nx // 2, ny // 2, nz // 2), (0, 0, 0), (nx - 1, ny - 1, nz - 1),
                             (10, 30, 5)]):
        g.add_source(PointSource3D(0.0, pos, 1.0 + q, "xyz"[q % 3], waveform="gaussian_sine",
                                   t0=20, spread=8, wavelength=12))
    dt = 0.5 / math.sqrt(3.0) / R.C0
    n = 41
    a = g.build_solver(dt, n)
    b = g.build_solver(dt, n)
    # per-cell coefficient arrays on one side, the material palette on the
    # other (when the grid has few enough materials)
    b.plan.set_palette(None, None)
    a.plan.set_variant(variant)
    a.plan.set_fused_rows(rows)
    a.plan.set_tile_order(xcd)
    b.plan.set_tile_order(xcd)
    a.plan.tune(kz=0, fused_kz=7)   # several z chunks per column
    b.plan.tune(kz=5, fused_kz=-1)
    a.run_fused(0, n)        # fused (odd count: ends in the second buffer)
    b.run_two_pass(0, n)
    for q in range(6):
        assert torch.equal(a.fields[q], b.fields[q]), NAMES[q]
    if a.psi is not None:
        assert torch.equal(a.psi, b.psi)
    assert float(a.fields[2].abs().max()) > 0


@pytest.mark.parametrize("boundary,lossy", [("cpml", True), ("absorbing", False)])
def test_two_pass_palette_bitwise_equals_arrays(gpu, boundary, lossy):
    """The E kernel reading uint8 material ids and the LDS table gives the
    same bits as reading the per-cell cb / ca arrays."""
    size = (70, 37, 29)
    rng = np.random.default_rng(5)
    eps = _boxes(size, rng, 10, 3, 14, 1.0, 9.0)
    sigma = None
    if lossy:
        sigma = np.zeros(size)
        sigma[:, 10:20, :] = 0.03 * R.EPS0 * R.C0
    g = EMGrid3D(size=size, resolution=1.0, boundary=boundary, pml_layers=6)
    g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=eps, sigma=sigma))
    g.add_source(PointSource3D(0.0, (30, 18, 14), 1.0, "z", waveform="gaussian_sine",
                               t0=20, spread=8, wavelength=12))
    dt = 0.5 / math.sqrt(3.0) / R.C0
    a = g.build_solver(dt, 30)
    b = g.build_solver(dt, 30)
    assert a.mat is not None and a.palette is not None
    b.plan.set_palette(None, None)
    a.run_two_pass(0, 30)
    b.run_two_pass(0, 30)
    for q in range(6):
        assert torch.equal(a.fields[q], b.fields[q]), NAMES[q]


def test_host_snapshots_equal_device_snapshots(gpu):
    """simulate(snapshots="host") streams the saved steps into pinned host
    memory on a side stream; the values are those of the device path."""
    size = (40, 24, 20)
    srcs = [R.Src(5e9, (20, 12, 10)), R.Src(7e9, (9, 5, 14), 0.5, "x")]
    dt = R.default_dt(0.01, np.float32)
    out = []
    for mode in ("device", "host"):
        g = EMGrid3D(size=size, resolution=0.01, boundary="absorbing", pml_layers=4)
        for s in srcs:
            g.add_source(PointSource3D(s.frequency, s.position, s.amplitude, s.polarization))
        out.append(g.simulate(dt=dt, save_every=6, n_steps=40, snapshots=mode))
    dev_h, host_h = out
    assert host_h.ez.device.type == "cpu"
    assert host_h.ez.shape == dev_h.ez.shape == (7, 40, 24, 20)
    for c in NAMES:
        assert torch.equal(getattr(host_h, c), getattr(dev_h, c).cpu()), c
    assert torch.equal(host_h.t, dev_h.t.cpu())


@pytest.mark.parametrize("boundary", ["absorbing", "periodic"])
def test_host_snapshots_match_oracle(gpu, boundary):
    """simulate(snapshots="host") against the oracle's own restatement of
    fdtd3d.py:169-417 (float32): every saved step, its time stamp and the
    grid coordinates; a 40x24x20 grid, two sources, a dielectric slab,
    snapshots every 6 of 41 steps (the last save not on the final step)."""
    size = (40, 24, 20)
    srcs = [R.Src(5e9, (20, 12, 10)), R.Src(7e9, (9, 5, 14), 0.5, "x")]
    m = np.zeros(size, bool)
    m[28:] = True
    mats = [R.Mat(m, 4.0)]
    dt = R.default_dt(0.01, np.float32)
    t_span = (0.0, 41.5 * dt)
    pml = 4 if boundary == "absorbing" else 0
    g = EMGrid3D(size=size, resolution=0.01, boundary=boundary, pml_layers=pml)
    for s in srcs:
        g.add_source(PointSource3D(s.frequency, s.position, s.amplitude, s.polarization))
    g.add_material(DielectricRegion(mask=m, epsilon_r=4.0))
    h = g.simulate(t_span=t_span, dt=dt, save_every=6, snapshots="host")
    t_ref, ref = R.simulate(size, 0.01, boundary, pml, srcs, mats, t_span=t_span, dt=dt,
                            save_every=6, dtype=np.float32)
    assert h.ez.device.type == "cpu" and h.ez.shape[0] == 7
    np.testing.assert_array_equal(h.t.numpy(), t_ref)
    np.testing.assert_allclose(h.grid_x.numpy(), np.arange(40) * 0.01, rtol=1e-6)
    for c in NAMES:
        got = getattr(h, c).numpy()
        assert got.shape == ref[c].shape
        for q in range(got.shape[0]):
            if boundary == "periodic":
                np.testing.assert_array_equal(got[q], ref[c][q])
            else:
                assert rel_l2(got[q], ref[c][q]) <= 1e-5, (c, q)
