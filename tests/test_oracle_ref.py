"""Pins of the CPU oracle ``oracle/ref_fdtd.py`` (no GPU).

The reference's own tests pin no numbers (tests/test_fdtd3d.py:16-139 are
qualitative), so the restatement is pinned here by an independent per-cell
loop, closed forms and invariants of the Yee scheme -- each chosen so that a
plausible slip (a dropped term, a wrong sign or index, a transposed operand)
fails at least one of them.
"""

import math

import numpy as np
import pytest

from oracle import ref_fdtd as R

DX = 0.01


def _run_vec(size, boundary, pml, sources, materials, n_steps, dt, dtype=np.float64):
    t, snaps = R.simulate(size, DX, boundary, pml, sources, materials, dt=dt,
                          save_every=n_steps + 1, n_steps=n_steps, dtype=dtype)
    return t, snaps


def _final_vec(size, boundary, pml, sources, materials, n_steps, dt):
    """Final fields via the vectorised restatement."""
    eps = R.build_eps_r(size, materials, np.float64)
    ch = dt / (R.MU0 * DX)
    ce = dt / (R.EPS0 * DX)
    damp = None
    if boundary == "absorbing":
        damp = R.damping_factor(R.sigma_avg(size, pml, DX, np.float64), dt, np.float64)
    f = [np.zeros(size) for _ in range(6)]
    for n in range(n_steps):
        f = R.step(f, eps, damp, boundary, ch, ce, np.float64)
        vals, _ = R.source_values(sources, n, 0.0, dt, np.float64)
        f = R.add_sources(f, sources, vals)
    return f


@pytest.mark.parametrize("boundary,pml", [("absorbing", 3), ("periodic", 0), ("walls", 0),
                                          ("absorbing", 6)])
def test_vectorised_matches_bruteforce_fp64(boundary, pml):
    size = (9, 8, 10)
    rng = np.random.default_rng(1)
    mask = rng.random(size) > 0.6
    mats = [R.Mat(mask, 4.0), R.Mat(rng.random(size) > 0.8, 2.5)]
    srcs = [R.Src(7e9, (4, 3, 5)), R.Src(5e9, (2, 6, 7), 0.5, "x"), R.Src(9e9, (6, 2, 2), 2.0, "y"),
            R.Src(7e9, (4, 3, 5), 1.5, "z")]
    dt = R.default_dt(DX, np.float64)
    n = 12
    a = _final_vec(size, boundary, pml, srcs, mats, n, dt)
    b = R.simulate_bruteforce(size, DX, boundary, pml, srcs, mats, n, dt)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert max(np.abs(x).max() for x in a[3:]) > 0


def test_sigma_profile_closed_form():
    s = R.build_sigma_1d(20, 4, DX, np.float64)
    smax = 3.2 / DX
    assert s[0] == pytest.approx(smax) and s[19] == pytest.approx(smax)
    assert s[3] == pytest.approx(smax / 64.0) and s[16] == pytest.approx(smax / 64.0)
    assert np.all(s[4:16] == 0)
    # overlapping slabs (pml > n/2): later mirror writes win, fdtd3d.py:153-156
    v = [smax * ((6 - i) / 6) ** 3 for i in range(6)]
    s = R.build_sigma_1d(8, 6, DX, np.float64)
    np.testing.assert_allclose(s, [v[0], v[1], v[5], v[4], v[4], v[5], v[1], v[0]])
    # damping exponent at the outermost cell underflows to 0 in float32
    d = R.damping_factor(R.sigma_avg((32, 32, 32), 8, DX, np.float32),
                         R.default_dt(DX, np.float32), np.float32)
    assert d[0, 0, 0] == 0.0 and d[16, 16, 16] == 1.0 and 0 < d[7, 16, 16] < 1


def test_dt_and_step_count():
    dt64 = R.default_dt(DX, np.float64)
    assert dt64 == pytest.approx(1.906575e-11, rel=1e-6)
    assert R.n_steps_for((0, 5e-10), dt64) == 26
    assert R.n_steps_for((0, 3e-10), dt64) == 15
    assert R.n_steps_for((0, 1e-9), dt64) == 52
    dt32 = R.default_dt(DX, np.float32)
    assert dt32 != dt64 and abs(dt32 / dt64 - 1) < 1e-7


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_source_value_closed_form(dtype):
    size = (12, 12, 12)
    src = R.Src(3e9, (5, 6, 7), 1.7, "z")
    dt = R.default_dt(DX, dtype)
    t, s = R.simulate(size, DX, "absorbing", 3, [src], [], dt=dt, save_every=1,
                      n_steps=3, dtype=dtype)
    assert np.all(s["ez"][0] == 0)  # step 0 adds A sin(0)
    e1 = s["ez"][1]
    expect = dtype(dtype(1.7) * np.sin(dtype(dtype(2 * math.pi * 3e9) * dtype(dt))))
    assert e1[5, 6, 7] == expect
    e1[5, 6, 7] = 0
    assert np.all(e1 == 0) and np.all(s["hx"][1] == 0)
    assert np.count_nonzero(s["hx"][2]) > 0  # the field starts to spread
    # t = t_c + dt
    assert t[0] == dtype(dt) and t[2] == dtype(dtype(dtype(2) * dtype(dt)) + dtype(dt))


def _rand_fields(size, rng):
    return [rng.standard_normal(size) for _ in range(6)]


def test_periodic_energy_conserved_fp64():
    """W = sum eps E^n.E^n / ce + sum H^{n-1/2}.H^{n+1/2} / ch is exactly
    conserved by the lossless periodic leapfrog (Taflove & Hagness ch. 3)."""
    size = (10, 12, 9)
    rng = np.random.default_rng(0)
    eps = 1.0 + 3.0 * rng.random(size)
    dt = R.default_dt(DX, np.float64)
    ch = dt / (R.MU0 * DX)
    ce = dt / (R.EPS0 * DX)
    f = _rand_fields(size, rng)
    ws = []
    for _ in range(40):
        h_old = [a.copy() for a in f[3:]]
        g = R.step(f, eps, None, "periodic", ch, ce, np.float64)
        w = (sum((eps * a * a).sum() for a in f[:3]) / ce
             + sum((x * y).sum() for x, y in zip(h_old, g[3:])) / ch)
        ws.append(w)
        f = g
    ws = np.array(ws)
    assert np.max(np.abs(ws / ws[0] - 1)) < 1e-11
    # a sign slip in either curl breaks conservation
    bad = [a.copy() for a in _rand_fields(size, np.random.default_rng(0))]
    e2 = []
    for _ in range(10):
        h_old = [a.copy() for a in bad[3:]]
        g = R.step(bad, eps, None, "periodic", -ch, ce, np.float64)
        e2.append(sum((eps * a * a).sum() for a in bad[:3]) / ce
                  + sum((x * y).sum() for x, y in zip(h_old, g[3:])) / ch)
        bad = g
    assert abs(e2[-1] / e2[0] - 1) > 1e-3


def _div_back(eps, ex, ey, ez):
    return ((eps * ex - np.roll(eps * ex, 1, 0)) + (eps * ey - np.roll(eps * ey, 1, 1))
            + (eps * ez - np.roll(eps * ez, 1, 2)))


def _div_fwd(hx, hy, hz):
    return ((np.roll(hx, -1, 0) - hx) + (np.roll(hy, -1, 1) - hy) + (np.roll(hz, -1, 2) - hz))


def test_periodic_divergence_preserved():
    size = (8, 10, 12)
    rng = np.random.default_rng(3)
    eps = 1.0 + rng.random(size)
    dt = R.default_dt(DX, np.float64)
    ch, ce = dt / (R.MU0 * DX), dt / (R.EPS0 * DX)
    f = _rand_fields(size, rng)
    d_e0, d_h0 = _div_back(eps, *f[:3]), _div_fwd(*f[3:])
    for _ in range(25):
        f = R.step(f, eps, None, "periodic", ch, ce, np.float64)
    np.testing.assert_allclose(_div_back(eps, *f[:3]), d_e0, atol=1e-9)
    np.testing.assert_allclose(_div_fwd(*f[3:]), d_h0, atol=1e-9)


@pytest.mark.parametrize("eps_r,m", [(1.0, 1), (1.0, 3), (4.0, 2)])
def test_yee_dispersion_single_mode(eps_r, m):
    """Ez = cos(k i) along x is a 2-D invariant subspace; its samples obey
    a_{n+1} + a_{n-1} = 2 cos(w dt) a_n with sin^2(w dt/2) = (S^2/eps_r)
    sin^2(k dx/2), S = c dt/dx (Taflove & Hagness ch. 4)."""
    n = 16
    size = (n, 4, 4)
    k = 2 * math.pi * m / n
    dt = R.default_dt(DX, np.float64)
    S = R.C0 * dt / DX
    ch, ce = dt / (R.MU0 * DX), dt / (R.EPS0 * DX)
    eps = np.full(size, eps_r)
    f = [np.zeros(size) for _ in range(6)]
    f[2][:] = np.cos(k * np.arange(n))[:, None, None]
    a = []
    for _ in range(30):
        f = R.step(f, eps, None, "periodic", ch, ce, np.float64)
        a.append(f[2][0, 0, 0])
    a = np.array(a)
    cos_wdt = 1 - 2 * (S * S / eps_r) * math.sin(k / 2) ** 2
    np.testing.assert_allclose(a[2:] + a[:-2], 2 * cos_wdt * a[1:-1], atol=1e-12)


def test_linearity_in_sources():
    size = (10, 10, 10)
    dt = R.default_dt(DX, np.float64)
    s1 = R.Src(4e9, (3, 4, 5))
    s2 = R.Src(6e9, (6, 5, 4), 0.7, "x")
    mats = [R.Mat(np.indices(size)[0] >= 6, 3.0)]
    a = _final_vec(size, "absorbing", 2, [s1], mats, 20, dt)
    b = _final_vec(size, "absorbing", 2, [s2], mats, 20, dt)
    c = _final_vec(size, "absorbing", 2, [s1, s2], mats, 20, dt)
    for x, y, z in zip(a, b, c):
        np.testing.assert_allclose(x + y, z, atol=1e-12 * max(1.0, np.abs(z).max()))


def test_fp32_restatement_close_to_fp64():
    size = (16, 16, 16)
    src = [R.Src(5e9, (8, 8, 8))]
    mats = [R.Mat(np.indices(size)[0] >= 10, 4.0)]
    dt = R.default_dt(DX, np.float64)
    _, a = R.simulate(size, DX, "absorbing", 8, src, mats, t_span=(0, 1e-9), dt=dt,
                      save_every=20, dtype=np.float64)
    _, b = R.simulate(size, DX, "absorbing", 8, src, mats, t_span=(0, 1e-9), dt=dt,
                      save_every=20, dtype=np.float32)
    rel = np.linalg.norm(a["ez"][-1] - b["ez"][-1]) / np.linalg.norm(a["ez"][-1])
    assert rel < 1e-4


def _step_old_h(f, eps, ch, ce):
    """A planted mistake (test only): the E update of a periodic step
    reads the old H instead of the new one."""
    ex, ey, ez, hx, hy, hz = f
    cx, cy, cz = R._curl_e_periodic(ex, ey, ez)
    ax, ay, az = R._curl_h_periodic(hx, hy, hz)
    cb = ce / eps
    return [ex + cb * ax, ey + cb * ay, ez + cb * az, hx - ch * cx, hy - ch * cy, hz - ch * cz]


def _reversal_error(step, size, n, seed=2):
    """Run n steps from random (E^0, H^-1/2), then n steps from
    (E^n, -H^{n+1/2}): a leapfrog scheme (H from the old E, then E from the
    new H) returns exactly to (E^0, -H^{1/2})."""
    rng = np.random.default_rng(seed)
    eps = 1.0 + 3.0 * rng.random(size)
    f0 = _rand_fields(size, rng)
    f = f0
    for _ in range(n):
        f = step(f, eps)
    h_next = step(f, eps)[3:]
    r = f[:3] + [-h for h in h_next]
    for _ in range(n):
        r = step(r, eps)
    want = f0[:3] + [-h for h in step(f0, eps)[3:]]
    scale = max(np.abs(a).max() for a in want)
    return max(np.abs(a - b).max() for a, b in zip(r, want)) / scale


def test_periodic_time_reversal_fp64():
    """Time reversal of the Yee leapfrog (fdtd3d.py:290-301: H from E^n,
    then E from the new H): 60 steps forward and 60 back from the
    H-reversed state recover the start to rounding; a step whose E update
    reads the old H does not."""
    size = (10, 12, 9)
    dt = R.default_dt(DX, np.float64)
    ch, ce = dt / (R.MU0 * DX), dt / (R.EPS0 * DX)
    good = _reversal_error(lambda f, eps: R.step(f, eps, None, "periodic", ch, ce, np.float64),
                           size, 60)
    assert good < 1e-11
    bad = _reversal_error(lambda f, eps: _step_old_h(f, eps, ch, ce), size, 60)
    assert bad > 1e-3


def _mirror_pairs(n, c):
    """Index maps of the x (or y) mirror about node c: integer-position
    nodes i -> 2c - i, half-integer ones i -> 2c - 1 - i (periodic)."""
    i = np.arange(n)
    return (2 * c - i) % n, (2 * c - 1 - i) % n


def _mirror_error(step, n=16, nz=10, steps=25):
    """A blob of Ez centred on node (c, c), in a medium layered along z,
    stays symmetric: under x -> 2c - x, Ey, Ez, Hx are even and Ex, Hy, Hz
    odd (E polar, H axial); under y -> 2c - y, Ex, Ez, Hy even and Ey, Hx,
    Hz odd.  (The medium varies along z only: eps_r of a cell index serves
    all three E components, which sit at different positions, so an eps
    that varies along x or y is not mirror-symmetric for all of them.)"""
    c = n // 2
    rng = np.random.default_rng(7)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dx, dy = np.minimum(abs(ii - c), n - abs(ii - c)), np.minimum(abs(jj - c), n - abs(jj - c))
    eps = np.broadcast_to((1.0 + 3.0 * rng.random(nz))[None, None, :], (n, n, nz)).copy()
    f = [np.zeros((n, n, nz)) for _ in range(6)]
    f[2] = np.exp(-0.3 * (dx ** 2 + dy ** 2))[:, :, None] * rng.standard_normal(nz)[None, None, :]
    for _ in range(steps):
        f = step(f, eps)
    ie, ih = _mirror_pairs(n, c)
    x_even = {1: ie, 2: ie, 3: ie}
    y_even = {0: ie, 2: ie, 4: ie}
    err = 0.0
    scale = max(np.abs(a).max() for a in f)
    for comp, a in enumerate(f):
        mx = a[x_even[comp]] if comp in x_even else -a[ih]
        my = a[:, y_even[comp]] if comp in y_even else -a[:, ih]
        err = max(err, np.abs(a - mx).max(), np.abs(a - my).max())
    return err / scale


def test_periodic_mirror_symmetry_fp64():
    """Mirror symmetry of a centred source (the build's oracle pin for the
    update's index offsets): 25 steps keep every component even or odd
    under the x and y mirrors to rounding; moving one difference of the E
    update by a cell (Hy(i+1) - Hy(i) instead of Hy(i) - Hy(i-1)) breaks
    it."""
    dt = R.default_dt(DX, np.float64)
    ch, ce = dt / (R.MU0 * DX), dt / (R.EPS0 * DX)
    good = _mirror_error(lambda f, eps: R.step(f, eps, None, "periodic", ch, ce, np.float64))
    assert good < 1e-12

    def shifted(f, eps):
        g = R.step(f, eps, None, "periodic", ch, ce, np.float64)
        # redo Ez with the Hy difference one cell off
        hx, hy = g[3], g[4]
        cz = (np.roll(hy, -1, 0) - hy) - (hx - np.roll(hx, 1, 1))
        g[2] = f[2] + (ce / eps) * cz
        return g

    assert _mirror_error(shifted) > 1e-3
