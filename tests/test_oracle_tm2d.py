"""Pins of the 2D TM oracle (oracle/ref_tm2d.py, a restatement of the
reference's src/neurosim/em/fdtd.py): brute-force per-cell loops, closed
forms, and invariants of the lossless scheme.  CPU only."""

import math

import numpy as np
import pytest

from oracle import ref_tm2d as R


def _rand_state(size, seed, dtype):
    g = np.random.default_rng(seed)
    return [g.standard_normal(size).astype(dtype) for _ in range(3)]


@pytest.mark.parametrize("boundary", ["absorbing", "periodic", "reflecting"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vectorized_matches_bruteforce_bitwise(boundary, dtype):
    size = (12, 11)
    waves = [R.Wave(4e9, 3, 0.7), R.Wave(2.5e9, 3, 1.3), R.Wave(6e9, 8)]
    walls = [R.Wall(6, 2, 5), R.Wall(9)]
    init = _rand_state(size, 1, dtype)
    dt = R.default_dt(0.01, dtype)
    _, out = R.simulate(size, 0.01, boundary, 3, waves, walls, dt=dt, n_steps=7, save_every=1,
                        dtype=dtype, init=init)
    ez, hx, hy = R.simulate_bruteforce(size, 0.01, boundary, 3, waves, walls, dt=dt, n_steps=7,
                                       dtype=dtype, init=init)
    np.testing.assert_array_equal(out["ez"][-1], ez)
    np.testing.assert_array_equal(out["hx"][-1], hx)
    np.testing.assert_array_equal(out["hy"][-1], hy)


def test_sigma_profile_closed_form():
    dx = 0.01
    s = R.sigma_1d(40, 10, dx, np.float64)
    smax = 3.2 / dx
    for i in range(10):
        v = smax * ((10 - i) / 10) ** 3
        assert s[i] == pytest.approx(v, rel=1e-15) and s[39 - i] == pytest.approx(v, rel=1e-15)
    assert np.all(s[10:30] == 0)
    # overlapping slabs: later (smaller) writes win in the middle
    s2 = R.sigma_1d(10, 8, dx, np.float64)
    # (i = 5 writes s[5] then s[4]; i = 7 writes s[7] then s[2])
    assert s2[4] == pytest.approx(smax * (3 / 8) ** 3) and s2[5] == pytest.approx(smax * (3 / 8) ** 3)
    assert s2[2] == pytest.approx(smax * (1 / 8) ** 3) and s2[0] == pytest.approx(smax)
    # float32 path: float32 sigma_max times a float32-rounded ratio
    s32 = R.sigma_1d(40, 10, dx, np.float32)
    assert s32.dtype == np.float32
    assert s32[1] == np.float32(np.float32(np.float32(3.2) / np.float32(dx)) * np.float32(0.729))
    assert R.sigma_1d(16, 0, dx, np.float64).sum() == 0


def test_default_dt_and_damping_closed_form():
    dx = 0.01
    assert R.default_dt(dx, np.float64) == 0.99 * dx / (R.C0 * math.sqrt(2.0))
    assert R.default_dt(dx, np.float32) == 0.99 * dx / (R.C0 * float(np.sqrt(np.float32(2.0))))
    dt = R.default_dt(dx, np.float64)
    d = R.damping(30, 20, 5, dx, dt, np.float64)
    assert d[15, 10] == 1.0
    sx0 = 3.2 / dx
    assert d[0, 10] == pytest.approx(math.exp(-0.5 * sx0 * dt / R.EPS0), rel=1e-12)
    assert d[0, 0] == pytest.approx(math.exp(-sx0 * dt / R.EPS0), rel=1e-12)


def test_conductor_mask_and_row_source_closed_form():
    nx, ny = 20, 16
    m = R.conductor_mask(nx, ny, [R.Wall(5, 7, 11), R.Wall(9), R.Wall(12, 3, None)])
    assert m[:, 5].sum() == nx - 4 and not m[7:11, 5].any()
    assert m[:, 9].all() and m[:, 12].all()
    assert m.sum() == (nx - 4) + nx + nx
    f = 3e9
    dt = R.default_dt(0.01, np.float64)
    _, out = R.simulate((nx, ny), 0.01, "absorbing", 4, [R.Wave(f, 8, 2.0)], dt=dt, n_steps=2,
                        save_every=1)
    assert np.all(out["ez"][0] == 0)                      # sin(0) after step 1
    row = out["ez"][1][:, 8]
    np.testing.assert_array_equal(row, np.full(nx, 2.0 * math.sin(2 * math.pi * f * dt)))
    assert np.count_nonzero(out["ez"][1]) == nx


def _interior_energy(ez, hx_prev, hx, hy_prev, hy):
    """eps0 |E^n|^2 + mu0 H^{n-1/2} . H^{n+1/2} over the periodic interior."""
    s = (slice(1, -1), slice(1, -1))
    return (R.EPS0 * float((ez[s] ** 2).sum())
            + R.MU0 * float((hx_prev[s] * hx[s]).sum() + (hy_prev[s] * hy[s]).sum()))


def test_periodic_energy_conserved_and_mode_frequency():
    """A standing Fourier mode of the (n-2)-periodic interior (the edge rows
    are the periodic copies) stays a mode and oscillates at the Yee
    frequency cos(w dt) = 1 - 2 S^2 (sin^2(kx dx/2) + sin^2(ky dx/2))."""
    nx, ny, dx = 34, 26, 0.01
    px, py = nx - 2, ny - 2
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    kx, ky = 2 * math.pi * 3 / (px * dx), 2 * math.pi * 2 / (py * dx)
    phi = np.cos(kx * (i - 1) * dx) * np.cos(ky * (j - 1) * dx)
    dt = R.default_dt(dx, np.float64)
    zero = np.zeros((nx, ny))
    _, out = R.simulate((nx, ny), dx, "periodic", 0, [R.Wave(1e9, 5, 0.0)], dt=dt, n_steps=60,
                        save_every=1, init=(phi, zero, zero))
    a = [1.0] + [float((e * phi)[1:-1, 1:-1].sum() / (phi * phi)[1:-1, 1:-1].sum())
                 for e in out["ez"]]
    for n, e in enumerate(out["ez"]):
        np.testing.assert_allclose(e[1:-1, 1:-1], a[n + 1] * phi[1:-1, 1:-1], atol=1e-12)
    S = R.C0 * dt / dx
    cw = 1 - 2 * S * S * (math.sin(kx * dx / 2) ** 2 + math.sin(ky * dx / 2) ** 2)
    for n in range(1, 59):
        assert a[n + 1] + a[n - 1] == pytest.approx(2 * cw * a[n], abs=1e-12)
    # energy of a random lossless state
    init = _rand_state((nx, ny), 4, np.float64)
    for f in init:  # make the copies consistent with the interior
        f[0, :], f[-1, :] = f[-2, :], f[1, :]
        f[:, 0], f[:, -1] = f[:, -2], f[:, 1]
    _, out = R.simulate((nx, ny), dx, "periodic", 0, [R.Wave(1e9, 5, 0.0)], dt=dt, n_steps=80,
                        save_every=1, init=init)
    # snapshot m holds (E^{m+1}, H^{m+1/2})
    w = [_interior_energy(out["ez"][n - 1], out["hx"][n - 1], out["hx"][n], out["hy"][n - 1],
                          out["hy"][n]) for n in range(1, 80)]
    assert max(abs(v - w[0]) for v in w) < 1e-12 * abs(w[0])
    # a dropped term would break it: the same sum with E^{n+1} drifts
    bad = [_interior_energy(out["ez"][n], out["hx"][n - 1], out["hx"][n], out["hy"][n - 1],
                            out["hy"][n]) for n in range(1, 80)]
    assert max(abs(v - bad[0]) for v in bad) > 1e-3 * abs(bad[0])


def test_linearity_and_reflecting_copies():
    size = (16, 14)
    dt = R.default_dt(0.01, np.float64)
    a = _rand_state(size, 7, np.float64)
    b = _rand_state(size, 8, np.float64)
    ab = [x + y for x, y in zip(a, b)]
    kw = dict(dt=dt, n_steps=9, save_every=1)
    w = [R.Wave(1e9, 4, 0.0)]
    for bnd in ("absorbing", "reflecting", "periodic"):
        _, oa = R.simulate(size, 0.01, bnd, 3, w, init=a, **kw)
        _, ob = R.simulate(size, 0.01, bnd, 3, w, init=b, **kw)
        _, oab = R.simulate(size, 0.01, bnd, 3, w, init=ab, **kw)
        for c in ("ez", "hx", "hy"):
            np.testing.assert_allclose(oab[c][-1], oa[c][-1] + ob[c][-1], rtol=0, atol=1e-9)
    _, o = R.simulate(size, 0.01, "reflecting", 3, [R.Wave(3e9, 6)], init=a, **kw)
    for c in ("ez", "hx", "hy"):
        f = o[c][-1]
        np.testing.assert_array_equal(f[0, 1:-1], f[1, 1:-1])
        np.testing.assert_array_equal(f[-1, 1:-1], f[-2, 1:-1])
        np.testing.assert_array_equal(f[:, 0], f[:, 1])
        np.testing.assert_array_equal(f[:, -1], f[:, -2])
