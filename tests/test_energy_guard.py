"""The float64 energy guard of ``EMGrid3D.simulate(energy_every=...)``.

``stability_energy_bound`` is checked on the CPU against the float64 C
oracle: stable runs (walls, periodic, sponge, CPML; sine and pulse sources;
Courant numbers up to 0.999 of the 3D limit) stay below it with room to
spare, and runs with dt above the 3D CFL limit cross it within a few tens
of steps, long before the fields overflow.  The GPU test runs the same
through ``simulate``: NumericalInstabilityError for dt above the limit,
none for a stable run."""

import math

import numpy as np
import pytest

from jaxphys_amd.em.fdtd3d import ENERGY_GUARD_MARGIN, stability_energy_bound
from oracle import c_oracle as CO
from oracle import ext_fdtd as X
from oracle import ref_fdtd as R


def _run(size, boundary, npml, eps, srcs, nsteps, S, every=10):
    """Largest energy/bound ratio over the records, and the first step at
    which the energy exceeds the bound (None if never)."""
    dt = S / math.sqrt(3.0) / R.C0
    o = CO.COracle(size, 1.0, dt, boundary, eps, npml=npml, dtype=np.float64)
    o.set_sources(srcs)
    eps_i = np.ascontiguousarray(eps.transpose(2, 1, 0))
    table = np.stack([t[:nsteps] for _, _, t in srcs], 1)
    bound = stability_energy_bound(table, eps_i, [p for _, p, _ in srcs], 1.0)
    worst, crossed = 0.0, None
    for n0 in range(0, nsteps, every):
        o.run(n0, every)
        n = n0 + every - 1
        e = sum(float((eps_i * a ** 2).sum()) for a in o.f[:3])
        h = sum(float((a ** 2).sum()) for a in o.f[3:])
        energy = 0.5 * R.EPS0 * e + 0.5 * R.MU0 * h
        if not math.isfinite(energy) or energy > bound[n]:
            crossed = n
            break
        worst = max(worst, energy / bound[n])
    return worst * ENERGY_GUARD_MARGIN, crossed


N = 24
C = N // 2
SINE = X.waveform_table("sine", 800, frequency=R.C0 / 20, dt=0.5 / math.sqrt(3) / R.C0)
GD = X.waveform_table("gaussian_derivative", 800, t0=60, spread=20)


def _sphere():
    i, j, k = np.indices((N, N, N))
    return np.where((i - C) ** 2 + (j - C) ** 2 + (k - C) ** 2 <= 25, 4.0, 1.0)


@pytest.mark.parametrize("boundary,npml,eps,src,S", [
    ("walls", 0, "ones", SINE, 0.5),
    ("periodic", 0, "ones", SINE, 0.99),
    ("absorbing", 6, "sphere", SINE, 0.5),
    ("cpml", 6, "sphere", GD, 0.999),
])
def test_stable_runs_stay_below_the_bound(boundary, npml, eps, src, S):
    e = np.ones((N, N, N)) if eps == "ones" else _sphere()
    ratio, crossed = _run((N, N, N), boundary, npml, e, [(2, (C, C, C), src)], 400, S)
    assert crossed is None
    # measured 0.35-0.65 of the unscaled bound: a factor 6 below the guard
    assert ratio < 0.7, ratio


@pytest.mark.parametrize("boundary,npml,S", [("walls", 0, 1.02), ("cpml", 6, 1.1),
                                             ("periodic", 0, 1.5)])
def test_unstable_runs_cross_the_bound_early(boundary, npml, S):
    e = np.ones((N, N, N))
    _, crossed = _run((N, N, N), boundary, npml, e, [(2, (C, C, C), SINE)], 400, S)
    assert crossed is not None and crossed < 100, crossed


def test_bound_closed_form():
    tab = np.array([[1.0, -2.0], [0.5, 0.0], [-3.0, 1.0]])
    eps = np.full((2, 3, 4), 2.0)
    eps[1, 2, 3] = 5.0
    b = stability_energy_bound(tab, eps, [(0, 0, 0), (3, 2, 1)], 0.5)
    cum = np.array([3.0, 3.5, 7.5])
    assert np.allclose(b, ENERGY_GUARD_MARGIN * 0.5 * R.EPS0 * 5.0 * 0.125 * cum ** 2,
                       rtol=1e-15)
    assert stability_energy_bound(np.zeros((3, 0)), eps, [], 1.0).shape == (3,)


@pytest.mark.gpu
def test_simulate_raises_on_unstable_dt(gpu):
    from jaxphys_amd import EMGrid3D, NumericalInstabilityError, PointSource3D

    g = EMGrid3D((24, 20, 16), 0.01, "absorbing", 4)
    g.add_source(PointSource3D(3e9, (12, 10, 8)))
    cfl = 0.01 / (R.C0 * math.sqrt(3.0))
    h = g.simulate(t_span=(0, 200 * cfl), dt=0.9 * cfl, save_every=50, energy_every=10)
    assert len(g.energy_log) == 22 and all(math.isfinite(e) for _, e in g.energy_log)
    with pytest.raises(NumericalInstabilityError, match="CFL"):
        g.simulate(t_span=(0, 400 * cfl), dt=1.1 * cfl, save_every=50, energy_every=10)
    # the guard records how far it got: every record before the raise is finite
    assert g.energy_log and all(math.isfinite(e) for _, e in g.energy_log)
    assert g.energy_log[-1][0] < 100
    # without energy records nothing is checked (the reference's behaviour)
    h = g.simulate(t_span=(0, 40 * cfl), dt=1.1 * cfl, save_every=50)
    assert h.ez.shape[0] == 1
