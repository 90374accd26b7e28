"""The HIP path through the EMGrid3D API against the committed golden
fixtures (the oracle's float32 output, tests/golden/cases.py): bit-exact
where no transcendental is evaluated per cell (walls, periodic), relative L2
<= 1e-5 on the stored planes and norms otherwise (sponge exp, CPML)."""

import os
import sys

import numpy as np
import pytest

from jaxphys_amd import DielectricRegion, EMGrid3D, PointSource3D

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

import cases as K  # noqa: E402

NAMES = ["ex", "ey", "ez", "hx", "hy", "hz"]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _gpu_snaps(name):
    c = K.CASES[name]
    size = c["size"]
    g = EMGrid3D(size=size, resolution=c["dx"], boundary=c["boundary"], pml_layers=c["pml"])
    if c["boundary"] == "cpml":
        g.add_material(DielectricRegion(mask=np.ones(size, bool), epsilon_r=K.eps_of(c),
                                        sigma=K.sigma_of(c)))
        for comp, pos, kind, kw in c["sources"]:
            g.add_source(PointSource3D(0.0, pos, kw.get("amplitude", 1.0), "xyz"[comp],
                                       waveform=kind, t0=kw.get("t0", 0.0),
                                       spread=kw.get("spread", 1.0),
                                       wavelength=kw.get("wavelength")))
        h = g.simulate(dt=K.dt_cpml(c), save_every=c["save_every"], n_steps=c["n_steps"])
    else:
        for f, p, a, pol in c["sources"]:
            g.add_source(PointSource3D(f, p, a, pol))
        for m, e in K.masks_of(c):
            g.add_material(DielectricRegion(mask=m, epsilon_r=e))
        h = g.simulate(save_every=c["save_every"], n_steps=c["n_steps"])
    return {n: getattr(h, n).cpu().numpy() for n in NAMES}


@pytest.mark.parametrize("name", list(K.CASES))
def test_gpu_matches_golden(gpu, name):
    c = K.CASES[name]
    print(f"[golden] {name}", flush=True)
    gold = np.load(os.path.join(GOLDEN, "fdtd3d_golden.npz"))
    got = _gpu_snaps(name)
    for n in NAMES:
        a = got[n][-1:] if c.get("last_only") else got[n]
        norm = np.sqrt((a.astype(np.float64) ** 2).sum(axis=(1, 2, 3)))
        if c["store"] == "final":
            np.testing.assert_array_equal(a[-1], gold[f"{name}/{n}"], err_msg=n)
        else:
            ref = gold[f"{name}/{n}/plane"]
            plane = a[:, :, :, a.shape[3] // 2]
            assert plane.shape == ref.shape
            for q in range(ref.shape[0]):
                assert _rel(plane[q], ref[q]) <= 1e-5, (n, q)
        np.testing.assert_allclose(norm, gold[f"{name}/{n}/norm"], rtol=1e-5, atol=0)
