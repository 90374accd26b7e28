"""Write tests/golden/fdtd3d_golden.npz from the oracle (float32).

    python tests/golden/make_fixtures.py

The reference holds no golden vectors for the FDTD path, so these fixtures
are the oracle's own outputs on the cases of cases.py; the CPU test
re-derives them bit for bit and the GPU test compares the HIP path with
them."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import cases as K  # noqa: E402
from oracle import ext_fdtd as X  # noqa: E402
from oracle import ref_fdtd as R  # noqa: E402

NAMES = ["ex", "ey", "ez", "hx", "hy", "hz"]


def run_case(name):
    """Float32 oracle snapshots of a case: dict comp -> (n_snap, nx, ny, nz)."""
    c = K.CASES[name]
    size = c["size"]
    if c["boundary"] == "cpml":
        dt = K.dt_cpml(c)
        S = R.C0 * dt / c["dx"]
        srcs = []
        for comp, pos, kind, kw in c["sources"]:
            tab = X.waveform_table(kind, c["n_steps"], amplitude=kw.get("amplitude", 1.0),
                                   t0=kw.get("t0", 0.0), spread=kw.get("spread", 1.0),
                                   wavelength=kw.get("wavelength"), courant=S)
            srcs.append((comp, pos, tab))
        st = X.run(size, c["dx"], dt, "cpml", K.eps_of(c), srcs, c["n_steps"],
                   sigma=K.sigma_of(c), npml=c["pml"], dtype=np.float32)
        return {n: f[None].astype(np.float32) for n, f in zip(NAMES, st.f)}
    srcs = [R.Src(f, p, a, pol) for f, p, a, pol in c["sources"]]
    mats = [R.Mat(m, e) for m, e in K.masks_of(c)]
    dt = R.default_dt(c["dx"], np.float32)
    _, snaps = R.simulate(size, c["dx"], c["boundary"], c["pml"], srcs, mats, dt=dt,
                          save_every=c["save_every"], n_steps=c["n_steps"], dtype=np.float32)
    return snaps


def record(name, snaps):
    """The stored part of a case's result."""
    c = K.CASES[name]
    out = {}
    for n in NAMES:
        a = snaps[n][-1:] if c.get("last_only") else snaps[n]
        if c["store"] == "final":
            out[f"{name}/{n}"] = a[-1]
        else:
            out[f"{name}/{n}/plane"] = a[:, :, :, a.shape[3] // 2]
        out[f"{name}/{n}/norm"] = np.sqrt((a.astype(np.float64) ** 2).sum(axis=(1, 2, 3)))
    return out


def main():
    data = {}
    for name in K.CASES:
        data.update(record(name, run_case(name)))
    path = os.path.join(HERE, "fdtd3d_golden.npz")
    np.savez_compressed(path, **data)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
