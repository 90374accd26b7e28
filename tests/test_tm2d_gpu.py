"""2D TM HIP step (libtm2d_hip.so) vs the float32 oracle (oracle/ref_tm2d.py,
the restatement of src/neurosim/em/fdtd.py): bit-exact on every snapshot,
every boundary mode, walls, several sources on one row, grid sizes that are
not tile multiples, and random initial states; at larger sizes, exact
translation invariance of the periodic step."""

import numpy as np
import pytest
import torch

from jaxphys_amd import EMGrid, PlaneWave, Wall
from oracle import ref_tm2d as R

pytestmark = pytest.mark.gpu

F32 = np.float32


def _grid(size, boundary, npml, waves, walls):
    g = EMGrid(size=size, resolution=0.01, boundary=boundary, pml_layers=npml)
    for w in waves:
        g.add_source(PlaneWave(w.frequency, w.y, w.amplitude))
    for w in walls:
        g.add_conductor(Wall(w.y, w.gap_start, w.gap_end))
    return g


@pytest.mark.parametrize("boundary", ["absorbing", "periodic", "reflecting"])
@pytest.mark.parametrize("size,walls", [
    ((67, 45), []),
    ((130, 37), [R.Wall(20, 30, 60), R.Wall(28), R.Wall(33, 5, None)]),
])
def test_simulate_bit_exact_vs_oracle(gpu, boundary, size, walls):
    waves = [R.Wave(3e9, 10, 1.0), R.Wave(5e9, 10, 0.5), R.Wave(4e9, size[1] - 1, 2.0)]
    print(f"[tm2d] {boundary} {size} walls={len(walls)}", flush=True)
    g = _grid(size, boundary, 6, waves, walls)
    h = g.simulate(t_span=(0.0, 1.2e-9), save_every=3)
    t_ref, ref = R.simulate(size, 0.01, boundary, 6, waves, walls, t_span=(0.0, 1.2e-9),
                            save_every=3, dtype=F32)
    np.testing.assert_array_equal(h.t.cpu().numpy(), t_ref)
    assert h.ez.shape == ref["ez"].shape and h.ez.shape[0] > 10
    for c in ("ez", "hx", "hy"):
        np.testing.assert_array_equal(getattr(h, c).cpu().numpy(), ref[c], err_msg=c)
    assert float(h.ez[-1].abs().max()) > 0


@pytest.mark.parametrize("boundary", ["absorbing", "periodic", "reflecting"])
def test_random_state_bit_exact(gpu, boundary):
    """All three fields nonzero everywhere: exercises every update term."""
    size = (93, 70)
    rng = np.random.default_rng(3)
    init = [rng.standard_normal(size).astype(F32) for _ in range(3)]
    waves = [R.Wave(2e9, 33, 0.3)]
    walls = [R.Wall(50, 10, 40)]
    g = _grid(size, boundary, 9, waves, walls)
    dt = R.default_dt(0.01, F32)
    sol = g.build_solver(dt, 25)
    sol.set_state(*init)
    sol.run(0, 25)
    _, ref = R.simulate(size, 0.01, boundary, 9, waves, walls, dt=dt, n_steps=25, save_every=1,
                        dtype=F32, init=init)
    for c, f in zip(("ez", "hx", "hy"), sol.fields):
        np.testing.assert_array_equal(f.T.cpu().numpy(), ref[c][-1], err_msg=c)


@pytest.mark.parametrize("boundary", ["absorbing", "periodic", "reflecting"])
@pytest.mark.parametrize("size", [(264, 70), (512, 41)])
def test_marching_wavefronts_bit_exact(gpu, boundary, size):
    """Several full 64-column marching wavefronts (the edge columns in the
    first / last lane) and partial ones, rows inside and at the edges,
    random state, walls."""
    rng = np.random.default_rng(size[0])
    init = [rng.standard_normal(size).astype(F32) for _ in range(3)]
    waves = [R.Wave(2e9, 17, 0.3), R.Wave(3e9, 17, 0.2)]
    walls = [R.Wall(20, 100, 140), R.Wall(33)]
    g = _grid(size, boundary, 7, waves, walls)
    dt = R.default_dt(0.01, F32)
    sol = g.build_solver(dt, 14)
    sol.set_state(*init)
    sol.run(0, 14)
    _, ref = R.simulate(size, 0.01, boundary, 7, waves, walls, dt=dt, n_steps=14, save_every=1,
                        dtype=F32, init=init)
    for c, f in zip(("ez", "hx", "hy"), sol.fields):
        np.testing.assert_array_equal(f.T.cpu().numpy(), ref[c][-1], err_msg=c)


def test_step_entry_point_matches_run(gpu):
    """tm2d_step (one step, explicit buffers) equals tm2d_run step by step."""
    from jaxphys_amd import _tm2d

    size = (75, 40)
    g = _grid(size, "absorbing", 5, [R.Wave(3e9, 12)], [R.Wall(30, 20, 25)])
    dt = R.default_dt(0.01, F32)
    a = g.build_solver(dt, 12)
    b = g.build_solver(dt, 12)
    a.run(0, 12)
    cur = [torch.zeros(size[::-1], device=gpu) for _ in range(3)]
    nxt = [torch.zeros(size[::-1], device=gpu) for _ in range(3)]
    for n in range(12):
        _tm2d.step(b.desc, cur, nxt, n)
        cur, nxt = nxt, cur
    for x, y in zip(a.fields, cur):
        assert torch.equal(x, y)


def test_medium_grid_bit_exact(gpu):
    size = (1024, 600)
    waves = [R.Wave(3e9, 300, 1.0), R.Wave(7e9, 100, 0.25)]
    walls = [R.Wall(400, 480, 540)]
    g = _grid(size, "absorbing", 10, waves, walls)
    h = g.simulate(t_span=(0.0, 1e-9), save_every=20)
    _, ref = R.simulate(size, 0.01, "absorbing", 10, waves, walls, t_span=(0.0, 1e-9),
                        save_every=20, dtype=F32)
    for c in ("ez", "hx", "hy"):
        np.testing.assert_array_equal(getattr(h, c)[-1].cpu().numpy(), ref[c][-1], err_msg=c)


def test_periodic_translation_invariance_large(gpu):
    """A cyclic shift of the (n-2)-periodic interior commutes with the
    periodic step, bit for bit, at a size the oracle would take minutes on."""
    nx, ny = 4096, 3074
    px, py = nx - 2, ny - 2
    rng = torch.Generator(device="cpu").manual_seed(11)
    core = [torch.randn(py, px, generator=rng) for _ in range(3)]

    def embed(c):  # (py, px) interior -> (ny, nx) with consistent copies
        f = torch.zeros(ny, nx)
        f[1:-1, 1:-1] = c
        f[0, :], f[-1, :] = f[-2, :], f[1, :]
        f[:, 0], f[:, -1] = f[:, -2], f[:, 1]
        return f

    sx, sy = 517, 1203
    g = EMGrid(size=(nx, ny), resolution=0.01, boundary="periodic", pml_layers=0)
    g.add_source(PlaneWave(3e9, 5, 0.0))
    dt = R.default_dt(0.01, F32)
    outs = []
    for shift in ((0, 0), (sy, sx)):
        sol = g.build_solver(dt, 40)
        for f, c in zip(sol.fields, core):
            f.copy_(embed(torch.roll(c, shifts=shift, dims=(0, 1))).to(gpu))
        sol.run(0, 40)
        outs.append([f[1:-1, 1:-1].cpu() for f in sol.fields])
        del sol
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(torch.roll(a, shifts=(sy, sx), dims=(0, 1)), b)
    assert float(outs[0][0].abs().max()) > 0


def test_empty_history_and_every_step(gpu):
    """fdtd.py:362-368: zero steps give empty histories; save_every <= 1
    keeps every step (t = t_c + dt)."""
    g = _grid((20, 12), "reflecting", 0, [R.Wave(3e9, 5, 1.0)], [])
    h = g.simulate(t_span=(0.0, 1e-12))
    assert h.t.shape == (0,) and h.ez.shape == (0, 20, 12) and h.hy.shape == (0, 20, 12)
    for every in (0, 1):
        h = g.simulate(t_span=(0.0, 1e-10), save_every=every)
        t_ref, ref = R.simulate((20, 12), 0.01, "reflecting", 0, [R.Wave(3e9, 5, 1.0)],
                                t_span=(0.0, 1e-10), save_every=every, dtype=F32)
        assert h.ez.shape[0] == ref["ez"].shape[0] == 4
        np.testing.assert_array_equal(h.t.cpu().numpy(), t_ref)
        for c in ("ez", "hx", "hy"):
            np.testing.assert_array_equal(getattr(h, c).cpu().numpy(), ref[c])
