"""Pins of the gradient oracle ``oracle/autograd_fdtd.py`` (CPU, float64).

Its forward pass must equal the NumPy extension oracle, and its reverse-mode
gradient must equal central finite differences of the NumPy oracle's
objective -- so a wrong index or sign in either restatement shows up."""

import math

import numpy as np
import pytest
import torch

from oracle import autograd_fdtd as A
from oracle import ext_fdtd as X
from oracle import ref_fdtd as R


def _objective_numpy(size, dx, dt, boundary, eps, srcs, n, npml, sigma, mon):
    acc = [0.0]

    def obs(_, st):
        acc[0] += float((st.f[2][mon] ** 2).sum())

    X.run(size, dx, dt, boundary, eps, srcs, n, sigma=sigma, npml=npml, dtype=np.float64,
          observe=obs)
    return acc[0]


@pytest.mark.parametrize("boundary,npml,lossy", [("cpml", 3, False), ("cpml", 3, True),
                                                 ("absorbing", 3, False), ("periodic", 0, False)])
def test_autograd_oracle_forward_and_fd(boundary, npml, lossy):
    size = (12, 11, 10)
    dx = 1.0
    dt = 0.5 / math.sqrt(3) / R.C0
    rng = np.random.default_rng(8)
    eps = 1.0 + 3.0 * rng.random(size)
    sigma = None
    if lossy:
        sigma = np.zeros(size)
        sigma[4:8, :, :] = 0.05 * R.EPS0 * R.C0
    n = 30
    tab = X.waveform_table("gaussian_derivative", n, t0=8, spread=3)
    srcs = [(2, (5, 5, 3), tab), (0, (7, 4, 5), 0.3 * tab)]
    mon = np.zeros(size, bool)
    mon[3:8, 4:8, 7] = True
    J, g = A.gradient(size, dx, dt, boundary, eps, srcs, n, npml, sigma, mon)
    J_np = _objective_numpy(size, dx, dt, boundary, eps, srcs, n, npml, sigma, mon)
    assert J == pytest.approx(J_np, rel=1e-12)
    for cell in [(5, 5, 5), (6, 7, 4), (4, 6, 6)]:
        h = 1e-4
        ep = eps.copy(); ep[cell] += h  # noqa: E702
        em = eps.copy(); em[cell] -= h  # noqa: E702
        fd = (_objective_numpy(size, dx, dt, boundary, ep, srcs, n, npml, sigma, mon)
              - _objective_numpy(size, dx, dt, boundary, em, srcs, n, npml, sigma, mon)) / (2 * h)
        assert g[cell] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(J)), (cell, g[cell], fd)
    assert np.abs(g).max() > 0
