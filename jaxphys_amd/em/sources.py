"""Per-step source value tables for the HIP step.

The reference's only waveform is ``amplitude * sin(omega * t_n)`` with
``t_n = t_start + n*dt`` evaluated in the working float type
(src/neurosim/em/fdtd3d.py:227, 236-238, 288).  The pulse waveforms are
extensions (no reference counterpart); time is measured in steps ``n`` and
the value of step ``n`` is added after the E update that produces E^{n+1},
exactly where the reference adds its sine.
"""

from __future__ import annotations

import math

import numpy as np

WAVEFORMS = ("sine", "gaussian", "gaussian_derivative", "gaussian_sine")


def reference_sine(frequency: float, amplitude: float, n_steps: int, t_start: float,
                   dt: float) -> np.ndarray:
    """float32 values in the reference's float32 operation order:
    ``t_c = t_start + f32(n)*dt``; ``val = A * sin(f32(omega) * t_c)``."""
    f32 = np.float32
    n = np.arange(n_steps, dtype=np.float32)
    t_c = (f32(t_start) + (n * f32(dt)).astype(np.float32)).astype(np.float32)
    omega = f32(2.0 * math.pi * frequency)
    return (f32(amplitude) * np.sin((omega * t_c).astype(np.float32))).astype(np.float32)


def pulse(kind: str, n_steps: int, amplitude: float, t0: float, spread: float,
          wavelength: float | None = None, courant: float | None = None) -> np.ndarray:
    """Extension pulses, computed in float64 and rounded once to float32.

    gaussian:            A exp(-u^2/2),                 u = (n - t0)/spread
    gaussian_derivative: -A u exp(-u^2/2)
    gaussian_sine:       A sin(2 pi S n / wavelength) exp(-u^2/2)
                         (wavelength in cells, S = c dt/dx the Courant number)
    """
    n = np.arange(n_steps, dtype=np.float64)
    u = (n - t0) / spread
    env = np.exp(-0.5 * u * u)
    if kind == "gaussian":
        v = env
    elif kind == "gaussian_derivative":
        v = -u * env
    elif kind == "gaussian_sine":
        if not wavelength or courant is None:
            raise ValueError("gaussian_sine needs wavelength (cells) and the Courant number")
        v = np.sin(2.0 * math.pi * courant * n / wavelength) * env
    else:
        raise ValueError(f"unknown waveform {kind!r}; expected one of {WAVEFORMS}")
    return (amplitude * v).astype(np.float32)
