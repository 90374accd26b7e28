"""Gradient-based optimisation over the GPU FDTD path.

Mirror of the reference's driver (src/neurosim/optimize.py:31-207): the
same ``optimize(objective, initial_guess, learning_rate, max_iterations,
tolerance, method, track_trajectory) -> OptimizeResult`` with gradient
descent or Adam (beta1 = 0.9, beta2 = 0.999, eps = 1e-8) and convergence on
the gradient norm.  The reference differentiates with ``jax.grad``; here the
objective is any torch callable whose gradient torch autograd can give --
for electromagnetic inverse design, ``FDTDObjective`` wraps the checkpointed
hand-written adjoint of the Yee step (jaxphys_amd/adjoint.py), so one
gradient costs about three forward runs instead of the stored scan.

``sensitivity`` mirrors the reference's Jacobian helper
(src/neurosim/optimize.py:210-227).  ``design_permittivity`` is the
inverse-design entry: it maximises or
minimises the probe objective of an ``AdjointProblem`` over the permittivity
of its design box, with optional bounds.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Any, Callable

import torch

logger = logging.getLogger(__name__)


@dataclass(frozen=True)
class OptimizeResult:
    """Result of an optimisation run (optimize.py:31-46).

    Attributes:
        x: optimal parameter value (torch tensor, float64 on the input's device).
        fun: objective value at ``x``.
        n_iterations: iterations performed.
        converged: whether the gradient norm fell below the tolerance.
        trajectory: stacked parameter history when tracked, else None.
    """

    x: torch.Tensor
    fun: float
    n_iterations: int
    converged: bool
    trajectory: torch.Tensor | None = None


def _value_and_grad(objective: Callable, x: torch.Tensor):
    xv = x.detach().clone().requires_grad_(True)
    val = objective(xv)
    if not torch.is_tensor(val):
        raise TypeError("objective must return a scalar torch tensor")
    (g,) = torch.autograd.grad(val.sum(), xv)
    return float(val.detach()), g.detach().to(x.dtype)


def _value(objective: Callable, x: torch.Tensor) -> float:
    with torch.no_grad():
        return float(objective(x.detach()))


def optimize(
    objective: Any,
    initial_guess: torch.Tensor | float,
    learning_rate: float = 0.01,
    max_iterations: int = 1000,
    tolerance: float = 1e-8,
    method: str = "gradient_descent",
    track_trajectory: bool = False,
    *,
    project: Callable[[torch.Tensor], torch.Tensor] | None = None,
) -> OptimizeResult:
    """Minimise a scalar objective with gradient descent or Adam
    (optimize.py:49-134, Adam :137-207).

    ``objective`` maps a float64 torch tensor to a scalar torch tensor and
    must be differentiable by torch autograd.  ``project`` (extension,
    keyword-only) is applied to the iterate after every update, e.g. a
    clamp to admissible permittivities.  As in the reference
    (optimize.py:95-134), any ``method`` other than "adam" runs gradient
    descent."""
    x = torch.as_tensor(initial_guess, dtype=torch.float64)
    if torch.is_tensor(initial_guess):
        x = initial_guess.detach().to(torch.float64).clone()
    logger.info("Starting optimization: method=%s, lr=%.2e, max_iter=%d", method,
                learning_rate, max_iterations)
    traj = [x.clone()] if track_trajectory else []

    def result(xr, n, conv):
        return OptimizeResult(x=xr, fun=_value(objective, xr), n_iterations=n, converged=conv,
                              trajectory=torch.stack(traj) if track_trajectory else None)

    if method == "adam":
        beta1, beta2, eps = 0.9, 0.999, 1e-8
        m = torch.zeros_like(x)
        v = torch.zeros_like(x)
        for i in range(1, max_iterations + 1):
            _, g = _value_and_grad(objective, x)
            if float(torch.sqrt(torch.sum(g * g))) < tolerance:
                return result(x, i, True)
            m = beta1 * m + (1 - beta1) * g
            v = beta2 * v + (1 - beta2) * g * g
            m_hat = m / (1 - beta1 ** i)
            v_hat = v / (1 - beta2 ** i)
            x = x - learning_rate * m_hat / (torch.sqrt(v_hat) + eps)
            if project is not None:
                x = project(x)
            if track_trajectory:
                traj.append(x.clone())
        return result(x, max_iterations, False)

    for i in range(max_iterations):
        _, g = _value_and_grad(objective, x)
        if float(torch.sqrt(torch.sum(g * g))) < tolerance:
            logger.info("Converged at iteration %d", i)
            return result(x, i, True)
        x = x - learning_rate * g
        if project is not None:
            x = project(x)
        if track_trajectory:
            traj.append(x.clone())
    logger.warning("Optimization did not converge after %d iterations", max_iterations)
    return result(x, max_iterations, False)


def sensitivity(simulation_fn: Callable, params) -> torch.Tensor:
    """Jacobian d(outputs)/d(params) of ``simulation_fn`` at ``params``
    (optimize.py:210-227, ``jax.jacobian`` there; torch autograd here, one
    backward pass per output).  ``params`` is taken as float64.  For an FDTD
    objective through ``FDTDObjective``, each output must come from its own
    ``AdjointProblem`` (the adjoint uses its problem's latest forward run)."""
    x = torch.as_tensor(params, dtype=torch.float64)
    if torch.is_tensor(params):
        x = params.detach().to(torch.float64).clone()
    return torch.autograd.functional.jacobian(simulation_fn, x)


def design_permittivity(problem, eps_init: torch.Tensor, *, maximize: bool = True,
                        learning_rate: float = 0.05, max_iterations: int = 20,
                        eps_min: float = 1.0, eps_max: float | None = None,
                        method: str = "adam", track_trajectory: bool = False) -> OptimizeResult:
    """Optimise the relative permittivity of ``problem``'s design box
    (an ``AdjointProblem``): maximise (default) or minimise its probe
    objective, keeping eps_r within [eps_min, eps_max].  Each iteration runs
    one checkpointed forward and one adjoint sweep on the GPU."""
    from jaxphys_amd.adjoint import FDTDObjective

    sign = -1.0 if maximize else 1.0

    def objective(e):
        return sign * FDTDObjective.apply(e, problem)

    def project(e):
        return e.clamp(min=eps_min, max=eps_max)

    res = optimize(objective, eps_init, learning_rate=learning_rate,
                   max_iterations=max_iterations, tolerance=0.0, method=method,
                   track_trajectory=track_trajectory, project=project)
    return OptimizeResult(x=res.x, fun=sign * res.fun, n_iterations=res.n_iterations,
                          converged=res.converged, trajectory=res.trajectory)
