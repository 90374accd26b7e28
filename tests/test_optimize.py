"""The optimisation driver (mirror of src/neurosim/optimize.py:49-207) on
closed-form objectives: gradient descent and Adam steps checked against
their update formulas, convergence bookkeeping, projection."""

import math

import pytest
import torch

from jaxphys_amd.optimize import OptimizeResult, optimize


def _quad(x):
    return ((x - 3.0) ** 2).sum()


def test_gradient_descent_matches_closed_form():
    # x_{t+1} = x_t - lr * 2 (x_t - 3)  =>  x_t - 3 = (1 - 2 lr)^t (x_0 - 3)
    res = optimize(_quad, 0.0, learning_rate=0.1, max_iterations=25, tolerance=0.0,
                   track_trajectory=True)
    assert isinstance(res, OptimizeResult) and not res.converged
    assert res.n_iterations == 25
    expect = 3.0 - 3.0 * 0.8 ** 25
    assert float(res.x) == pytest.approx(expect, rel=1e-12)
    assert res.trajectory.shape == (26,)
    assert float(res.trajectory[1]) == pytest.approx(0.6, rel=1e-12)
    assert res.fun == pytest.approx((expect - 3.0) ** 2, rel=1e-9)


def test_gradient_descent_convergence_count():
    # converges when |g| = 2 |x - 3| < tol; iteration index of the check
    res = optimize(_quad, 0.0, learning_rate=0.25, max_iterations=100, tolerance=1e-3)
    assert res.converged
    # |x_t - 3| = 3 * 0.5^t ; first t with 6 * 0.5^t < 1e-3
    t = math.ceil(math.log(6.0 / 1e-3) / math.log(2.0))
    assert res.n_iterations == t


def test_adam_first_steps():
    # the first Adam step moves by lr * sign(g) (bias-corrected moments)
    res = optimize(_quad, torch.tensor([0.0, 10.0]), learning_rate=0.5, max_iterations=1,
                   tolerance=0.0, method="adam")
    assert torch.allclose(res.x, torch.tensor([0.5, 9.5], dtype=torch.float64), atol=1e-7)
    res = optimize(_quad, torch.tensor([0.0]), learning_rate=0.1, max_iterations=300,
                   tolerance=1e-6, method="adam")
    assert abs(float(res.x) - 3.0) < 1e-2


def test_projection_and_errors():
    res = optimize(lambda x: -(x ** 2).sum(), torch.tensor([1.0]), learning_rate=1.0,
                   max_iterations=10, tolerance=0.0,
                   project=lambda x: x.clamp(max=4.0))
    assert float(res.x) == 4.0
    # optimize.py:95-134: any method other than "adam" is gradient descent
    a = optimize(_quad, 0.0, learning_rate=0.1, max_iterations=7, tolerance=0.0,
                 method="newton")
    b = optimize(_quad, 0.0, learning_rate=0.1, max_iterations=7, tolerance=0.0)
    assert float(a.x) == float(b.x) and a.n_iterations == b.n_iterations == 7


def test_sensitivity_is_the_jacobian():
    # optimize.py:210-227: d(outputs)/d(params), float64
    from jaxphys_amd.optimize import sensitivity

    A = torch.tensor([[1.0, 2.0, -1.0], [0.5, 0.0, 3.0]], dtype=torch.float64)

    def f(p):
        return A @ p + torch.stack([p[0] * p[1], p[2] ** 3])

    p = torch.tensor([1.5, -2.0, 0.5])
    J = sensitivity(f, p)
    assert J.dtype == torch.float64 and J.shape == (2, 3)
    expect = A.clone()
    expect[0, 0] += -2.0
    expect[0, 1] += 1.5
    expect[1, 2] += 3 * 0.25
    assert torch.allclose(J, expect, rtol=0, atol=1e-12)
    # a scalar output gives the gradient
    g = sensitivity(lambda p: (p ** 2).sum(), [1.0, 2.0])
    assert torch.allclose(g, torch.tensor([2.0, 4.0], dtype=torch.float64))
