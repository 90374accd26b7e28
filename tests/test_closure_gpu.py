"""The learned closure end to end at a small size: fine/coarse pairs from the
FDTD solver, training on the BF16-MFMA kernels, and the gain over the plain
coarse grid on held-out seeds (fine FDTD is the judge)."""

import numpy as np
import pytest
import torch

from jaxphys_amd import closure as CL

pytestmark = pytest.mark.gpu


def test_restricted_fine_start_matches_coarse_start(gpu):
    """Before any step both grids are zero; after the first coarse interval
    the restricted fine state and the coarse state agree to first order
    (same pulse, same medium): the relative difference is well below 1."""
    prob = CL.ClosureProblem(seed=5, n_fine=64, npml_fine=8, device=gpu)
    states, _ = prob.fine_states(40, 10)
    sol = CL.coarse_rollout(prob, None, 40, 10)
    err = CL.rel_err(sol.fields, states[-1], margin=prob.npml_c)
    print(f"[closure] plain coarse vs restricted fine after 40 coarse steps: {err:.3f}", flush=True)
    assert 0.0 < err < 0.8


def test_closure_training_beats_plain_coarse(gpu):
    kw = dict(npml_fine=8)
    net, hist = CL.train_closure(gpu, seeds=list(range(6)), n_fine=64, n_coarse=100, every=10,
                                 epochs=3, crop=32, lr=2e-3, **kw)
    # the least-squares source model was fitted (the crops held whole boxes)
    assert float(net.ks.abs().max()) > 0
    assert np.isfinite(hist).all()
    assert np.mean(hist[-20:]) < np.mean(hist[:20])
    res = [CL.evaluate_closure(net, s, n_fine=64, n_coarse=100, every=10, **kw)
           for s in (1000, 1001)]
    for r in res:
        print(f"[closure] seed {r['seed']}: coarse {r['err_coarse']:.4f} "
              f"closure {r['err_closure']:.4f}", flush=True)
    assert np.mean([r["err_closure"] for r in res]) < np.mean([r["err_coarse"] for r in res])
