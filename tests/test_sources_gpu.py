"""A rejected source update leaves the previous sources in place: valid
set_sources, then one that is rejected (out-of-grid position), then a run
that must equal the run with the valid sources (bit for bit)."""

import numpy as np
import pytest
import torch

from jaxphys_amd import ConfigurationError, EMGrid3D, PointSource3D


@pytest.mark.gpu
@pytest.mark.parametrize("boundary", ["absorbing", "cpml"])
def test_rejected_set_sources_keeps_previous_sources(gpu, boundary):
    g = EMGrid3D((24, 20, 16), 0.01, boundary, 4)
    g.add_source(PointSource3D(5e9, (10, 10, 8)))
    g.add_source(PointSource3D(3e9, (6, 12, 9), polarization="x"))
    n = 30
    dt = 0.5 * 0.01 / (2.99792458e8 * np.sqrt(3.0))
    ref = g.build_solver(dt, n)
    ref.run(0, n)
    sol = g.build_solver(dt, n)
    table = sol.source_table.copy()
    with pytest.raises(ConfigurationError, match="out of grid"):
        sol.set_sources([2, 0, 1], [(10, 10, 8), (6, 12, 9), (24, 0, 0)],
                        np.zeros((n, 3), np.float32))
    # the solver's own record of its sources was not changed either
    assert sol.n_sources == 2 and np.array_equal(sol.source_table, table)
    sol.run(0, n)
    torch.cuda.synchronize()
    for a, b in zip(sol.fields, ref.fields):
        assert torch.equal(a, b)
    assert float(sol.fields[2].abs().max()) > 0
