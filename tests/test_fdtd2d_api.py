"""The reference's 2D FDTD tests (tests/test_em.py:15-77 and
tests/test_edge_cases.py:1523-1547), ported.

Validation tests run on the CPU (they raise before any device work); the
ones that simulate run on the GPU through libtm2d_hip.so.
"""

import pytest
import torch
from pydantic import ValidationError

from jaxphys_amd import ConfigurationError, EMGrid, PlaneWave, Wall

gpu = pytest.mark.gpu


class TestEMGrid:
    @gpu
    def test_basic_simulation(self, gpu) -> None:
        grid = EMGrid(size=(50, 50), resolution=0.01)
        grid.add_source(PlaneWave(frequency=3e9, y=10))
        fields = grid.simulate(t_span=(0, 1e-9), save_every=10)
        assert fields.ez.shape[1] == 50
        assert fields.ez.shape[2] == 50
        assert fields.t.shape[0] > 0

    @gpu
    def test_with_conductor(self, gpu) -> None:
        grid = EMGrid(size=(50, 50), resolution=0.01)
        grid.add_source(PlaneWave(frequency=3e9, y=10))
        grid.add_conductor(Wall(y=25, gap_start=20, gap_end=30))
        fields = grid.simulate(t_span=(0, 5e-10), save_every=5)
        assert fields.ez.shape[0] > 0
        # Ez is zero on the wall outside the gap, and the wave passes the gap
        wall = fields.ez[-1, :, 25]
        assert float(wall[:20].abs().max()) == 0.0 and float(wall[30:].abs().max()) == 0.0

    def test_no_source_error(self) -> None:
        grid = EMGrid(size=(50, 50))
        with pytest.raises(ConfigurationError):
            grid.simulate()

    def test_small_grid_error(self) -> None:
        with pytest.raises(ConfigurationError):
            EMGrid(size=(5, 5))

    @gpu
    def test_multiple_sources_are_applied(self, gpu) -> None:
        grid = EMGrid(size=(50, 50), resolution=0.01)
        grid.add_source(PlaneWave(frequency=3e9, y=10, amplitude=1.0))
        grid.add_source(PlaneWave(frequency=5e9, y=35, amplitude=0.5))
        fields = grid.simulate(t_span=(0, 2e-10), save_every=1)
        top_row = torch.max(torch.abs(fields.ez[-1, :, 10]))
        bottom_row = torch.max(torch.abs(fields.ez[-1, :, 35]))
        assert top_row > 0.0
        assert bottom_row > 0.0

    @gpu
    def test_boundary_modes_affect_solution(self, gpu) -> None:
        periodic = EMGrid(size=(40, 40), resolution=0.01, boundary="periodic", pml_layers=0)
        reflecting = EMGrid(size=(40, 40), resolution=0.01, boundary="reflecting", pml_layers=0)
        source = PlaneWave(frequency=3e9, y=2, amplitude=1.0)
        periodic.add_source(source)
        reflecting.add_source(source)
        periodic_fields = periodic.simulate(t_span=(0, 2e-10), save_every=1)
        reflecting_fields = reflecting.simulate(t_span=(0, 2e-10), save_every=1)
        assert not torch.allclose(periodic_fields.ez[-1], reflecting_fields.ez[-1])


class TestFDTDEdgeCases:
    def test_source_out_of_bounds_raises(self) -> None:
        grid = EMGrid(size=(50, 50))
        with pytest.raises(ConfigurationError, match="out of grid"):
            grid.add_source(PlaneWave(frequency=1e9, y=100))

    def test_conductor_out_of_bounds_raises(self) -> None:
        grid = EMGrid(size=(50, 50))
        with pytest.raises(ConfigurationError, match="out of grid"):
            grid.add_conductor(Wall(y=100))

    def test_em_grid_properties(self) -> None:
        grid = EMGrid(size=(100, 80))
        assert grid.size == (100, 80)


@pytest.mark.parametrize("kw", [dict(resolution=0.0), dict(boundary="cpml"),
                                dict(pml_layers=-1)])
def test_config_validation_matches_reference(kw):
    """config.py:96-110: pydantic rejects these before any grid exists."""
    with pytest.raises(ValidationError):
        EMGrid(size=(20, 20), **kw)


def test_float64_rejected():
    with pytest.raises(ConfigurationError, match="float32"):
        EMGrid(size=(20, 20), dtype=torch.float64)
