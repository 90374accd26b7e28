"""The reference's 12 FDTD3D tests (tests/test_fdtd3d.py:16-139), ported.

Validation tests run on the CPU (they raise before any device work); the
ones that simulate run on the GPU through libyee_hip.so.
"""

import numpy as np
import pytest
import torch

from jaxphys_amd import ConfigurationError, DielectricRegion, EMGrid3D, PointSource3D

gpu = pytest.mark.gpu


class TestEMGrid3D:
    @gpu
    def test_basic_simulation(self, gpu) -> None:
        grid = EMGrid3D(size=(20, 20, 20), resolution=0.01)
        grid.add_source(PointSource3D(frequency=3e9, position=(10, 10, 10)))
        fields = grid.simulate(t_span=(0, 5e-10), save_every=10)
        assert fields.ez.shape[1] == 20
        assert fields.ez.shape[2] == 20
        assert fields.ez.shape[3] == 20
        assert fields.t.shape[0] > 0
        assert fields.t.shape[0] == 3  # 26 steps, saved at 0, 10, 20

    @gpu
    def test_source_excites_field(self, gpu) -> None:
        grid = EMGrid3D(size=(20, 20, 20), resolution=0.01)
        grid.add_source(PointSource3D(frequency=3e9, position=(10, 10, 10), amplitude=1.0))
        fields = grid.simulate(t_span=(0, 3e-10), save_every=5)
        assert float(torch.max(torch.abs(fields.ez[-1]))) > 0.0

    @gpu
    def test_all_six_components(self, gpu) -> None:
        grid = EMGrid3D(size=(16, 16, 16), resolution=0.01)
        grid.add_source(PointSource3D(frequency=5e9, position=(8, 8, 8)))
        fields = grid.simulate(t_span=(0, 2e-10), save_every=5)
        assert fields.ex.shape == fields.ey.shape == fields.ez.shape
        assert fields.hx.shape == fields.hy.shape == fields.hz.shape
        assert fields.ex.shape[1:] == (16, 16, 16)

    @gpu
    def test_polarization_x(self, gpu) -> None:
        grid = EMGrid3D(size=(16, 16, 16), resolution=0.01)
        grid.add_source(PointSource3D(frequency=3e9, position=(8, 8, 8), polarization="x"))
        fields = grid.simulate(t_span=(0, 2e-10), save_every=5)
        assert float(torch.max(torch.abs(fields.ex[-1]))) > 0.0

    @gpu
    def test_polarization_y(self, gpu) -> None:
        grid = EMGrid3D(size=(16, 16, 16), resolution=0.01)
        grid.add_source(PointSource3D(frequency=3e9, position=(8, 8, 8), polarization="y"))
        fields = grid.simulate(t_span=(0, 2e-10), save_every=5)
        assert float(torch.max(torch.abs(fields.ey[-1]))) > 0.0

    @gpu
    def test_dielectric_material(self, gpu) -> None:
        grid_free = EMGrid3D(size=(16, 16, 16), resolution=0.01)
        grid_free.add_source(PointSource3D(frequency=5e9, position=(8, 8, 8)))
        grid_diel = EMGrid3D(size=(16, 16, 16), resolution=0.01)
        grid_diel.add_source(PointSource3D(frequency=5e9, position=(8, 8, 8)))
        mask = np.zeros((16, 16, 16), dtype=bool)
        mask[10:, :, :] = True
        grid_diel.add_material(DielectricRegion(mask=mask, epsilon_r=4.0))
        fields_free = grid_free.simulate(t_span=(0, 1e-9), save_every=20)
        fields_diel = grid_diel.simulate(t_span=(0, 1e-9), save_every=20)
        assert not torch.allclose(fields_free.ez[-1], fields_diel.ez[-1])

    @gpu
    def test_periodic_boundary(self, gpu) -> None:
        grid_abs = EMGrid3D(size=(16, 16, 16), resolution=0.01, boundary="absorbing")
        grid_per = EMGrid3D(size=(16, 16, 16), resolution=0.01, boundary="periodic",
                            pml_layers=0)
        source = PointSource3D(frequency=5e9, position=(8, 8, 8))
        grid_abs.add_source(source)
        grid_per.add_source(source)
        fields_abs = grid_abs.simulate(t_span=(0, 2e-10), save_every=5)
        fields_per = grid_per.simulate(t_span=(0, 2e-10), save_every=5)
        assert not torch.allclose(fields_abs.ez[-1], fields_per.ez[-1])

    def test_no_source_error(self) -> None:
        grid = EMGrid3D(size=(16, 16, 16))
        with pytest.raises(ConfigurationError):
            grid.simulate()

    def test_small_grid_error(self) -> None:
        with pytest.raises(ConfigurationError):
            EMGrid3D(size=(4, 4, 4))

    def test_source_out_of_bounds(self) -> None:
        grid = EMGrid3D(size=(16, 16, 16))
        with pytest.raises(ConfigurationError):
            grid.add_source(PointSource3D(frequency=1e9, position=(20, 8, 8)))

    def test_material_shape_mismatch(self) -> None:
        grid = EMGrid3D(size=(16, 16, 16))
        bad_mask = np.zeros((10, 10, 10), dtype=bool)
        with pytest.raises(ConfigurationError):
            grid.add_material(DielectricRegion(mask=bad_mask))

    @gpu
    def test_grid_coordinates(self, gpu) -> None:
        grid = EMGrid3D(size=(16, 16, 16), resolution=0.02)
        grid.add_source(PointSource3D(frequency=1e9, position=(8, 8, 8)))
        fields = grid.simulate(t_span=(0, 1e-10), save_every=5)
        assert fields.grid_x.shape == (16,)
        assert float(fields.grid_x[1] - fields.grid_x[0]) == pytest.approx(0.02)


def test_error_messages_match_reference() -> None:
    """Same message text as fdtd3d.py:105, 117-118, 127-128, 198-199."""
    with pytest.raises(ConfigurationError, match=r"Grid must be at least 8x8x8, got \(4, 4, 4\)"):
        EMGrid3D(size=(4, 4, 4))
    grid = EMGrid3D(size=(16, 16, 16))
    with pytest.raises(ConfigurationError,
                       match=r"Source position \(20, 8, 8\) out of grid bounds \(16, 16, 16\)"):
        grid.add_source(PointSource3D(frequency=1e9, position=(20, 8, 8)))
    with pytest.raises(ConfigurationError, match=r"doesn't match grid size \(16, 16, 16\)"):
        grid.add_material(DielectricRegion(mask=np.zeros((10, 10, 10), dtype=bool)))
    with pytest.raises(ConfigurationError, match="At least one source must be added"):
        grid.simulate()


def test_extension_validation() -> None:
    with pytest.raises(ConfigurationError):
        EMGrid3D(size=(16, 16, 16), boundary="cpml", pml_layers=9)
    with pytest.raises(ConfigurationError):
        EMGrid3D(size=(16, 16, 16), dtype=torch.float64)
