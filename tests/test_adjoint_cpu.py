"""Host-side logic of the adjoint: merging of probe terms and the cell ->
probe map the GPU kernels use (no GPU needed)."""

import numpy as np
import pytest

from jaxphys_amd.adjoint import probe_terms
from jaxphys_amd.exceptions import ConfigurationError


def test_probe_terms_merge_duplicates_and_build_the_map():
    probes = [("ez", (3, 4, 5)), ("ex", (3, 4, 5)), ("ez", (3, 4, 5)), (1, (6, 2, 5))]
    keys, w, box, pmap = probe_terms(probes, [1.0, 2.0, 0.5, 3.0], (10, 8, 7))
    assert keys == [(2, 3, 4, 5), (0, 3, 4, 5), (1, 6, 2, 5)]
    assert w == [1.5, 2.0, 3.0]
    assert box == (3, 2, 5, 4, 3, 1)               # (x0, y0, z0, nx, ny, nz)
    expect = np.full((3, 1, 3, 4), -1, dtype=np.int32)   # (c, z, y, x) over the box
    expect[2, 0, 2, 0] = 0
    expect[0, 0, 2, 0] = 1
    expect[1, 0, 0, 3] = 2
    np.testing.assert_array_equal(pmap, expect)


def test_probe_terms_reject_h_components_outside_cells_and_bad_weights():
    with pytest.raises(ConfigurationError, match="E components"):
        probe_terms([("hx", (1, 1, 1))], None, (4, 4, 4))
    with pytest.raises(ConfigurationError, match="outside"):
        probe_terms([("ez", (4, 1, 1))], None, (4, 4, 4))
    with pytest.raises(ConfigurationError, match="one weight"):
        probe_terms([("ez", (1, 1, 1))], [1.0, 2.0], (4, 4, 4))


def test_probe_terms_without_map_when_the_box_is_large():
    keys, w, box, pmap = probe_terms([("ez", (0, 0, 0)), ("ez", (2047, 2047, 1))], None,
                                     (2048, 2048, 2))
    assert len(keys) == 2 and w == [1.0, 1.0]
    assert box is None and pmap is None
    assert probe_terms([], None, (4, 4, 4)) == ([], [], None, None)
