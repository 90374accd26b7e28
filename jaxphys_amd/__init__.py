"""jaxphys_amd: MI355X-native FDTD Maxwell solvers with the API of
sushaan-k/jaxphys (package ``neurosim``): the 3D ``EMGrid3D``,
``PointSource3D``, ``DielectricRegion``, ``EMFieldHistory3D`` and the 2D TM
``EMGrid``, ``PlaneWave``, ``Wall``, ``EMFieldHistory``
(src/neurosim/__init__.py:60-61, 111-112, 170, 202-207), running on
hand-written HIP kernels for gfx950.

Unlike ``import neurosim`` (src/neurosim/__init__.py:20-24), importing this
package changes no global precision flag: the GPU step computes in float32.
"""

from jaxphys_amd.config import EMConfig
from jaxphys_amd.em.constants import C0, EPS0, MU0
from jaxphys_amd.em.fdtd import EMGrid, PlaneWave, Wall
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D
from jaxphys_amd.exceptions import (
    ConfigurationError,
    NativeLibraryError,
    NeurosimError,
    NumericalInstabilityError,
    SimulationError,
)
from jaxphys_amd.state import EMFieldHistory, EMFieldHistory3D

__version__ = "0.1.0"

__all__ = [
    "C0", "EPS0", "MU0",
    "EMGrid3D", "PointSource3D", "DielectricRegion", "EMFieldHistory3D",
    "EMGrid", "PlaneWave", "Wall", "EMFieldHistory", "EMConfig",
    "NeurosimError", "SimulationError", "NumericalInstabilityError",
    "ConfigurationError", "NativeLibraryError",
]
