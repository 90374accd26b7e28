"""jaxphys_amd: MI355X-native 3D FDTD Maxwell step with the API of
sushaan-k/jaxphys (package ``neurosim``): ``EMGrid3D``, ``PointSource3D``,
``DielectricRegion`` and ``EMFieldHistory3D`` (src/neurosim/__init__.py:61,
112, 205-207), running on hand-written HIP kernels for gfx950.

Unlike ``import neurosim`` (src/neurosim/__init__.py:20-24), importing this
package changes no global precision flag: the GPU step computes in float32.
"""

from jaxphys_amd.em.constants import C0, EPS0, MU0
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D
from jaxphys_amd.exceptions import (
    ConfigurationError,
    NativeLibraryError,
    NeurosimError,
    NumericalInstabilityError,
    SimulationError,
)
from jaxphys_amd.state import EMFieldHistory3D

__version__ = "0.1.0"

__all__ = [
    "C0", "EPS0", "MU0",
    "EMGrid3D", "PointSource3D", "DielectricRegion", "EMFieldHistory3D",
    "NeurosimError", "SimulationError", "NumericalInstabilityError",
    "ConfigurationError", "NativeLibraryError",
]
