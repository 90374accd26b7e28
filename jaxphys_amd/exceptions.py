"""Exception names of the reference package (src/neurosim/exceptions.py:8-45).

Only the ones the FDTD path raises are kept; the hierarchy is the same so
``except NeurosimError`` in user code keeps working.
"""


class NeurosimError(Exception):
    """Base exception (exceptions.py:8)."""


class SimulationError(NeurosimError):
    """A simulation failed at run time (exceptions.py:12).  The HIP library's
    runtime errors (YEE_EHIP) map here."""


class NumericalInstabilityError(SimulationError):
    """Unstable integration, e.g. a time step above the 3D CFL limit whose
    float64 energy check blows up (exceptions.py:16)."""


class ConfigurationError(NeurosimError):
    """Invalid or inconsistent parameters (exceptions.py:24)."""


class NativeLibraryError(SimulationError):
    """The HIP library is missing or cannot run here (no reference
    counterpart: the reference has no native code)."""


class VisualizationError(NeurosimError):
    """Plotting is unavailable or was asked for something that does not
    exist (exceptions.py:44)."""
