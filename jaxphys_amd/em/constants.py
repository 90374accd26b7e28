"""SI constants of the reference (src/neurosim/em/fdtd3d.py:35-38).

``MU0 = 4e-7*pi`` uses the Python float pi (``jnp.pi`` is ``numpy.pi``).
"""

import math

C0 = 299792458.0
MU0 = 4.0e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)
ETA0 = math.sqrt(MU0 / EPS0)
