"""Electromagnetics: the 3D FDTD solver (src/neurosim/em/__init__.py:8)."""

from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D

__all__ = ["EMGrid3D", "PointSource3D", "DielectricRegion"]
