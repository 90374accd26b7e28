"""Electromagnetics: the 2D TM and 3D FDTD solvers (src/neurosim/em/__init__.py:7-8)."""

from jaxphys_amd.em.fdtd import EMGrid, PlaneWave, Wall
from jaxphys_amd.em.fdtd3d import DielectricRegion, EMGrid3D, PointSource3D

__all__ = ["EMGrid", "PlaneWave", "Wall", "EMGrid3D", "PointSource3D", "DielectricRegion"]
