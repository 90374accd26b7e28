"""Validated configuration of the 2D FDTD grid (src/neurosim/config.py:96-110).

``EMGrid`` builds an ``EMConfig`` from its arguments exactly as the
reference does (fdtd.py:96-101), so invalid values raise the same pydantic
``ValidationError`` (a ``ValueError``): ``resolution`` must be > 0,
``courant_number`` in (0, 1], ``boundary`` one of the three literals and
``pml_layers`` >= 0.
"""

from __future__ import annotations

from typing import Literal

from pydantic import BaseModel, Field


class EMConfig(BaseModel):
    """Configuration for electromagnetic FDTD simulations (config.py:96)."""

    resolution: float = Field(0.01, gt=0.0)
    courant_number: float = Field(0.5, gt=0.0, le=1.0)
    boundary: Literal["absorbing", "periodic", "reflecting"] = "absorbing"
    pml_layers: int = Field(10, ge=0)
