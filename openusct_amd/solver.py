"""Absorbing-boundary mask of the 2D solver (simulation/ringfwi/solver.py:22-42).

The mask multiplies the new pressure level and the adjoint level every step.
The reference's gradient with a sponge is not the exact adjoint of the damped
scheme; this build reproduces the reference's formula, not a corrected one.
"""

from __future__ import annotations

import numpy as np


def make_sponge(shape, width=25, strength=0.08):
    """Outer product of 1D tapers exp(-(strength (width - i))^2), i < width."""
    ny, nx = shape
    idx = np.arange(width)
    taper = np.exp(-((strength * (width - idx)) ** 2))
    px = np.ones(nx)
    px[:width] = taper
    px[-width:] = taper[::-1]
    py = np.ones(ny)
    py[:width] = taper
    py[-width:] = taper[::-1]
    return np.outer(py, px)
