"""Source wavelets (simulation/ringfwi/sources.py:8-26)."""

from __future__ import annotations

import numpy as np


def ricker(nt, dt, f0, t0=None):
    """Ricker wavelet (1 - 2a) exp(-a), a = (pi f0 (n dt - t0))^2, t0 = 1/f0 by default."""
    if t0 is None:
        t0 = 1.0 / f0
    t = np.arange(nt) * dt - t0
    a = (np.pi * f0 * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)
