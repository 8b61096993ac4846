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


def gabor(nt, dt, f0, frac_bw=0.6, t0=None):
    """Gaussian-windowed cosine of centre frequency f0 whose -6 dB fractional
    bandwidth is frac_bw (simulation/ringfwi/sources.py:29-43): the frequency
    FWHM frac_bw * f0 gives the spectral sigma (FWHM / 2.3548) and the time
    sigma 1 / (2 pi sigma_f); the pulse is centred at t0 (default 3 time sigmas)."""
    sig_f = frac_bw * f0 / 2.3548
    sig_t = 1.0 / (2.0 * np.pi * sig_f)
    t = np.arange(nt) * dt - (3.0 * sig_t if t0 is None else t0)
    return np.exp(-t ** 2 / (2.0 * sig_t ** 2)) * np.cos(2.0 * np.pi * f0 * t)


def hann_tone_burst(nt, dt, f0, n_cycles=1, t0=None):
    """n_cycles of a sine at f0 under a Hann window, starting at t0 (default 0),
    zero outside (simulation/ringfwi/sources.py:46-62)."""
    t0 = 0.0 if t0 is None else t0
    t = np.arange(nt) * dt
    dur = n_cycles / f0
    out = np.zeros(nt)
    on = (t >= t0) & (t <= t0 + dur)
    tau = t[on] - t0
    out[on] = 0.5 * (1.0 - np.cos(2.0 * np.pi * tau / dur)) * np.sin(2.0 * np.pi * f0 * tau)
    return out
