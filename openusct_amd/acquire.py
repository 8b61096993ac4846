"""Simulated acquisition (simulation/ringfwi/acquire.py:18-74): run the device
forward model over a sound-speed model and package the full-matrix capture as
a :class:`~openusct_amd.dataset.Dataset`, the same object a hardware
acquisition produces."""

from __future__ import annotations

import numpy as np

from .dataset import ArrayGeometry, Dataset
from .fwi import forward_fmc
from .phantom import velocity_to_m


def simulate_dataset(c_model, array_geom, wavelet, dt, nominal_speed_m_s, src_list=None,
                     store_ground_truth=True, description=""):
    """Full-matrix capture of ``c_model`` (m/s, 2D ring or 3D cylinder) on the
    grid-mapped array ``array_geom``; sample rate 1/dt, time axis len(wavelet)."""
    wavelet = np.asarray(wavelet, dtype=np.float64)
    traces = forward_fmc(velocity_to_m(c_model), array_geom, wavelet, dt, array_geom.h, len(wavelet),
                         src_list=src_list)
    f_peak = peak_frequency(wavelet, dt)
    geom = ArrayGeometry(element_pos=array_geom.element_positions, radius_m=array_geom.radius_m,
                         centre_freq_hz=f_peak, array_type=getattr(array_geom, "array_type", "ring"))
    tx = list(range(array_geom.n_elements)) if src_list is None else list(src_list)
    truth = {"c": c_model, "h_m": array_geom.h} if store_ground_truth else None
    return Dataset(geometry=geom, data=traces, sample_rate_hz=1.0 / dt, tx_wavelet=wavelet,
                   tx_centre_freq_hz=f_peak, nominal_speed_m_s=nominal_speed_m_s, tx_elements=tx,
                   ground_truth=truth,
                   metadata={"description": description, "created_by": "OpenUSCT simulator"})


def peak_frequency(wavelet, dt):
    """Frequency (Hz) of the largest bin of the wavelet's amplitude spectrum
    (acquire.py:68-74: first maximum of |rfft| on the rfftfreq grid)."""
    spec = np.abs(np.fft.rfft(wavelet))
    return float(np.fft.rfftfreq(len(wavelet), dt)[int(np.argmax(spec))])
