"""GPU: the reference's end-to-end backend benchmark (software/python/bench_backend.py:19-62).
The 3D cylinder-array inversion of the demonstration (30^3, 12 transmitters,
nt = 300, 6 iterations, mask and bounds) runs twice through fwi.invert: once
with the reference's own C++ core as the backend (libuap compiled from its
sources into oracle/_ref, behind an adapter with uap.py:34-67's signatures),
once on the device.  The models agree as the reference's benchmark requires
its two backends to (max |m_a - m_b| / max |m_a|, here <= 1e-5 for FP32 fields
against FP64) and the wall-clock speed-up is printed.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class _RefCore:
    """The reference core behind the backend protocol of invert (fwi.py:594-603)."""

    def __init__(self, mod):
        self.mod = mod

    def _lin(self, geom, shape, src_list):
        from openusct_amd.plan import linear_indices
        rec = linear_indices(geom.idx, shape).astype(np.int32)
        tx = rec if src_list is None else rec[list(src_list)].copy()
        return tx, rec

    def misfit_and_gradient(self, m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
                            misfit_type="l2"):
        tx, rec = self._lin(geom, m.shape, src_list)
        J, g = self.mod.misfit_and_gradient(np.ascontiguousarray(m, float), float(h), float(dt), int(nt),
                                            tx, rec, np.ascontiguousarray(wavelet, float),
                                            np.ascontiguousarray(dobs, float),
                                            {"l2": 0, "gcn": 1}[misfit_type])
        return float(J), g

    def misfit(self, m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
               misfit_type="l2"):
        return self.misfit_and_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge, src_list,
                                        misfit_type)[0]


def test_backend_speedup_3d_inversion():
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    if ref_dir not in sys.path:
        sys.path.insert(0, ref_dir)
    try:
        import _uap
    except ImportError:
        pytest.skip("reference C++ core not built (oracle/_ref)")
    from openusct_amd import fwi, phantom
    from openusct_amd.geometry import CylinderArray
    from openusct_amd.sources import ricker
    h = 1.25e-3
    cyl = CylinderArray(n_rings=3, per_ring=12, radius_m=0.015, height_m=0.014, domain_m=0.036, h=h)
    n = cyl.n
    dt, nt = 1.2e-7, 300
    w = ricker(nt, dt, 0.3e6)
    c_bg = phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.012, h)
    c_true = phantom.add_sphere(c_bg, (0.56, 0.5, 0.5), 0.005, 2650.0, h)
    src = list(range(0, cyl.n_elements, 3))
    ref = _RefCore(_uap)
    tx, rec = ref._lin(cyl, c_true.shape, src)
    dobs = _uap.forward_fmc(phantom.velocity_to_m(c_true), h, dt, nt, tx, rec, w)
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n].astype(float) * h
    cc = (n - 1) * h / 2
    mask = (np.hypot(xx - cc, yy - cc) <= 0.012 * 0.95).astype(float)
    common = dict(src_list=src, n_iter=6, step_frac=0.04, update_mask=mask,
                  m_bounds=(phantom.velocity_to_m(3600.0), phantom.velocity_to_m(2500.0)))
    m0 = phantom.velocity_to_m(c_bg)
    print(f"[case] backend speed-up: {n}^3, {len(src)} transmitters, nt={nt}", flush=True)
    fwi.invert(m0, cyl, w, dt, h, nt, dobs, n_iter=1, src_list=src)  # plan + tuning outside the clock
    t = time.perf_counter()
    m_dev, h_dev = fwi.invert(m0, cyl, w, dt, h, nt, dobs, **common)
    t_dev = time.perf_counter() - t
    t = time.perf_counter()
    m_cpp, h_cpp = fwi.invert(m0, cyl, w, dt, h, nt, dobs, backend=ref, **common)
    t_cpp = time.perf_counter() - t
    err = float(np.max(np.abs(m_dev - m_cpp)) / np.max(np.abs(m_cpp)))
    print(f"{len(h_cpp) - 1} iterations; final misfit reference core {h_cpp[-1]:.4e}, device "
          f"{h_dev[-1]:.4e}; model agreement {err:.2e}; wall clock reference core {t_cpp:.2f} s, "
          f"device {t_dev:.2f} s, speed-up {t_cpp / t_dev:.1f}x")
    assert len(h_dev) == len(h_cpp)
    assert err < 1e-5
    fwi.clear_plans()
