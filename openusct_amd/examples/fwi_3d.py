"""Cylinder-array FWI of a spherical flaw in 3D
(simulation/examples/run_fwi_demo_3d.py:30-115).

A 3000 m/s cylinder of radius 12 mm in 1480 m/s couplant holds a 5 mm,
2650 m/s sphere; 3 rings of 12 elements (radius 15 mm, 14 mm tall) record
from every third element (1.25 mm grid, 300 steps of 120 ns, 300 kHz Ricker).
Eight descent steps with the gradient smoothed (sigma = 1 cell) and updates
inside 95 % of the cylinder radius, bounded to 2500-3600 m/s.
"""

from __future__ import annotations

import argparse
import time

import numpy as np

from .. import fwi, phantom
from ..geometry import CylinderArray
from ..sources import ricker
from . import _plot


def main(out_dir="figures", n_iter=8, verbose=True):
    h, domain = 1.25e-3, 0.036
    geom = CylinderArray(n_rings=3, per_ring=12, radius_m=0.015, height_m=0.014, domain_m=domain, h=h)
    n = geom.n
    dt, nt = 1.2e-7, 300
    wavelet = ricker(nt, dt, 0.3e6)
    c0 = phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.012, h)
    c_true = phantom.add_sphere(c0, (0.56, 0.50, 0.50), 0.005, 2650.0, h)
    m_true, m0 = phantom.velocity_to_m(c_true), phantom.velocity_to_m(c0)
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n].astype(float) * h
    cxy = (n - 1) * h / 2.0
    mask = (np.hypot(xx - cxy, yy - cxy) <= 0.95 * 0.012).astype(float)
    bounds_c = (2500.0, 3600.0)
    bounds = (phantom.velocity_to_m(bounds_c[1]), phantom.velocity_to_m(bounds_c[0]))
    src = list(range(0, geom.n_elements, 3))
    if verbose:
        print(f"grid {n}^3, {geom.n_elements} elements, {len(src)} transmitters, nt={nt}")
    t0 = time.time()
    dobs = fwi.forward_fmc(m_true, geom, wavelet, dt, h, nt, src_list=src)
    t_fwd = time.time() - t0
    t0 = time.time()
    m_rec, history = fwi.invert(m0, geom, wavelet, dt, h, nt, dobs, src_list=src, n_iter=n_iter,
                                step_frac=0.04, update_mask=mask, m_bounds=bounds, smooth_sigma=1.0,
                                verbose=verbose)
    t_inv = time.time() - t0
    c_rec = phantom.m_to_velocity(m_rec)
    iz, iy, ix = (int(round(f * (n - 1))) for f in (0.50, 0.50, 0.56))
    out = {"history": history, "c_flaw": float(c_rec[iz, iy, ix]), "c_rec": c_rec,
           "forward_s": t_fwd, "invert_s": t_inv, "figure": None}
    if verbose:
        print(f"forward FMC {t_fwd:.2f} s, inversion {t_inv:.2f} s; misfit "
              f"{history[-1] / history[0] * 100:.1f} % of initial, flaw {out['c_flaw']:.0f} m/s")
    plt = _plot.pyplot()
    if plt is not None and out_dir:
        # one row: the true and reconstructed axial plane through the flaw, the
        # reconstruction's two vertical planes through it, and the misfit curve
        panels = [(c_true[iz], f"true, z index {iz}"), (c_rec[iz], f"FWI, z index {iz}"),
                  (c_rec[:, iy, :], f"FWI, y index {iy}"), (c_rec[:, :, ix], f"FWI, x index {ix}")]
        fig = _plot.velocity_panels(plt, panels, [0, domain * 1e3, 0, domain * 1e3],
                                    bounds_c[0], bounds_c[1], history)
        out["figure"] = _plot.save(fig, out_dir, "reconstruction_3d.png")
        plt.close(fig)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out-dir", default="figures")
    ap.add_argument("--n-iter", type=int, default=8)
    a = ap.parse_args()
    main(a.out_dir, a.n_iter)
