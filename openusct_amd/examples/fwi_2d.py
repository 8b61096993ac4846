"""Ring-array FWI of a flaw in 2D (simulation/examples/run_fwi_demo.py:29-110).

A 3000 m/s specimen of radius 20 mm in a 1480 m/s couplant holds a 7 mm,
2650 m/s inclusion; 16 elements on a 24 mm ring record a full-matrix capture
(0.6 mm grid, 800 steps of 70 ns, 250 kHz Ricker).  Starting from the
background without the inclusion, steepest descent on the device recovers it
from the traces alone; updates are confined to 95 % of the specimen radius and
bounded to 2600-4200 m/s.
"""

from __future__ import annotations

import argparse
import time

import numpy as np

from .. import fwi, phantom
from ..geometry import RingArray
from ..sources import ricker
from . import _plot


def main(out_dir="figures", n_iter=12, verbose=True):
    h, domain = 6.0e-4, 0.06
    geom = RingArray(n_elements=16, radius_m=0.024, domain_m=domain, h=h)
    n = geom.n
    dt, nt = 7.0e-8, 800
    wavelet = ricker(nt, dt, 0.25e6)
    c_spec, c_coup, r_spec = 3000.0, 1480.0, 0.020
    c_start = phantom.coupling_background((n, n), c_spec, c_coup, r_spec, h)
    c_true = phantom.add_inclusion(c_start, (0.58, 0.46), 0.007, 2650.0, h)
    m_true, m0 = phantom.velocity_to_m(c_true), phantom.velocity_to_m(c_start)
    yy, xx = np.mgrid[0:n, 0:n] * h
    centre = (n - 1) * h / 2.0
    mask = (np.hypot(xx - centre, yy - centre) <= 0.95 * r_spec).astype(float)
    bounds = (phantom.velocity_to_m(4200.0), phantom.velocity_to_m(2600.0))
    if verbose:
        print(f"grid {n}x{n}, {geom.n_elements} elements, nt={nt}")
    t0 = time.time()
    dobs = fwi.forward_fmc(m_true, geom, wavelet, dt, h, nt)
    t_fwd = time.time() - t0
    t0 = time.time()
    m_rec, history = fwi.invert(m0, geom, wavelet, dt, h, nt, dobs, n_iter=n_iter, step_frac=0.03,
                                update_mask=mask, m_bounds=bounds, verbose=verbose)
    t_inv = time.time() - t0
    c_rec = phantom.m_to_velocity(m_rec)
    iy, ix = int(round(0.46 * (n - 1))), int(round(0.58 * (n - 1)))
    out = {"history": history, "c_flaw": float(c_rec[iy, ix]), "c_rec": c_rec,
           "forward_s": t_fwd, "invert_s": t_inv, "figure": None}
    if verbose:
        print(f"forward FMC {t_fwd:.2f} s, inversion ({len(history)} iterations) {t_inv:.2f} s; "
              f"misfit {history[-1] / history[0] * 100:.1f} % of initial, flaw "
              f"{out['c_flaw']:.0f} m/s (true 2650)")
    plt = _plot.pyplot()
    if plt is not None and out_dir:
        fig = _plot.velocity_panels(plt, [(c_true, "True model"), (c_start, "Starting model"),
                                          (c_rec, "FWI reconstruction")],
                                    [0, domain * 1e3, 0, domain * 1e3], 2600, 3100, history,
                                    marks=geom.xy)
        out["figure"] = _plot.save(fig, out_dir, "reconstruction.png")
        plt.close(fig)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out-dir", default="figures")
    ap.add_argument("--n-iter", type=int, default=12)
    a = ap.parse_args()
    main(a.out_dir, a.n_iter)
