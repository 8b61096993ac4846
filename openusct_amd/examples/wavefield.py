"""Pressure snapshots of one transmit (simulation/examples/wavefield_snapshot.py:25-55):
element 0 of a 16-element ring fires into the 2D specimen of the FWI
demonstration; the field is shown at steps 180, 360 and 560 as the wavefront
crosses the specimen and scatters off the inclusion.
"""

from __future__ import annotations

import argparse

import numpy as np

from .. import fwi, phantom
from ..geometry import RingArray
from ..sources import ricker
from . import _plot

FRAMES = (180, 360, 560)


def main(out_dir="figures", verbose=True):
    h, domain = 6.0e-4, 0.06
    geom = RingArray(n_elements=16, radius_m=0.024, domain_m=domain, h=h)
    n = geom.n
    dt, nt = 7.0e-8, 800
    wavelet = ricker(nt, dt, 0.25e6)
    c = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.020, h)
    c = phantom.add_inclusion(c, (0.58, 0.46), 0.007, 2650.0, h)
    hist = fwi.simulate_wavefield(phantom.velocity_to_m(c), geom, 0, wavelet, dt, h, nt)
    out = {"frames": {fr: np.array(hist[fr]) for fr in FRAMES}, "figure": None}
    if verbose:
        print("peak |p| at " + ", ".join(f"step {fr}: {np.abs(hist[fr]).max():.3e}" for fr in FRAMES))
    plt = _plot.pyplot()
    if plt is not None and out_dir:
        fig = _plot.snapshots(plt, out["frames"], [0, domain * 1e3, 0, domain * 1e3], geom.xy, dt)
        out["figure"] = _plot.save(fig, out_dir, "wavefield.png")
        plt.close(fig)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out-dir", default="figures")
    a = ap.parse_args()
    main(a.out_dir)
