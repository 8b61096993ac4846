"""The acquisition loop through the dataset file
(simulation/examples/run_pipeline_demo.py:31-87): simulate a full-matrix
capture on the device, write it as an HDF5 dataset with the native writer,
read it back and reconstruct it through the "fwi" plugin -- the same call a
hardware acquisition would take.

16-element ring (radius 20 mm) on a 1 mm grid, 500 steps of 80 ns, 300 kHz
Ricker; a 5 mm, 2600 m/s flaw in a 3000 m/s specimen of radius 16 mm.
"""

from __future__ import annotations

import argparse
import os

from .. import phantom, plugins
from ..acquire import simulate_dataset
from ..dataset import Dataset
from ..geometry import RingArray
from ..sources import ricker
from . import _plot


def main(out_dir="figures", n_iter=12, verbose=True):
    h, domain = 1.0e-3, 0.050
    ring = RingArray(n_elements=16, radius_m=0.020, domain_m=domain, h=h)
    n = ring.n
    dt, nt = 8.0e-8, 500
    wavelet = ricker(nt, dt, 0.3e6)
    c_bg = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.016, h)
    c_true = phantom.add_inclusion(c_bg, (0.60, 0.50), 0.005, 2600.0, h)
    if verbose:
        print("1. simulating the acquisition on the device")
    ds = simulate_dataset(c_true, ring, wavelet, dt, nominal_speed_m_s=3000.0,
                          description="pipeline demo specimen with flaw")
    os.makedirs(out_dir or ".", exist_ok=True)
    path = os.path.join(out_dir or ".", "acquisition.h5")
    ds.save(path)
    loaded = Dataset.load(path)
    if verbose:
        print(f"2. saved and read back {path} (data {loaded.data.shape}, "
              f"{loaded.geometry.dim}D {loaded.geometry.array_type})")
        print("3. reconstructing with the 'fwi' plugin")
    bounds = (phantom.velocity_to_m(3600.0), phantom.velocity_to_m(2500.0))
    result = plugins.run("fwi", loaded, start_c=c_bg, m_bounds=bounds, n_iter=n_iter,
                         step_frac=0.03, verbose=verbose)
    history = result["history"]
    iy, ix = int(round(0.50 * (n - 1))), int(round(0.60 * (n - 1)))
    out = {"history": history, "c_flaw": float(result["c"][iy, ix]), "dataset": path,
           "figure": None}
    if verbose:
        print(f"   misfit {history[-1] / history[0] * 100:.1f} % of initial, flaw "
              f"{out['c_flaw']:.0f} m/s (true 2600)")
    plt = _plot.pyplot()
    if plt is not None and out_dir:
        lo, hi = result["extent"]
        fig = _plot.velocity_panels(plt, [(c_true, "True model"), (c_bg, "Starting model"),
                                          (result["c"], "FWI via pipeline")],
                                    [lo * 1e3, hi * 1e3, lo * 1e3, hi * 1e3], 2500, 3100, history)
        out["figure"] = _plot.save(fig, out_dir, "pipeline.png")
        plt.close(fig)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out-dir", default="figures")
    ap.add_argument("--n-iter", type=int, default=12)
    a = ap.parse_args()
    main(a.out_dir, a.n_iter)
