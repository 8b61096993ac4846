"""GPU: the reference's demonstrations of the acoustic path
(simulation/examples/run_fwi_demo.py, run_fwi_demo_3d.py, run_pipeline_demo.py,
wavefield_snapshot.py) run end to end through openusct_amd.examples with the
reference's settings.  They check what the demonstrations show: the misfit
falls, the flaw comes back slower than the specimen, the wavefield is finite
and travels, and every figure is written.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_plans():
    from openusct_amd import fwi
    yield
    fwi.clear_plans()


def test_fwi_2d_demo(tmp_path):
    from openusct_amd.examples import fwi_2d
    print("[case] examples.fwi_2d (101^2, 16 elements, nt 800, 12 iterations)", flush=True)
    r = fwi_2d.main(str(tmp_path), verbose=True)
    h = r["history"]
    assert h[-1] < 0.5 * h[0]
    assert r["c_flaw"] < 2950.0  # true 2650 inside a 3000 m/s specimen
    assert os.path.getsize(r["figure"]) > 0


def test_fwi_3d_demo(tmp_path):
    from openusct_amd.examples import fwi_3d
    print("[case] examples.fwi_3d (cylinder array, 12 transmitters, nt 300, 8 iterations)",
          flush=True)
    r = fwi_3d.main(str(tmp_path), verbose=True)
    h = r["history"]
    assert h[-1] < 0.8 * h[0]
    assert r["c_flaw"] < 2990.0
    assert os.path.getsize(r["figure"]) > 0


def test_pipeline_demo(tmp_path):
    from openusct_amd.dataset import Dataset
    from openusct_amd.examples import pipeline
    print("[case] examples.pipeline (simulate -> HDF5 -> plugin FWI)", flush=True)
    r = pipeline.main(str(tmp_path), verbose=True)
    h = r["history"]
    assert h[-1] < 0.5 * h[0]
    assert r["c_flaw"] < 2900.0
    ds = Dataset.load(r["dataset"])
    assert ds.data.shape == (16, 500, 16) and ds.data.dtype == np.float32
    assert os.path.getsize(r["figure"]) > 0


def test_wavefield_demo(tmp_path):
    from openusct_amd.examples import wavefield
    print("[case] examples.wavefield (one transmit, 800 steps)", flush=True)
    r = wavefield.main(str(tmp_path), verbose=True)
    peaks = []
    for fr, p in r["frames"].items():
        assert np.all(np.isfinite(p))
        peaks.append(np.unravel_index(np.argmax(np.abs(p)), p.shape))
        assert np.abs(p).max() > 0.0
    assert len(set(peaks)) == len(peaks)  # the wavefront moves between the frames
    assert os.path.getsize(r["figure"]) > 0
