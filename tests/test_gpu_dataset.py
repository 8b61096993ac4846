"""GPU: the acquisition chain around the solver -- simulate_dataset on the
device, the dataset file written and read by the native HDF5 library, the
channel data streamed into HBM, and FWI through the plugin on the loaded
dataset (simulation/tests/test_pipeline.py:23-58, test_dataset.py:22-86).

The traces a simulated dataset holds are the device forward's own (bit for
bit); on disk they are float32, as the reference stores them, and read back
bit-exactly.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_plans():
    from openusct_amd import fwi
    yield
    fwi.clear_plans()


def _ring_case():
    from openusct_amd import phantom
    from openusct_amd.geometry import RingArray
    from openusct_amd.sources import ricker
    h = 1.25e-3
    ring = RingArray(n_elements=12, radius_m=0.020, domain_m=0.050, h=h)
    n = ring.n
    dt, nt = 1.2e-7, 360
    wavelet = ricker(nt, dt, 0.3e6)
    c_bg = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.016, h)
    c_true = phantom.add_inclusion(c_bg, (0.60, 0.50), 0.005, 2600.0, h)
    return ring, n, dt, wavelet, c_bg, c_true


def test_pipeline_simulate_save_load_fwi(tmp_path):
    """test_pipeline.py:23-58 on the device: simulate, write, read back,
    reconstruct through the "fwi" plugin; the misfit halves and the flaw
    comes back slower than the background."""
    from openusct_amd import phantom, plugins
    from openusct_amd.acquire import simulate_dataset
    from openusct_amd.dataset import Dataset
    ring, n, dt, wavelet, c_bg, c_true = _ring_case()
    print("[case] pipeline 12-element ring, 41^2, nt 360", flush=True)
    ds = simulate_dataset(c_true, ring, wavelet, dt, nominal_speed_m_s=3000.0)
    path = str(tmp_path / "acq.h5")
    ds.save(path)
    loaded = Dataset.load(path)
    m_bounds = (phantom.velocity_to_m(3600.0), phantom.velocity_to_m(2500.0))
    result = plugins.run("fwi", loaded, start_c=c_bg, m_bounds=m_bounds, n_iter=6, step_frac=0.04)
    hist = result["history"]
    ix, iy = int(round(0.60 * (n - 1))), int(round(0.50 * (n - 1)))
    print(f"misfit {hist[-1] / hist[0] * 100:.1f}%, flaw {result['c'][iy, ix]:.0f} m/s")
    assert hist[-1] < 0.5 * hist[0]
    assert result["c"][iy, ix] < 2900.0


def test_simulated_dataset_holds_the_device_traces(tmp_path):
    from openusct_amd import fwi, phantom
    from openusct_amd.acquire import peak_frequency, simulate_dataset
    from openusct_amd.dataset import Dataset
    ring, n, dt, wavelet, _, c_true = _ring_case()
    src = [0, 3, 7]
    print("[case] simulate_dataset fields", flush=True)
    ds = simulate_dataset(c_true, ring, wavelet, dt, 1480.0, src_list=src, description="t")
    d = fwi.forward_fmc(phantom.velocity_to_m(c_true), ring, wavelet, dt, ring.h, len(wavelet), src_list=src)
    assert np.array_equal(ds.data, d)
    assert list(ds.tx_elements) == src and ds.n_tx == 3 and ds.n_rx == 12
    assert ds.tx_centre_freq_hz == peak_frequency(wavelet, dt) == ds.geometry.centre_freq_hz
    assert np.array_equal(ds.geometry.element_pos, ring.element_positions)
    assert ds.ground_truth["h_m"] == ring.h and ds.metadata["created_by"] == "OpenUSCT simulator"
    path = str(tmp_path / "s.h5")
    ds.save(path)
    back = Dataset.load(path)
    assert np.array_equal(back.data, ds.data.astype(np.float32))
    assert back.metadata["description"] == "t"


def test_simulated_3d_dataset_roundtrip(tmp_path):
    """test_dataset.py:56-86: a cylinder acquisition through the file."""
    from openusct_amd import phantom
    from openusct_amd.acquire import simulate_dataset
    from openusct_amd.dataset import Dataset
    from openusct_amd.geometry import CylinderArray
    from openusct_amd.sources import ricker
    h = 2.0e-3
    cyl = CylinderArray(n_rings=2, per_ring=6, radius_m=0.012, height_m=0.010, domain_m=0.03, h=h)
    n = cyl.n
    dt, nt = 2.0e-7, 120
    c = phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.009, h)
    c = phantom.add_sphere(c, (0.5, 0.5, 0.5), 0.004, 2700.0, h)
    print("[case] 3D dataset round trip", flush=True)
    ds = simulate_dataset(c, cyl, ricker(nt, dt, 0.3e6), dt, nominal_speed_m_s=1480.0, src_list=[0, 6])
    assert ds.geometry.dim == 3 and ds.geometry.element_pos.shape == (12, 3)
    assert float(np.abs(ds.data).max()) > 0.0
    path = str(tmp_path / "ds3d.h5")
    ds.save(path)
    loaded = Dataset.load(path)
    assert loaded.geometry.array_type == "cylinder" and loaded.ground_truth["c"].ndim == 3
    assert np.array_equal(loaded.data, ds.data.astype(np.float32))
    assert np.array_equal(loaded.ground_truth["c"], c)


def test_read_fmc_streams_into_device_memory(tmp_path):
    """read_fmc: double-buffered pinned staging (several blocks), transmit
    subsets and the float32 / float64 device types equal the host read."""
    from openusct_amd.dataset import ArrayGeometry, Dataset, read_fmc
    rng = np.random.default_rng(8)
    data = rng.standard_normal((37, 500, 24))
    ds = Dataset(ArrayGeometry.ring(24, 0.02, 5e5), data, 2e7, rng.standard_normal(500), 5e5, 1480.0)
    path = str(tmp_path / "big.h5")
    ds.save(path)
    host = Dataset.load(path).data
    print("[case] read_fmc 37 x 500 x 24 in blocks of 3 transmits", flush=True)
    dev = read_fmc(path, "cuda", block_bytes=3 * 500 * 24 * 4)
    torch.cuda.synchronize()
    assert dev.dtype == torch.float64 and dev.is_cuda
    assert torch.equal(dev.cpu(), torch.from_numpy(host.astype(np.float64)))
    sub = read_fmc(path, "cuda", dtype=torch.float32, rows=(5, 19), block_bytes=4 * 500 * 24 * 4)
    assert torch.equal(sub.cpu(), torch.from_numpy(host[5:19]))
    assert read_fmc(path, "cuda", rows=(4, 4)).shape == (0, 500, 24)
    with pytest.raises(ValueError):
        read_fmc(path, "cuda", rows=(30, 40))
