"""GPU: optimiser trajectories of fwi.invert (fwi.py:535-667) against the
reference's own, recorded in tests/golden/invert.npz by make_golden.py: steepest
descent with the masked, Gaussian-smoothed gradient (scipy.ndimage, as the
reference), bounds and the halving line search -- L2, GCN, with a border
sponge, in 3D -- and truncated Gauss-Newton (fwi.py:434-531).

The iterates follow from gradients that match the reference's to about 1e-6,
so the line search takes the same decisions (same number of iterations) and
the misfit history agrees to 1e-5 relative; the model update m - m0
compounds the gradient error over the iterations and is held to 1e-4.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import IdxGeom, golden, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_plans():
    from openusct_amd import fwi
    yield
    fwi.clear_plans()
    fwi.METHOD = "batched"


def _case(name):
    from openusct_amd import phantom
    inv = golden("invert.npz")
    bounds = (phantom.velocity_to_m(3600.0), phantom.velocity_to_m(2500.0))
    if name == "gd_sponge":
        f = golden("sponge_2d.npz")
        return f, f["dobs"], f["sponge"], dict(n_iter=3, step_frac=0.02), inv
    if name == "gd3d":
        f = golden("fmc_3d.npz")
        return f, f["dobs"], None, dict(n_iter=3, step_frac=0.005), inv
    f = golden("fmc_2d.npz")
    kw = {"gd": dict(n_iter=3, step_frac=0.02),
          "gd_mask": dict(n_iter=3, step_frac=0.03, update_mask=inv["mask"], m_bounds=bounds),
          "gd_gcn": dict(n_iter=3, step_frac=0.02, misfit_type="gcn"),
          "gn": dict(n_iter=2, optimizer="gn", n_cg=4, update_mask=inv["mask"],
                     m_bounds=bounds)}[name]
    return f, f["nosponge_dobs"], None, kw, inv


@pytest.mark.parametrize("name", ["gd", "gd_mask", "gd_gcn", "gn", "gd_sponge", "gd3d"])
def test_invert_trajectory_vs_reference(name):
    from openusct_amd import fwi
    f, dobs, sponge, kw, inv = _case(name)
    print(f"[case] invert trajectory {name} {f['m0'].shape} {kw.get('optimizer', 'gd')}", flush=True)
    m0 = f["m0"]
    m, hist = fwi.invert(m0, IdxGeom(f["idx"]), f["wavelet"], float(f["dt"]), float(f["h"]),
                         int(f["nt"]), dobs, sponge, [int(s) for s in f["src_list"]], **kw)
    ref_m, ref_h = inv[f"{name}_m"], inv[f"{name}_history"]
    hist = np.asarray(hist)
    e_h = float(np.max(np.abs(hist - ref_h) / np.abs(ref_h))) if hist.shape == ref_h.shape else np.inf
    e_m = rel_l2(m - m0, ref_m - m0)
    print(f"{len(hist)} misfits (reference {len(ref_h)}), history {e_h:.2e}, update {e_m:.2e}, "
          f"final misfit {hist[-1] / hist[0]:.4f} of initial")
    assert hist.shape == ref_h.shape
    assert e_h < 1e-5 and e_m < 1e-4
