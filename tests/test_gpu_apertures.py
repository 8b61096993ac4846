"""GPU: finite-aperture elements (footprints) and time windows against the
reference's golden vectors (tests/golden/aperture_2d.npz, aperture_3d.npz,
written by simulation/ringfwi/fwi.py itself) and the FP64 oracle.

Tolerance: relative L2 <= 1e-5 for traces, misfits and gradients, as for point
elements.  Between the device's own plans (checkpointed / whole history) the
results are equal bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import IdxGeom, footprints_from, golden, rel_l2

TOL = 1e-5

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_plans():
    from openusct_amd import fwi
    yield
    fwi.clear_plans()
    fwi.METHOD = "batched"


def _aperture():
    f = golden("aperture_2d.npz")
    geom = IdxGeom(f["idx"])
    fps = footprints_from(f["fp_cells"], f["fp_counts"], f["fp_w"])
    return f, geom, fps, float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"], \
        [int(s) for s in f["src_list"]]


@pytest.mark.parametrize("tag", ["nosponge", "sponge"])
@pytest.mark.parametrize("method", ["batched", "plain"])
def test_apertures_2d_golden(method, tag):
    from openusct_amd import fwi
    fwi.METHOD = method
    f, geom, fps, h, dt, nt, w, src = _aperture()
    sp = f["sponge"] if tag == "sponge" else None
    print(f"[case] apertures 2D {tag} method={method}", flush=True)
    d = fwi.forward_fmc(f["m_true"], geom, w, dt, h, nt, sp, src, footprints=fps)
    e_d = rel_l2(d, f[f"{tag}_dobs"])
    errs = [e_d]
    for mt in ("l2", "gcn"):
        J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f[f"{tag}_dobs"], sp, src,
                                       mt, footprints=fps)
        e_J = abs(J - float(f[f"{tag}_{mt}_J"])) / abs(float(f[f"{tag}_{mt}_J"]))
        e_g = rel_l2(g, f[f"{tag}_{mt}_g"])
        print(f"{mt}: traces {e_d:.2e} J {e_J:.2e} gradient {e_g:.2e}")
        errs += [e_J, e_g]
    assert max(errs) < TOL


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_apertures_3d_golden(method):
    from openusct_amd import fwi
    fwi.METHOD = method
    f = golden("aperture_3d.npz")
    geom = IdxGeom(f["idx"])
    fps = footprints_from(f["fp_cells"], f["fp_counts"], f["fp_w"])
    src = [int(s) for s in f["src_list"]]
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    print(f"[case] apertures 3D method={method}", flush=True)
    d = fwi.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=fps)
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["dobs"], None, src,
                                   footprints=fps)
    e_d, e_J, e_g = rel_l2(d, f["dobs"]), abs(J - float(f["J"])) / float(f["J"]), rel_l2(g, f["g"])
    print(f"traces {e_d:.2e} J {e_J:.2e} gradient {e_g:.2e}")
    assert e_d < TOL and e_J < TOL and e_g < TOL


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_time_windows_golden(method):
    from openusct_amd import fwi
    fwi.METHOD = method
    f, geom, fps, h, dt, nt, w, src = _aperture()
    tw = list(f["tw"])
    print(f"[case] time windows method={method}", flush=True)
    errs = []
    for mt in ("l2", "gcn"):
        J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src, mt,
                                       time_weights=tw)
        Jm = fwi.misfit(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src, mt, time_weights=tw)
        e = (abs(J - float(f[f"tw_{mt}_J"])) / abs(float(f[f"tw_{mt}_J"])),
             rel_l2(g, f[f"tw_{mt}_g"]),
             abs(Jm - float(f[f"tw_{mt}_misfit"])) / abs(float(f[f"tw_{mt}_misfit"])))
        print(f"{mt}: J {e[0]:.2e} gradient {e[1]:.2e} misfit {e[2]:.2e}")
        errs += list(e)
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src,
                                   "l2", footprints=fps, time_weights=list(f["twf"]))
    e = (abs(J - float(f["twf_J"])) / float(f["twf_J"]), rel_l2(g, f["twf_g"]))
    print(f"windows + footprints: J {e[0]:.2e} gradient {e[1]:.2e}")
    errs += list(e)
    assert max(errs) < TOL


def test_unit_windows_are_the_identity():
    """test_time_weights.py:31-37: unit weights reproduce the unwindowed misfit
    and gradient (here bit for bit: x * 1 = x in the residual kernels)."""
    from openusct_amd import fwi
    f, geom, fps, h, dt, nt, w, src = _aperture()
    J0, g0 = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src)
    ones = [np.ones((nt, geom.n_elements)) for _ in src]
    J1, g1 = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src,
                                     time_weights=ones)
    assert J0 == J1 and np.array_equal(g0, g1)


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_gauss_newton_golden_sponge_and_footprints(method):
    """The reference's Gauss-Newton products (fwi.py:410-431) with a sponge over
    the whole grid and with footprints."""
    from openusct_amd import fwi
    fwi.METHOD = method
    f, geom, fps, h, dt, _, w, src = _aperture()
    nt = int(f["gn_nt"])
    print(f"[case] GN golden sponge / footprints method={method}", flush=True)
    hv = fwi.gn_hessian_vector(f["m0"], geom, w[:nt], dt, h, nt, f["gn_v"], f["sponge"], src)
    e1 = rel_l2(hv, f["gn_sponge_hv"])
    hv = fwi.gn_hessian_vector(f["m0"], geom, w[:nt], dt, h, nt, f["gn_v"], None, src,
                               footprints=fps)
    e2 = rel_l2(hv, f["gn_fp_hv"])
    print(f"Hv sponge {e1:.2e} footprints {e2:.2e}")
    assert e1 < TOL and e2 < TOL


def _groups_plan_problem(n_src=8, nt=300):
    from openusct_amd.geometry import RingArray, build_footprints
    from openusct_amd.plan import linear_indices
    f, _, _, h, dt, _, w, _ = _aperture()
    geom = RingArray(n_elements=24, radius_m=0.024, domain_m=0.06, h=h)
    shape = f["m0"].shape
    fps = build_footprints(geom, 0.004, "rect")
    groups = [(linear_indices(c, shape), wt) for c, wt in fps]
    dev = torch.device("cuda")
    return (f, geom, fps, groups, h, dt, nt, torch.tensor(w[:nt], device=dev),
            torch.tensor(f["m0"], device=dev), torch.tensor(f["m_true"], device=dev), n_src)


def test_groups_checkpointed_and_multipass():
    """Checkpoint windows recompute the footprint sources exactly (bit-identical
    gradient); shots split over passes select their own transmit lists."""
    from openusct_amd.plan import Plan
    f, geom, fps, groups, h, dt, nt, wd, m0, mt, n_src = _groups_plan_problem(n_src=24)
    shape = tuple(m0.shape)
    tx = groups[:n_src]
    with Plan(shape, h, dt, nt, None, None, tx_groups=tx, rec_groups=groups) as full:
        dobs = full.forward_fmc(mt, wd)
        J0, g0 = full.misfit_and_gradient(m0, wd, dobs)
        hv0 = full.gn_hessian_vector(m0, wd, g0 * 1e3)
        assert full.info["n_windows"] == 1
    for K in (7, 50):
        with Plan(shape, h, dt, nt, None, None, tx_groups=tx, rec_groups=groups,
                  ckpt_every=K) as ck:
            print(f"[case] groups checkpoint K={K} {ck.info}", flush=True)
            J1, g1 = ck.misfit_and_gradient(m0, wd, dobs)
            assert torch.equal(J0, J1) and torch.equal(g0, g1), K
            assert torch.equal(ck.gn_hessian_vector(m0, wd, g0 * 1e3), hv0), K
    budget = 60 * shape[0] * shape[1] * 4 * 16 + (8 << 20)
    with Plan(shape, h, dt, nt, None, None, tx_groups=tx, rec_groups=groups,
              mem_budget=budget) as many:
        print(f"[case] groups multi-pass {many.info}", flush=True)
        assert many.info["shot_passes"] > 1
        d2 = many.forward_fmc(mt, wd)
        assert torch.equal(d2, dobs)
        J2, g2 = many.misfit_and_gradient(m0, wd, dobs)
    assert abs(J2.item() - J0.item()) <= 1e-6 * abs(J0.item())
    assert rel_l2(g2.cpu().numpy(), g0.cpu().numpy()) < 1e-6


def test_groups_vs_oracle_many_shots():
    """24 rectangular apertures, 24 shots: traces and gradient against the oracle."""
    from oracle import oracle as O
    from openusct_amd import fwi
    f, geom, fps, groups, h, dt, nt, wd, m0, mt, n_src = _groups_plan_problem(n_src=24)
    w = wd.cpu().numpy()
    src = list(range(n_src))
    print("[case] groups vs oracle, 24 rect apertures", flush=True)
    dobs = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=fps)
    d = fwi.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=fps)
    J_ref, g_ref = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, dobs, None, src,
                                         footprints=fps)
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, dobs, None, src, footprints=fps)
    e = (rel_l2(d, dobs), abs(J - J_ref) / J_ref, rel_l2(g, g_ref))
    print(f"traces {e[0]:.2e} J {e[1]:.2e} gradient {e[2]:.2e}")
    assert max(e) < TOL


def test_invert_with_footprints_and_windows_reduces_misfit():
    from openusct_amd import fwi
    f, geom, fps, h, dt, nt, w, src = _aperture()
    tw = fwi.p_window_weights(f["nosponge_dobs"], dt, float(f["f0"]))
    m, hist = fwi.invert(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], src_list=src,
                         n_iter=3, footprints=fps, time_weights=tw)
    print(f"[case] invert footprints + windows: {hist}")
    assert hist[-1] < hist[0] and np.all(np.isfinite(m))
