"""GPU: the envelope, normalised-envelope (egcn), traveltime and graph-space
optimal transport (gsot) misfits (simulation/ringfwi/fwi.py:134-243) through
the device -- forward traces, adjoint source on the device (gsot: the
library's assignment kernel), the library's adjoint sweep on it
(usct_gradient_from_source).

Against the reference's own misfit / misfit_and_gradient on the problem of
fmc_2d.npz (tests/golden/misfits_2d.npz): relative L2 <= 1e-5 for the gradient
and misfit, as for L2.  The directional finite-difference checks of
test_gradient_misfits.py:38-74 with the reference's tolerances (1e-3 for the
envelope misfits, 0.15 for the stationary-phase traveltime adjoint, 1e-2 for
gsot); the step is 1e-4 of m (the reference's 1e-6 assumes FP64 fields), for
gsot the reference's 1e-7, which keeps the assignment fixed.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import IdxGeom, golden, rel_l2

TOL = 1e-5

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_plans():
    from openusct_amd import fwi
    yield
    fwi.clear_plans()
    fwi.METHOD = "batched"


def _problem():
    f = golden("fmc_2d.npz")
    return (f, IdxGeom(f["idx"]), float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"],
            [int(s) for s in f["src_list"]])


@pytest.mark.parametrize("method", ["batched", "plain"])
@pytest.mark.parametrize("mt", ["envelope", "egcn", "traveltime", "gsot"])
def test_misfit_gradient_vs_reference(mt, method):
    from openusct_amd import fwi
    f, geom, h, dt, nt, w, src = _problem()
    g_ref = golden("misfits_2d.npz")
    fwi.METHOD = method
    print(f"[case] {mt} {method}", flush=True)
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src, mt)
    Jm = fwi.misfit(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src, mt)
    e = (abs(J - float(g_ref[f"{mt}_J"])) / abs(float(g_ref[f"{mt}_J"])), rel_l2(g, g_ref[f"{mt}_g"]),
         abs(Jm - float(g_ref[f"{mt}_misfit"])) / abs(float(g_ref[f"{mt}_misfit"])))
    print(f"J {e[0]:.2e} gradient {e[1]:.2e} misfit {e[2]:.2e}")
    assert max(e) < TOL


def test_windowed_envelope_and_sponge_traveltime_vs_reference():
    from openusct_amd import fwi
    f, geom, h, dt, nt, w, src = _problem()
    g_ref = golden("misfits_2d.npz")
    print("[case] windowed envelope, sponge traveltime", flush=True)
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src,
                                   "envelope", time_weights=list(g_ref["tw"]))
    e1 = (abs(J - float(g_ref["tw_envelope_J"])) / float(g_ref["tw_envelope_J"]),
          rel_l2(g, g_ref["tw_envelope_g"]))
    J, g = fwi.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["sponge_dobs"], g_ref["sponge"],
                                   src, "traveltime")
    e2 = (abs(J - float(g_ref["sponge_traveltime_J"])) / float(g_ref["sponge_traveltime_J"]),
          rel_l2(g, g_ref["sponge_traveltime_g"]))
    print(f"windows {e1[0]:.2e} {e1[1]:.2e}; sponge {e2[0]:.2e} {e2[1]:.2e}")
    assert max(e1 + e2) < TOL


@pytest.mark.parametrize("mt,tol,eps", [("envelope", 1e-3, 1e-4), ("egcn", 1e-3, 1e-4),
                                        ("traveltime", 0.15, 1e-4), ("gsot", 1e-2, 1e-7)])
def test_directional_finite_difference(mt, tol, eps):
    """test_gradient_misfits.py:21-74 on the device."""
    from openusct_amd import fwi, phantom
    from openusct_amd.geometry import RingArray
    from openusct_amd.sources import ricker
    h = 1.0e-3
    geom = RingArray(n_elements=8, radius_m=0.024, domain_m=0.06, h=h)
    n = geom.n
    dt, nt = 1.2e-7, 350
    wav = ricker(nt, dt, 0.4e6)
    c_true = phantom.add_inclusion(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h),
                                   (0.5, 0.5), 0.006, 3400.0, h)
    m0 = phantom.velocity_to_m(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h))
    src = [0, 4]
    dobs = fwi.forward_fmc(phantom.velocity_to_m(c_true), geom, wav, dt, h, nt, src_list=src)
    rng = np.random.default_rng(1)
    _, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, dobs, src_list=src, misfit_type=mt)
    dm = rng.standard_normal((n, n))
    dm[:5, :] = dm[-5:, :] = dm[:, :5] = dm[:, -5:] = 0.0
    dm *= float(np.mean(m0))
    analytic = float(np.sum(g * dm))
    Jp = fwi.misfit(m0 + eps * dm, geom, wav, dt, h, nt, dobs, src_list=src, misfit_type=mt)
    Jm = fwi.misfit(m0 - eps * dm, geom, wav, dt, h, nt, dobs, src_list=src, misfit_type=mt)
    fd = (Jp - Jm) / (2 * eps)
    rel = abs(analytic - fd) / (abs(fd) + 1e-30)
    print(f"[case] FD {mt}: analytic {analytic:.6e} fd {fd:.6e} rel {rel:.2e}")
    assert rel < tol


def test_gradient_from_source_matches_fused_l2():
    """The adjoint-source entry point with the L2 residual as its source gives
    the fused L2 gradient (the residual is formed in FP64 here and in FP32 hi +
    lo in the library, hence agreement to rounding, not bits), on batched and
    plain plans, and on a multi-pass plan."""
    from openusct_amd.plan import Plan, linear_indices
    f, geom, h, dt, nt, w, src = _problem()
    shape = f["m0"].shape
    rec = linear_indices(geom.idx, shape)
    dev = torch.device("cuda")
    md, wd = torch.tensor(f["m0"], device=dev), torch.tensor(w, device=dev)
    tx = np.array([(6 + i) * shape[1] + 9 + i for i in range(40)], dtype=np.int64)
    for kw in (dict(method="batched"), dict(method="plain"),
               dict(method="batched", mem_budget=60 * shape[0] * shape[1] * 4 * 16 + (8 << 20))):
        with Plan(shape, h, dt, nt, tx, rec, **kw) as p:
            print(f"[case] gradient_from_source {kw} {p.info['shot_passes']} pass(es)", flush=True)
            dobs = p.forward_fmc(torch.tensor(f["m_true"], device=dev), wd)
            d0 = p.forward_fmc(md, wd)
            _, g_l2 = p.misfit_and_gradient(md, wd, dobs)
            g = p.gradient_from_source(md, wd, d0 - dobs)
            e = rel_l2(g.cpu().numpy(), g_l2.cpu().numpy())
            print(f"rel {e:.2e}")
            assert e < 1e-6


def test_invert_with_envelope_reduces_misfit():
    from openusct_amd import fwi
    f, geom, h, dt, nt, w, src = _problem()
    print("[case] invert envelope", flush=True)
    m, hist = fwi.invert(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], src_list=src, n_iter=3,
                         misfit_type="envelope")
    print(hist)
    assert hist[-1] < hist[0] and np.all(np.isfinite(m))


def test_gsot_flags_non_finite_traces():
    from openusct_amd import misfits
    d = torch.randn(2, 40, 3, dtype=torch.float64, device="cuda")
    o = torch.randn(2, 40, 3, dtype=torch.float64, device="cuda")
    J, adj = misfits.gsot(d, o)
    assert torch.isfinite(J) and torch.isfinite(adj).all()
    d[1, 7, 2] = float("nan")
    with pytest.raises(ValueError):
        misfits.gsot(d, o)


def test_traces_autograd_gives_the_device_adjoint():
    """autograd.traces: a misfit written in torch on the device traces gets its
    m-gradient from the device adjoint -- the L2 misfit reproduces the fused
    L2 gradient, the envelope misfit through torch's FFT autograd the
    reference-form envelope adjoint source (whose regularised envelope differs
    only at 1e-12 of the largest envelope)."""
    from openusct_amd import autograd as A
    from openusct_amd import misfits
    from openusct_amd.plan import Plan, linear_indices
    f, geom, h, dt, nt, w, src = _problem()
    shape = f["m0"].shape
    rec = linear_indices(geom.idx, shape)
    dev = torch.device("cuda")
    md, wd = torch.tensor(f["m0"], device=dev), torch.tensor(w, device=dev)
    dobs = torch.tensor(f["nosponge_dobs"], device=dev)
    print("[case] traces autograd: L2 and envelope", flush=True)
    with Plan(shape, h, dt, nt, rec[src], rec) as p:
        m = md.clone().requires_grad_(True)
        J = 0.5 * torch.sum((A.traces(m, p, wd) - dobs) ** 2)
        J.backward()
        _, g_l2 = p.misfit_and_gradient(md, wd, dobs)
        e1 = rel_l2(m.grad.cpu().numpy(), g_l2.cpu().numpy())
        m2 = md.clone().requires_grad_(True)
        env = torch.abs(misfits.analytic(A.traces(m2, p, wd)))
        J2 = 0.5 * torch.sum((env - torch.abs(misfits.analytic(dobs))) ** 2)
        J2.backward()
        J3, g3 = A._misfit_and_gradient(p, md, wd, dobs, None, "envelope")
        e2 = rel_l2(m2.grad.cpu().numpy(), g3.cpu().numpy())
        print(f"L2 {e1:.2e} envelope {e2:.2e} (J {J2.item():.6e} vs {J3.item():.6e})")
        assert e1 < 1e-6 and e2 < 1e-5
        c0 = torch.sqrt(1.0 / md)
        _, hist = A.gradient_descent(c0, p, wd, dobs, steps=2, step_frac=0.005, misfit_type="gsot")
        print("gsot descent", hist)
        assert hist[-1] <= hist[0]


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_envelope_with_footprints_vs_reference(method):
    """Finite-aperture elements through the adjoint-source path (the source
    spread over each receiver's footprint, fwi.py:316-319)."""
    from helpers import footprints_from
    from openusct_amd import fwi
    a = golden("aperture_2d.npz")
    g_ref = golden("misfits_2d.npz")
    fps = footprints_from(a["fp_cells"], a["fp_counts"], a["fp_w"])
    fwi.METHOD = method
    print(f"[case] envelope + footprints {method}", flush=True)
    J, g = fwi.misfit_and_gradient(a["m0"], IdxGeom(a["idx"]), a["wavelet"], float(a["dt"]),
                                   float(a["h"]), int(a["nt"]), a["nosponge_dobs"], None,
                                   [int(s) for s in a["src_list"]], "envelope", footprints=fps)
    e = (abs(J - float(g_ref["fp_envelope_J"])) / float(g_ref["fp_envelope_J"]),
         rel_l2(g, g_ref["fp_envelope_g"]))
    print(f"J {e[0]:.2e} gradient {e[1]:.2e}")
    assert max(e) < TOL
