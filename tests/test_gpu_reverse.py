"""GPU: 3D gradients that keep no forward levels.  The adjoint sweep runs the
leapfrog backwards alongside it (k_radj3d: p^{k-2} = 2p^{k-1} - p^k + inc, the
forward's own FP64 update with p^k and p^{k-2} exchanged), so the gradient needs
neither the whole history nor checkpoints and re-computed forward levels.

Against the reference's 3D golden vectors (simulation/tests/test_gradient_3d.py
problem) and the FP64 oracle: relative L2 <= 1e-5, the bar every gradient
meets.  Against the device's plans that keep levels: the re-created levels
differ from the forward's own by rounding only (the FP64 update of the stored
hi + lo state run backwards), so gradients agree to <= 1e-6, not bit for bit;
between reversing plans (graph replay / eager) they agree bit for bit.
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


def _golden_plan_inputs(name, nt=None):
    from openusct_amd.plan import linear_indices
    f = golden(name)
    geom = IdxGeom(f["idx"])
    shape = f["m0"].shape
    rec = linear_indices(geom.idx, shape)
    src = [int(s) for s in f["src_list"]]
    dev = torch.device("cuda")
    nt = int(f["nt"]) if nt is None else nt
    return (f, shape, float(f["h"]), float(f["dt"]), nt, rec[src], rec,
            torch.tensor(f["m0"], device=dev), torch.tensor(f["m_true"], device=dev),
            torch.tensor(f["wavelet"][:nt], device=dev))


@pytest.mark.parametrize("name", ["fmc_3d.npz", "fmc_3d_fine.npz"])
def test_reversing_gradient_vs_golden_and_stored_levels(name):
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, tx, rec, m0, mt, wd = _golden_plan_inputs(name)
    dobs = torch.tensor(f["dobs"], device=m0.device)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True) as rp:
        info = rp.info
        print(f"[case] reversing 3D gradient {name} {info}", flush=True)
        assert info["reverse"] == 1 and info["ckpt_every"] == 0
        J, g = rp.misfit_and_gradient(m0, wd, dobs)
        d = rp.forward_fmc(mt, wd)
    with Plan(shape, h, dt, nt, tx, rec, reverse=False) as kp:
        assert kp.info["reverse"] == 0
        Jk, gk = kp.misfit_and_gradient(m0, wd, dobs)
        dk = kp.forward_fmc(mt, wd)
    g, gk = g.cpu().numpy(), gk.cpu().numpy()
    e = (abs(J.item() - float(f["J"])) / float(f["J"]), rel_l2(g, f["g"]), rel_l2(g, gk))
    print(f"J {e[0]:.2e} gradient vs golden {e[1]:.2e} vs stored levels {e[2]:.2e}")
    assert torch.equal(d, dk)  # the forward is the same launch sequence
    assert torch.equal(J, Jk)
    assert e[0] < TOL and e[1] < TOL and e[2] < 1e-6


@pytest.mark.parametrize("nt", [2, 3, 4, 5, 60])
def test_reversing_short_time_axes(nt):
    """The first levels (k <= 2 re-create levels 0 and -1, which stay the zero
    level) and the shortest axes against the whole-history plan."""
    from openusct_amd.plan import Plan
    f, shape, h, dt, _, tx, rec, m0, mt, wd = _golden_plan_inputs("fmc_3d.npz", nt=nt)
    with Plan(shape, h, dt, nt, tx, rec, reverse=False) as kp:
        dobs = kp.forward_fmc(mt, wd)
        Jk, gk = kp.misfit_and_gradient(m0, wd, dobs)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True) as rp:
        print(f"[case] reversing nt={nt} {rp.info}", flush=True)
        J, g = rp.misfit_and_gradient(m0, wd, dobs)
    e = rel_l2(g.cpu().numpy(), gk.cpu().numpy())
    print(f"gradient vs stored levels {e:.2e}")
    assert torch.equal(J, Jk)
    assert e < 1e-6


def test_reversing_passes_graph_and_eager():
    """40 shots in several passes against one pass; graph replays against eager
    launches, bit for bit, with eager work between the replays."""
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, _, rec, m0, mt, wd = _golden_plan_inputs("fmc_3d.npz")
    n = shape[0]
    tx = np.array([((3 + i % 15) * n + 4 + i % 13) * n + 5 + i % 11 for i in range(40)], np.int64)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False) as one:
        info1 = one.info
        dobs = one.forward_fmc(mt, wd)
        J1, g1 = one.misfit_and_gradient(m0, wd, dobs)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False,
              mem_budget=int(0.55 * info1["workspace_bytes"])) as many:
        info = many.info
        print(f"[case] reversing passes {info1} -> {info}", flush=True)
        assert info["shot_passes"] > 1 and info["reverse"] == 1
        J2, g2 = many.misfit_and_gradient(m0, wd, dobs)
    assert abs(J1.item() - J2.item()) <= 1e-6 * abs(J1.item())
    assert rel_l2(g2.cpu().numpy(), g1.cpu().numpy()) < 1e-6
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False, graph=False) as eager:
        Je, ge = eager.misfit_and_gradient(m0, wd, dobs)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False) as gp:
        for rep in range(3):
            Jg, gg = gp.misfit_and_gradient(m0, wd, dobs)
            torch.cuda.synchronize()
            _ = (gg * 2).sum().item()
            assert torch.equal(Je, Jg) and torch.equal(ge, gg), rep


@pytest.mark.parametrize("name", ["fmc_3d.npz", "fmc_3d_fine.npz"])
def test_reversing_work_groups_sharing_edge_rows(name, monkeypatch):
    """Tile kind 4: the reversing adjoint in work-groups of four strips stacked
    along y that pass their edge rows through LDS (forced through the autotuner's
    USCT_TUNE_KIND).  Against the golden vectors, the default reversing plan
    (other tiles: the gradient's shot sums run in another order, so <= 1e-6) and,
    bit for bit, a second call of the same plan."""
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, tx, rec, m0, mt, wd = _golden_plan_inputs(name)
    dobs = torch.tensor(f["dobs"], device=m0.device)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False) as dp:
        Jd, gd = dp.misfit_and_gradient(m0, wd, dobs)
    monkeypatch.setenv("USCT_TUNE_KIND", "4")
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=True) as yp:
        info = yp.info
        print(f"[case] reversing, edge rows through LDS, {name} {info}", flush=True)
        assert info["tile_kind"] == 4 and info["reverse"] == 1
        J, g = yp.misfit_and_gradient(m0, wd, dobs)
        Jr, gr = yp.misfit_and_gradient(m0, wd, dobs)
    monkeypatch.delenv("USCT_TUNE_KIND")
    g_np = g.cpu().numpy()
    e = (abs(J.item() - float(f["J"])) / float(f["J"]), rel_l2(g_np, f["g"]),
         rel_l2(g_np, gd.cpu().numpy()))
    print(f"J {e[0]:.2e} gradient vs golden {e[1]:.2e} vs default tiles {e[2]:.2e}")
    assert torch.equal(J, Jd) and torch.equal(J, Jr) and torch.equal(g, gr)
    assert e[0] < TOL and e[1] < TOL and e[2] < 1e-6


def test_forward_work_groups_sharing_edge_rows(monkeypatch):
    """Tile kind 4 for the 3D forward (64-lane groups): 64 transmitters on the
    golden 3D problem's grid, traces bit for bit against the default tiles, the
    reversing gradient (both kernels of kind 4) within 1e-6 of the default plan's."""
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, _, rec, m0, mt, wd = _golden_plan_inputs("fmc_3d_fine.npz")
    tx = np.array([rec[(7 * i) % len(rec)] for i in range(64)], np.int64)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=False) as dp:
        d0 = dp.forward_fmc(mt, wd)
        J0, g0 = dp.misfit_and_gradient(m0, wd, d0 * 1.01)
    monkeypatch.setenv("USCT_TUNE_KIND", "4")
    with Plan(shape, h, dt, nt, tx, rec, reverse=True, autotune=True) as yp:
        info = yp.info
        print(f"[case] forward and reversing adjoint, edge rows through LDS {info}", flush=True)
        assert info["fwd_tile_kind"] == 4 and info["tile_kind"] == 4
        d4 = yp.forward_fmc(mt, wd)
        J4, g4 = yp.misfit_and_gradient(m0, wd, d0 * 1.01)
    monkeypatch.delenv("USCT_TUNE_KIND")
    e = rel_l2(g4.cpu().numpy(), g0.cpu().numpy())
    print(f"gradient vs default tiles {e:.2e}")
    assert torch.equal(d4, d0) and torch.equal(J4, J0)
    assert e < 1e-6


def test_reversing_c4_reduced_vs_oracle():
    """C4 spacing, time step and array on a 64^3 grid, 300 steps, 4 shots."""
    from oracle import oracle as O
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c4(n=64, nt_scale=0.15, shots=4)
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    dobs = O.forward_fmc(prob.m_true, prob.geom, prob.wavelet, prob.dt, prob.h, prob.nt, None,
                         prob.src_list)
    J_ref, g_ref = O.misfit_and_gradient(prob.m0, prob.geom, prob.wavelet, prob.dt, prob.h,
                                         prob.nt, dobs, None, prob.src_list)
    dev = torch.device("cuda")
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=True) as plan:
        print(f"[case] reversing C4 reduced {plan.info}", flush=True)
        J, g = plan.misfit_and_gradient(torch.tensor(prob.m0, device=dev),
                                        torch.tensor(prob.wavelet, device=dev),
                                        torch.tensor(dobs, device=dev))
    e = (abs(J.item() - J_ref) / J_ref, rel_l2(g.cpu().numpy(), g_ref))
    print(f"J {e[0]:.2e} gradient {e[1]:.2e}")
    assert max(e) < TOL


@pytest.mark.parametrize("misfit", ["l2", "gcn"])
def test_reversing_windowed_and_gcn(misfit):
    """Time windows (fwi.py:298-301) and the GCN misfit only change the adjoint
    source: with the forward run backwards they give the gradients of the plan
    that keeps levels."""
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, tx, rec, m0, mt, wd = _golden_plan_inputs("fmc_3d_fine.npz")
    dobs = torch.tensor(f["dobs"], device=m0.device)
    t = torch.arange(nt, dtype=torch.float64, device=m0.device)[:, None]
    win = torch.clamp((0.7 * nt - t) / (0.1 * nt), 0.0, 1.0).expand(nt, len(rec))
    tw = torch.stack([win, win.flip(0)]).contiguous()
    out = {}
    for rev in (True, False):
        with Plan(shape, h, dt, nt, tx, rec, reverse=rev) as p:
            print(f"[case] reversing={rev} {misfit} windowed", flush=True)
            out[rev] = (p.misfit_and_gradient(m0, wd, dobs, misfit_type=misfit),
                        p.misfit_and_gradient(m0, wd, dobs, misfit_type=misfit, time_weights=tw))
    for k in range(2):
        (J1, g1), (J2, g2) = out[True][k], out[False][k]
        e = rel_l2(g1.cpu().numpy(), g2.cpu().numpy())
        print(f"{'windowed' if k else 'plain'}: J {abs(J1.item() - J2.item()) / abs(J2.item()):.2e} "
              f"gradient {e:.2e}")
        assert torch.equal(J1, J2) and e < 1e-6


def test_reversing_long_time_axis():
    """C4's spacing, time step and full axis (2000 levels) on a 64^3 grid with
    its ellipsoid phantom, 8 shots: the levels re-created over 2000 steps still
    give the gradient of the plan that keeps checkpoints, to <= 1e-6."""
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c4(n=64, nt_scale=1.0, shots=8)
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    dev = torch.device("cuda")
    m0, mt = torch.tensor(prob.m0, device=dev), torch.tensor(prob.m_true, device=dev)
    wd = torch.tensor(prob.wavelet, device=dev)
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=True) as rp:
        print(f"[case] reversing over {prob.nt} levels {rp.info}", flush=True)
        dobs = rp.forward_fmc(mt, wd)
        J, g = rp.misfit_and_gradient(m0, wd, dobs)
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=False, ckpt_every=50) as kp:
        Jk, gk = kp.misfit_and_gradient(m0, wd, dobs)
    e = (abs(J.item() - Jk.item()) / Jk.item(), rel_l2(g.cpu().numpy(), gk.cpu().numpy()))
    print(f"J {e[0]:.2e} gradient {e[1]:.2e}")
    assert e[0] < 1e-9 and e[1] < 1e-6


def test_reversing_gauss_newton_and_adjoint_source():
    """The Hessian-vector product (background levels re-created for the
    adjoint's imaging) against the oracle, and the caller-supplied adjoint
    source against the plan that keeps levels."""
    from oracle import oracle as O
    from openusct_amd.plan import Plan
    f, shape, h, dt, nt, tx, rec, m0, mt, wd = _golden_plan_inputs("fmc_3d_fine.npz")
    geom = IdxGeom(f["idx"])
    src = [int(s) for s in f["src_list"]]
    rng = np.random.default_rng(5)
    v = np.zeros(shape)
    v[5:-5, 5:-5, 5:-5] = rng.standard_normal(shape)[5:-5, 5:-5, 5:-5]
    v *= float(np.mean(f["m0"]))
    ref = O.gn_hessian_vector(f["m0"], geom, f["wavelet"], dt, h, nt, v, None, src)
    with Plan(shape, h, dt, nt, tx, rec, reverse=True) as rp:
        print(f"[case] reversing GN and adjoint source {rp.info}", flush=True)
        hv = rp.gn_hessian_vector(m0, wd, torch.tensor(v, device=m0.device))
        adj = torch.randn(len(tx), nt, len(rec), dtype=torch.float64, device=m0.device,
                          generator=torch.Generator(device=m0.device).manual_seed(3))
        gs = rp.gradient_from_source(m0, wd, adj)
    with Plan(shape, h, dt, nt, tx, rec, reverse=False) as kp:
        gk = kp.gradient_from_source(m0, wd, adj)
    e = (rel_l2(hv.cpu().numpy(), ref), rel_l2(gs.cpu().numpy(), gk.cpu().numpy()))
    print(f"Hv vs oracle {e[0]:.2e}; adjoint-source gradient vs stored levels {e[1]:.2e}")
    assert e[0] < TOL and e[1] < 1e-6


def test_reverse_needs_a_3d_point_element_gradient_plan():
    from openusct_amd.plan import Plan, linear_indices
    f = golden("fmc_2d.npz")
    shape = f["m0"].shape
    rec = linear_indices(IdxGeom(f["idx"]).idx, shape)
    with pytest.raises(NotImplementedError):
        Plan(shape, float(f["h"]), float(f["dt"]), 50, rec[:2], rec, reverse=True)
    f3, shape3, h, dt, nt, tx, rec3, *_ = _golden_plan_inputs("fmc_3d.npz")
    with pytest.raises(NotImplementedError):
        Plan(shape3, h, dt, nt, tx, rec3, reverse=True, forward_only=True)
    groups = [([int(c)], [1.0]) for c in rec3]
    with pytest.raises(NotImplementedError):
        Plan(shape3, h, dt, nt, None, None, reverse=True, tx_groups=groups[:2], rec_groups=groups)
