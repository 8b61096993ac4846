"""GPU: three time levels per launch (k_fwd2d_xn / k_adj2d_xn) against two.

Each level a three-level launch computes is the same arithmetic as a
single-level launch, on halo columns and rows wide enough that every value it
hands on equals the owner's; the gradient accumulators take their updates in
the single-level order.  So a plan that launches three levels at a time must
give the traces, the misfit and the gradient of a two-level plan with the same
tiles bit for bit: over time axes whose length leaves every remainder (one or
two levels to finish with), checkpoint windows of every length modulo three,
several shot passes, the GCN misfit, time weights, the Gauss-Newton product,
a caller-supplied adjoint source, forward-only plans and FP32 low parts.  The
reference computes one level per step (software/python/uap_gpu.py:128-140,
211-223).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import IdxGeom, golden, rel_l2

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def _problem(n_src=24, nt=200):
    """The 61x61 golden ring (16 receivers); sources on the ring elements, then on
    cells spread over the grid, so that a pass holds up to 64-lane shot groups."""
    from openusct_amd.plan import linear_indices
    f = golden("fmc_2d.npz")
    geom = IdxGeom(f["idx"])
    n = f["m0"].shape[0]
    rec = linear_indices(geom.idx, f["m0"].shape)
    extra = [(5 + (7 * i) % 51) * n + 4 + (11 * i) % 53 for i in range(max(n_src - len(rec), 0))]
    tx = np.concatenate([rec, np.array(extra, dtype=np.int64)])[:n_src]
    w = np.asarray(f["wavelet"], dtype=np.float64)
    if len(w) < nt:
        w = np.concatenate([w, np.zeros(nt - len(w))])
    return (f, float(f["h"]), float(f["dt"]), tx, rec, torch.tensor(f["m0"], device=DEV),
            torch.tensor(f["m_true"], device=DEV), torch.tensor(w[:nt], device=DEV))


def _pair(shape, h, dt, nt, tx, rec, **kw):
    from openusct_amd.plan import Plan
    return (Plan(shape, h, dt, nt, tx, rec, autotune=False, levels=2, **kw),
            Plan(shape, h, dt, nt, tx, rec, autotune=False, levels=3, **kw))


@pytest.mark.parametrize("nt,K", [(200, 0), (201, 0), (202, 0), (4, 0), (3, 0), (2, 0),
                                  (200, 7), (201, 9), (202, 8), (100, 3), (100, 4), (100, 2)])
def test_three_level_gradient_bit_identical(nt, K):
    f, h, dt, tx, rec, m0, mt, wd = _problem(n_src=70, nt=nt)
    two, three = _pair(m0.shape, h, dt, nt, tx, rec, ckpt_every=K)
    with two, three:
        i2, i3 = two.info, three.info
        print(f"[case] levels 3 vs 2, nt={nt} K={K} {i3}", flush=True)
        assert i2["fwd_levels_per_launch"] == 2 and i3["fwd_levels_per_launch"] == 3
        assert i2["levels_per_launch"] == 2 and i3["levels_per_launch"] == 3
        assert i3["ckpt_every"] == i2["ckpt_every"] and i3["shots_per_pass"] == i2["shots_per_pass"]
        d2 = two.forward_fmc(mt, wd)
        d3 = three.forward_fmc(mt, wd)
        assert torch.equal(d2, d3)
        J2, g2 = two.misfit_and_gradient(m0, wd, d2)
        J3, g3 = three.misfit_and_gradient(m0, wd, d2)
    assert torch.equal(J2, J3)
    assert torch.equal(g2, g3)
    if nt >= 100:  # the scattered waves have reached the ring
        assert g3.abs().max().item() > 0.0


def test_three_level_misfits_windows_and_passes():
    """GCN misfit, time weights, graph replays and a plan split over shot passes."""
    f, h, dt, tx, rec, m0, mt, wd = _problem(n_src=40, nt=151)
    nt = 151
    n = m0.shape[0]
    tw = torch.rand((len(tx), nt, len(rec)), generator=torch.Generator().manual_seed(3),
                    dtype=torch.float64).to(DEV)
    field = n * n * 4
    for kw in (dict(), dict(ckpt_every=10), dict(mem_budget=60 * field * 16 + (8 << 20))):
        two, three = _pair(m0.shape, h, dt, nt, tx, rec, **kw)
        with two, three:
            print(f"[case] levels 3 vs 2 {kw} {three.info}", flush=True)
            dobs = two.forward_fmc(mt, wd)
            for misfit in ("l2", "gcn"):
                r2 = two.misfit_and_gradient(m0, wd, dobs, misfit_type=misfit)
                for rep in range(2):  # graph replay
                    r3 = three.misfit_and_gradient(m0, wd, dobs, misfit_type=misfit)
                    torch.cuda.synchronize()
                    assert torch.equal(r2[0], r3[0]) and torch.equal(r2[1], r3[1]), (kw, misfit, rep)
            r2 = two.misfit_and_gradient(m0, wd, dobs, time_weights=tw)
            r3 = three.misfit_and_gradient(m0, wd, dobs, time_weights=tw)
            assert torch.equal(r2[0], r3[0]) and torch.equal(r2[1], r3[1]), kw


def test_three_level_gauss_newton_adjoint_source_forward_only_fp32_lo():
    f, h, dt, tx, rec, m0, mt, wd = _problem(n_src=8, nt=120)
    nt = 120
    v = torch.tensor(np.random.default_rng(4).standard_normal(m0.shape) * 1e-9, device=DEV)
    src = torch.tensor(np.random.default_rng(5).standard_normal((len(tx), nt, len(rec))), device=DEV)
    two, three = _pair(m0.shape, h, dt, nt, tx, rec, ckpt_every=11)
    with two, three:
        print(f"[case] Gauss-Newton and adjoint source, levels 3 vs 2 {three.info}", flush=True)
        assert torch.equal(two.gn_hessian_vector(m0, wd, v), three.gn_hessian_vector(m0, wd, v))
        assert torch.equal(two.gradient_from_source(m0, wd, src),
                           three.gradient_from_source(m0, wd, src))
    two, three = _pair(m0.shape, h, dt, nt, tx, rec, forward_only=True)
    with two, three:
        print(f"[case] forward-only, levels 3 vs 2 {three.info}", flush=True)
        assert torch.equal(two.forward_fmc(mt, wd), three.forward_fmc(mt, wd))
    two, three = _pair(m0.shape, h, dt, nt, tx, rec, fp32_lo=True)
    with two, three:
        print(f"[case] FP32 low parts, levels 3 vs 2 {three.info}", flush=True)
        d = two.forward_fmc(mt, wd)
        assert torch.equal(d, three.forward_fmc(mt, wd))
        r2 = two.misfit_and_gradient(m0, wd, d * 1.01)
        r3 = three.misfit_and_gradient(m0, wd, d * 1.01)
        assert torch.equal(r2[0], r3[0]) and torch.equal(r2[1], r3[1])


def test_three_level_autotuned_and_rejections():
    """Autotuned plans (two-level by default, three-level on request): traces do
    not change, and the misfit and gradient change only by the order in which
    lane groups of other widths sum over shots.  Plans that cannot fuse launches
    refuse three levels."""
    from openusct_amd.plan import Plan
    f, h, dt, tx, rec, m0, mt, wd = _problem(n_src=64, nt=90)
    nt = 90
    with Plan(m0.shape, h, dt, nt, tx, rec, autotune=False, levels=2) as ref:
        dobs = ref.forward_fmc(mt, wd)
        J, g = ref.misfit_and_gradient(m0, wd, dobs)
    for levels in (None, 3):
        with Plan(m0.shape, h, dt, nt, tx, rec, autotune=True, levels=levels) as p:
            print(f"[case] autotuned levels={levels} {p.info}", flush=True)
            if levels == 3:
                assert p.info["levels_per_launch"] == 3 and p.info["fwd_levels_per_launch"] == 3
            assert torch.equal(p.forward_fmc(mt, wd), dobs)
            J1, g1 = p.misfit_and_gradient(m0, wd, dobs)
            assert abs(J1.item() - J.item()) <= 1e-9 * abs(J.item())
            assert rel_l2(g1.cpu().numpy(), g.cpu().numpy()) < 1e-6
    for kw in (dict(sponge=True), dict(fuse=False), dict(method="plain")):
        with pytest.raises(NotImplementedError) as e:
            Plan(m0.shape, h, dt, nt, tx, rec, levels=3, **kw)
        assert "three-level" in str(e.value)
