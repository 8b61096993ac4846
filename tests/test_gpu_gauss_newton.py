"""GPU: the Gauss-Newton Hessian-vector product (J^T J) v and the truncated
Gauss-Newton optimizer, against the FP64 oracle on the problem of
simulation/tests/test_gauss_newton.py:21-28 (8-element ring, 61^2 at 1 mm,
nt = 300).  Tolerance: relative L2 <= 1e-5, as for the gradient; between the
device's own plans (checkpointed / whole history, one or two levels per
launch) the product is equal bit for bit.
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


def _setup():
    from openusct_amd import phantom
    from openusct_amd.geometry import RingArray
    from openusct_amd.sources import ricker
    h = 1.0e-3
    geom = RingArray(n_elements=8, radius_m=0.024, domain_m=0.06, h=h)
    n = geom.n
    dt, nt = 1.2e-7, 300
    w = ricker(nt, dt, 0.4e6)
    m0 = phantom.velocity_to_m(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h))
    return geom, n, h, dt, nt, w, m0


def _v(shape, m0, seed=2):
    """test_gauss_newton.py:36-39: a random direction, zero on a 5-cell margin."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape)
    sl = tuple(slice(5, -5) for _ in shape)
    out = np.zeros(shape)
    out[sl] = v[sl]
    return out * float(np.mean(m0))


@pytest.mark.parametrize("sponge", [False, True])
@pytest.mark.parametrize("method", ["batched", "plain"])
def test_hessian_vector_vs_oracle(method, sponge):
    from oracle import oracle as O
    from openusct_amd import fwi, solver
    geom, n, h, dt, nt, w, m0 = _setup()
    sp = solver.make_sponge((n, n), width=8) if sponge else None
    sl = [0, 4]
    v = _v((n, n), m0)
    print(f"[case] GN Hv 2D {n}x{n} nt={nt} method={method} sponge={sponge}", flush=True)
    fwi.METHOD = method
    hv = fwi.gn_hessian_vector(m0, geom, w, dt, h, nt, v, sp, sl)
    ref = O.gn_hessian_vector(m0, geom, w, dt, h, nt, v, sp, sl)
    e = rel_l2(hv, ref)
    print(f"Hv rel L2 {e:.2e}")
    assert e < TOL


@pytest.mark.parametrize("name", ["fmc_3d.npz", "fmc_3d_fine.npz"])
def test_hessian_vector_3d_vs_oracle(name):
    from oracle import oracle as O
    from openusct_amd import fwi
    f = golden(name)
    geom = IdxGeom(f["idx"])
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    src = [int(s) for s in f["src_list"]]
    m0 = f["m0"]
    v = _v(m0.shape, m0, seed=5)
    print(f"[case] GN Hv 3D {name} {m0.shape} nt={nt}", flush=True)
    hv = fwi.gn_hessian_vector(m0, geom, w, dt, h, nt, v, None, src)
    ref = O.gn_hessian_vector(m0, geom, w, dt, h, nt, v, None, src)
    e = rel_l2(hv, ref)
    print(f"Hv rel L2 {e:.2e}")
    assert e < TOL


def test_hessian_vector_plans_bit_identical():
    """Checkpoint windows and two-level launches do not change the product; a
    multi-pass plan (shots split over passes) only changes the summation order."""
    from openusct_amd.plan import Plan, linear_indices
    geom, n, h, dt, nt, w, m0 = _setup()
    rec = linear_indices(geom.idx, (n, n))
    tx = rec[[0, 2, 4, 6]]
    dev = torch.device("cuda")
    md = torch.tensor(m0, device=dev)
    wd = torch.tensor(w, device=dev)
    vd = torch.tensor(_v((n, n), m0), device=dev)
    with Plan((n, n), h, dt, nt, tx, rec, autotune=False, fuse=False) as p:
        ref = p.gn_hessian_vector(md, wd, vd)
        info = p.info
    for kw in (dict(ckpt_every=7), dict(ckpt_every=50), dict(fuse=True), dict(graph=False)):
        with Plan((n, n), h, dt, nt, tx, rec, autotune=False, **kw) as p:
            print(f"[case] GN Hv plan {kw} {p.info}", flush=True)
            assert p.info["shots_per_pass"] == info["shots_per_pass"]
            for rep in range(2):  # graph replay
                hv = p.gn_hessian_vector(md, wd, vd)
                torch.cuda.synchronize()
                assert torch.equal(hv, ref), (kw, rep)
    # 40 shots over several passes
    tx = np.array([(5 + i) * n + 7 + i for i in range(40)], dtype=np.int64)
    nt2 = 200
    with Plan((n, n), h, dt, nt2, tx, rec) as one:
        ref = one.gn_hessian_vector(md, wd[:nt2], vd)
    budget = 60 * n * n * 4 * 16 + (8 << 20)
    with Plan((n, n), h, dt, nt2, tx, rec, mem_budget=budget) as many:
        print(f"[case] GN Hv multi-pass {many.info}", flush=True)
        assert many.info["shot_passes"] > 1
        hv = many.gn_hessian_vector(md, wd[:nt2], vd)
    assert rel_l2(hv.cpu().numpy(), ref.cpu().numpy()) < 1e-6


def test_hessian_vector_3d_fused_bit_identical():
    """3D: background and Born levels in one launch (k_born3d) against separate
    k_fwd3d and k_born launches (fuse=False), bit for bit -- keeping the whole
    history, with checkpoints (whose re-computed background levels come from
    k_fwd3d) and with the forward run backwards."""
    from openusct_amd.plan import Plan, linear_indices
    f = golden("fmc_3d_fine.npz")
    shape = f["m0"].shape
    rec = linear_indices(IdxGeom(f["idx"]).idx, shape)
    tx = rec[[int(s) for s in f["src_list"]]]
    h, dt, nt = float(f["h"]), float(f["dt"]), int(f["nt"])
    dev = torch.device("cuda")
    md, wd = torch.tensor(f["m0"], device=dev), torch.tensor(f["wavelet"], device=dev)
    vd = torch.tensor(_v(shape, f["m0"], seed=5), device=dev)
    for rev in (False, True):
        with Plan(shape, h, dt, nt, tx, rec, autotune=False, fuse=False, reverse=rev) as p:
            ref = p.gn_hessian_vector(md, wd, vd)
        cases = [dict(reverse=rev)] + ([dict(reverse=False, ckpt_every=20)] if not rev else [])
        for kw in cases:
            with Plan(shape, h, dt, nt, tx, rec, autotune=False, **kw) as p:
                print(f"[case] GN Hv 3D fused {kw} {p.info}", flush=True)
                for rep in range(2):
                    hv = p.gn_hessian_vector(md, wd, vd)
                    torch.cuda.synchronize()
                    assert torch.equal(hv, ref), (kw, rep)


def test_hessian_is_symmetric_positive():
    """<u, Hv> = <Hu, v> and <v, Hv> > 0 on the device."""
    from openusct_amd import fwi
    geom, n, h, dt, nt, w, m0 = _setup()
    u = _v((n, n), m0, seed=11)
    v = _v((n, n), m0, seed=12)
    hu = fwi.gn_hessian_vector(m0, geom, w, dt, h, nt, u, None, [1, 5])
    hv = fwi.gn_hessian_vector(m0, geom, w, dt, h, nt, v, None, [1, 5])
    a, b = float(np.sum(u * hv)), float(np.sum(hu * v))
    print(f"[case] GN symmetry <u,Hv> {a:.6e} <Hu,v> {b:.6e}")
    assert abs(a - b) <= 1e-5 * max(abs(a), abs(b))
    assert float(np.sum(v * hv)) > 0.0 and float(np.sum(u * hu)) > 0.0


def test_hessian_vector_rejects_forward_only_plan():
    from openusct_amd.plan import Plan, linear_indices
    geom, n, h, dt, nt, w, m0 = _setup()
    rec = linear_indices(geom.idx, (n, n))
    dev = torch.device("cuda")
    with Plan((n, n), h, dt, nt, rec[:2], rec, forward_only=True) as p:
        with pytest.raises(NotImplementedError):
            p.gn_hessian_vector(torch.tensor(m0, device=dev), torch.tensor(w, device=dev),
                                torch.zeros((n, n), dtype=torch.float64, device=dev))


def test_gauss_newton_beats_gradient_descent():
    """test_gauss_newton.py:54-79 on the device: four Gauss-Newton iterations
    (8 CG steps) reduce the misfit below 5% and below four gradient-descent
    iterations."""
    from openusct_amd import fwi, phantom
    geom, n, h, dt, nt, w, _ = _setup()
    c_true = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h)
    c_true = phantom.add_inclusion(c_true, (0.55, 0.5), 0.006, 3350.0, h)
    m_true = phantom.velocity_to_m(c_true)
    m0 = phantom.velocity_to_m(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h))
    sl = [0, 2, 4, 6]
    dobs = fwi.forward_fmc(m_true, geom, w, dt, h, nt, src_list=sl)
    yy, xx = np.mgrid[0:n, 0:n].astype(float) * h
    cc = (n - 1) * h / 2
    mask = (np.hypot(xx - cc, yy - cc) <= 0.019).astype(float)
    _, hist_gd = fwi.invert(m0, geom, w, dt, h, nt, dobs, src_list=sl, n_iter=4,
                            update_mask=mask)
    _, hist_gn = fwi.invert(m0, geom, w, dt, h, nt, dobs, src_list=sl, n_iter=4,
                            update_mask=mask, optimizer="gauss-newton", n_cg=8)
    red_gd = hist_gd[-1] / hist_gd[0]
    red_gn = hist_gn[-1] / hist_gn[0]
    print(f"[case] 4 iterations: GD misfit -> {red_gd * 100:.2f}% | GN -> {red_gn * 100:.3f}%")
    assert red_gn < red_gd
    assert red_gn < 0.05


def test_lazy_buffers_are_held_to_the_plan_budget():
    """A plan whose budget fits the gradient but not the Born state (or the time
    windows' traces) refuses the Gauss-Newton product (and the windowed call) with
    MemoryError before allocating; gradients on it keep working, and a plan with
    room runs the same product.  The 256^2 problem, 8 shots, whole history kept:
    the Born state needs 24 B per shot-cell (25 MB), the windows 8 B per trace
    sample (8 MB), and the tight budget leaves 4 MB beyond the plan."""
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c1()
    shape, h, dt, nt = prob.shape, prob.h, prob.dt, prob.nt
    rec = linear_indices(prob.geom.idx, shape)
    tx = rec[prob.src_list]
    dev = torch.device("cuda")
    md, wd = torch.tensor(prob.m0, device=dev), torch.tensor(prob.wavelet, device=dev)
    mt = torch.tensor(prob.m_true, device=dev)
    kw = dict(autotune=False, ckpt_every=nt - 1)
    with Plan(shape, h, dt, nt, tx, rec, **kw) as free:
        held = free.info["workspace_bytes"]
        dobs = free.forward_fmc(mt, wd)
        v = torch.ones_like(md) * md.mean()
        hv = free.gn_hessian_vector(md, wd, v)
    budget = held + (4 << 20)
    print(f"[case] 256^2, 8 shots: plan of {held} bytes, budget {budget}", flush=True)
    with Plan(shape, h, dt, nt, tx, rec, mem_budget=budget, **kw) as tight:
        assert tight.info["workspace_bytes"] == held
        with pytest.raises(MemoryError, match="Gauss-Newton"):
            tight.gn_hessian_vector(md, wd, v)
        with pytest.raises(MemoryError, match="time-window"):
            tight.misfit(md, wd, dobs, time_weights=torch.ones_like(dobs))
        assert tight.info["workspace_bytes"] == held
        J, g = tight.misfit_and_gradient(md, wd, dobs)
        assert J.item() > 0.0 and torch.isfinite(g).all()
    assert torch.isfinite(hv).all() and hv.abs().max().item() > 0.0
