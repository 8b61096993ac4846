"""GPU: the configurations the metric is quoted on, at their full sizes.

The reference checks its GPU backend on its own problem sizes
(software/python/test_gpu.py:43-91); these do the same for the batched path on
the workloads of the benchmark:

* the 256^3 cylinder, 64 shots, nt = 2000: the gradient that keeps no forward
  levels (the adjoint sweep runs the forward backwards) against the plan that
  keeps checkpoints every 50 levels, <= 1e-6;
* the 1024^2 ring, 256 shots, nt = 4000 (checkpointed): the one-pass plan
  against a plan forced into several passes by its memory budget, and the full
  grid with its time axis cut to 2000 levels against the FP64 oracle, <= 1e-5;
* the 512^2 ring at 128 and at 256 shots with the benchmark's own plan
  (autotuned, default budget): traces bit for bit and the gradient against the
  plain per-shot path, which sums one shot's gradient after another;
* a serving mix drawn from the serving workload's own generator covering every
  grid family (256^2, 512^2, 1024^2, 128^3, 192^3, 256^3) and both request
  kinds: each result the same alone and in the mix, bit for bit.

Each test prints the errors it measures.
"""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def _tensors(prob):
    return (torch.tensor(prob.m_true, device=DEV), torch.tensor(prob.m0, device=DEV),
            torch.tensor(prob.wavelet, device=DEV))


def test_c4_full_reversing_vs_checkpointed():
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c4()
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    mt, m0, wd = _tensors(prob)
    t0 = time.time()
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=True) as rp:
        print(f"[case] C4 256^3, 64 shots, nt 2000, forward run backwards {rp.info}", flush=True)
        dobs = rp.forward_fmc(mt, wd)
        J, g = rp.misfit_and_gradient(m0, wd, dobs)
        g = g.cpu().numpy()
    torch.cuda.empty_cache()
    print(f"  reversing gradient done ({time.time() - t0:.0f} s)", flush=True)
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=False, ckpt_every=50) as kp:
        print(f"[case] C4 checkpointed, K = 50 {kp.info}", flush=True)
        Jk, gk = kp.misfit_and_gradient(m0, wd, dobs)
        gk = gk.cpu().numpy()
    e = (abs(J.item() - Jk.item()) / Jk.item(), rel_l2(g, gk))
    print(f"C4: J {e[0]:.2e}  gradient {e[1]:.2e} ({time.time() - t0:.0f} s)")
    assert np.isfinite(g).all() and np.abs(g).max() > 0.0
    assert e[0] < 1e-9 and e[1] < 1e-6


def test_c4_full_edge_rows_through_lds(monkeypatch):
    """C4 at full size with both 3D kernels of tile kind 4 (work-groups of four
    y-stacked strips passing edge rows through LDS) against the default tiles:
    traces and misfit bit for bit, gradient within 1e-6."""
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c4()
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    mt, m0, wd = _tensors(prob)
    t0 = time.time()
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=True, autotune=False) as dp:
        print(f"[case] C4 default tiles {dp.info}", flush=True)
        dobs = dp.forward_fmc(mt, wd)
        J0, g0 = dp.misfit_and_gradient(m0, wd, dobs * 1.01)
        g0 = g0.cpu().numpy()
    torch.cuda.empty_cache()
    monkeypatch.setenv("USCT_TUNE_KIND", "4")
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, reverse=True, autotune=True) as yp:
        info = yp.info
        print(f"[case] C4 tile kind 4 {info}", flush=True)
        assert info["tile_kind"] == 4 and info["fwd_tile_kind"] == 4
        d4 = yp.forward_fmc(mt, wd)
        J4, g4 = yp.misfit_and_gradient(m0, wd, dobs * 1.01)
        g4 = g4.cpu().numpy()
    monkeypatch.delenv("USCT_TUNE_KIND")
    e = rel_l2(g4, g0)
    print(f"C4 kind 4: traces equal {torch.equal(d4, dobs)}, gradient {e:.2e}, "
          f"bitwise {np.array_equal(g4, g0)} ({time.time() - t0:.0f} s)")
    assert torch.equal(d4, dobs) and torch.equal(J4, J0)
    assert e < 1e-6


def test_c3_full_one_pass_vs_multi_pass():
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c3()
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    mt, m0, wd = _tensors(prob)
    t0 = time.time()
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec) as p1:
        info1 = p1.info
        print(f"[case] C3 1024^2, 256 shots, nt 4000 {info1}", flush=True)
        dobs = p1.forward_fmc(mt, wd)
        J1, g1 = p1.misfit_and_gradient(m0, wd, dobs)
        g1 = g1.cpu().numpy()
    torch.cuda.empty_cache()
    # a budget of about a third of the one-pass plan forces several passes
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec,
              mem_budget=info1["workspace_bytes"] // 3) as p2:
        info2 = p2.info
        print(f"[case] C3 under a third of the budget {info2}", flush=True)
        d2 = p2.forward_fmc(mt, wd)
        J2, g2 = p2.misfit_and_gradient(m0, wd, dobs)
        g2 = g2.cpu().numpy()
    assert info2["shot_passes"] > info1["shot_passes"]
    assert torch.equal(d2, dobs)
    e = (abs(J1.item() - J2.item()) / J1.item(), rel_l2(g2, g1))
    print(f"C3: passes {info1['shot_passes']} -> {info2['shot_passes']}, K {info1['ckpt_every']} -> "
          f"{info2['ckpt_every']}: traces equal, J {e[0]:.2e}, gradient {e[1]:.2e} "
          f"({time.time() - t0:.0f} s)")
    assert np.isfinite(g1).all() and np.abs(g1).max() > 0.0
    assert e[0] < 1e-9 and e[1] < 1e-6


def test_c3_grid_short_axis_vs_oracle():
    """The 1024^2 grid, array and phantom with the time axis cut to 2000 levels
    (75 mm of travel: the transmitted and scattered waves cross the inclusions
    and reach the ring) and 3 shots, against the FP64 oracle."""
    from oracle import oracle as O
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.c3(nt_scale=0.5)
    src = [0, 85, 170]
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[src]
    t0 = time.time()
    dobs = O.forward_fmc(prob.m_true, prob.geom, prob.wavelet, prob.dt, prob.h, prob.nt, None, src)
    J_ref, g_ref = O.misfit_and_gradient(prob.m0, prob.geom, prob.wavelet, prob.dt, prob.h, prob.nt,
                                         dobs, None, src)
    print(f"[case] C3 grid, nt {prob.nt}, shots {src}: oracle in {time.time() - t0:.0f} s",
          flush=True)
    mt, m0, wd = _tensors(prob)
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, ckpt_every=40) as p:
        d = p.forward_fmc(mt, wd).cpu().numpy()
        J, g = p.misfit_and_gradient(m0, wd, torch.tensor(dobs, device=DEV))
    e = (rel_l2(d, dobs), abs(J.item() - J_ref) / J_ref, rel_l2(g.cpu().numpy(), g_ref))
    print(f"C3 grid nt {prob.nt}: traces {e[0]:.2e}  J {e[1]:.2e}  gradient {e[2]:.2e} "
          f"(J = {J_ref:.3e})")
    assert J_ref > 0.0
    assert max(e) < 1e-5


@pytest.mark.parametrize("name", ["C2", "C2-256"])
def test_512_batched_bench_plan_vs_plain_per_shot(name):
    """The benchmark's plan (autotuned, default budget: whole history, two-level
    launches, 64-lane groups) against the plain path, which runs one shot after
    another, one launch per level, and sums their gradients."""
    from openusct_amd import configs
    from openusct_amd.plan import Plan, linear_indices
    prob = configs.BUILDERS[name]()
    rec = linear_indices(prob.geom.idx, prob.shape)
    tx = rec[prob.src_list]
    mt, m0, wd = _tensors(prob)
    t0 = time.time()
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, method="batched", autotune=True) as bp:
        print(f"[case] {name} 512^2, {len(tx)} shots, bench plan {bp.info}", flush=True)
        dobs = bp.forward_fmc(mt, wd)
        J, g = bp.misfit_and_gradient(m0, wd, dobs)
        g = g.cpu().numpy()
    torch.cuda.empty_cache()
    with Plan(prob.shape, prob.h, prob.dt, prob.nt, tx, rec, method="plain") as pp:
        dp = pp.forward_fmc(mt, wd)
        Jp, gp = pp.misfit_and_gradient(m0, wd, dobs)
        gp = gp.cpu().numpy()
    assert torch.equal(dp, dobs)
    e = (abs(J.item() - Jp.item()) / Jp.item(), rel_l2(g, gp))
    print(f"{name}: traces equal, J {e[0]:.2e}, gradient {e[1]:.2e} ({time.time() - t0:.0f} s)")
    assert np.isfinite(g).all() and np.abs(g).max() > 0.0
    assert e[0] < 1e-9 and e[1] < 1e-6


def test_serving_mix_from_the_serving_generator_every_grid_family():
    """Requests of make_mix (the serving workload's generator, seed 5): the first
    request of every grid family and kind among 200 drawn, with time axes x0.1 and
    16-48 shots so that the 256^3 requests stay short; all of them in one mix, then
    each alone."""
    from openusct_amd.serve import Server, fill_observed, make_mix, summarize
    pool = make_mix(200, seed=5, nt_scale=0.1, shots=(16, 48))
    picked, seen = [], set()
    for r in pool:
        key = (r.shape, r.kind)
        if key not in seen:
            seen.add(key)
            picked.append(r)
    shapes = {r.shape for r in picked}
    assert {(256, 256), (512, 512), (1024, 1024), (128,) * 3, (192,) * 3, (256,) * 3} <= shapes
    for i, r in enumerate(picked):
        r.arrival = 0.02 * i
    fill_observed(picked)
    t0 = time.time()
    mixed = Server(max_inflight=6).run(picked)
    print(f"[case] serving mix of {len(picked)} requests "
          f"({sorted((r.shape, r.kind, len(r.tx)) for r in picked)}): {summarize(mixed)}",
          flush=True)
    assert sorted(mixed) == sorted(r.rid for r in picked)
    for r in picked:
        r0, r.arrival = r.arrival, 0.0
        alone = Server(max_inflight=1).run([r])[r.rid]
        r.arrival = r0
        b = mixed[r.rid]
        if r.kind == "gradient":
            assert alone.J == b.J, r.rid
            assert torch.equal(alone.grad, b.grad), r.rid
            assert torch.isfinite(b.grad).all()
        else:
            assert torch.equal(alone.data, b.data), r.rid
    print(f"every result equal alone and in the mix ({time.time() - t0:.0f} s)")
