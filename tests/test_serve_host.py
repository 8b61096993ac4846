"""Serving loop host logic (CPU only): the request mix, footprints, statistics."""

from __future__ import annotations

import numpy as np


def test_make_mix_is_seeded_and_poisson():
    from openusct_amd.serve import make_mix
    a = make_mix(30, seed=5, nt_scale=0.01, rate=100.0)
    b = make_mix(30, seed=5, nt_scale=0.01, rate=100.0)
    assert [(r.kind, r.shape, len(r.tx)) for r in a] == [(r.kind, r.shape, len(r.tx)) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) >= 0) and arr[-1] > 0
    kinds = {r.kind for r in a}
    assert kinds == {"gradient", "forward"}
    for r in a:
        assert 16 <= len(r.tx) <= 256
        assert len(set(map(int, r.tx))) == len(r.tx)
        assert r.nt >= 8


def test_footprint_depends_on_kind_and_cap():
    from openusct_amd.serve import make_mix
    r = make_mix(4, seed=1, nt_scale=0.05, grids=[("2d", 512)])[0]
    r.kind = "gradient"
    g = r.footprint()
    r.kind = "forward"
    f = r.footprint()
    assert f < g
    assert r.footprint(cap=f // 2) == f // 2


def test_summarize():
    from openusct_amd.serve import Result, summarize
    res = {i: Result(i, "forward", None, None, None, arrival=0.1 * i, admitted=0.1 * i,
                     done=0.1 * i + 0.5 + 0.01 * i, updates=10 ** 9) for i in range(100)}
    s = summarize(res)
    assert s["requests"] == 100
    assert abs(s["p50_s"] - (0.5 + 0.01 * 49.5)) < 0.02
    assert s["p99_s"] >= s["p50_s"]
    assert abs(s["gpt_updates_per_s"] - 100 / s["makespan_s"]) < 1e-9
