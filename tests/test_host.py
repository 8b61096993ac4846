"""Host-side logic (CPU only): input builders against the reference's golden
vectors, index handling, argument checks, and the C-ABI library's exports."""

from __future__ import annotations

import ctypes
import json
import os
import re

import numpy as np
import pytest

from helpers import IdxGeom, golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ring_indices_match_reference_including_ties():
    from openusct_amd.geometry import RingArray
    g = golden("geometry.npz")
    i = 0
    while f"ring{i}_params" in g:
        ne, r, d, h = g[f"ring{i}_params"]
        ring = RingArray(n_elements=int(ne), radius_m=float(r), domain_m=float(d), h=float(h))
        assert ring.n == int(g[f"ring{i}_n"])
        assert np.array_equal(np.array(ring.idx), g[f"ring{i}_idx"]), i
        assert np.array_equal(ring.element_positions, g[f"ring{i}_pos"])
        i += 1
    assert i >= 7


def test_cylinder_indices_match_reference():
    from openusct_amd.geometry import CylinderArray
    g = golden("geometry.npz")
    i = 0
    while f"cyl{i}_params" in g:
        nr, pr, r, ht, d, h = g[f"cyl{i}_params"]
        cyl = CylinderArray(n_rings=int(nr), per_ring=int(pr), radius_m=float(r),
                            height_m=float(ht), domain_m=float(d), h=float(h))
        assert cyl.n == int(g[f"cyl{i}_n"])
        assert np.array_equal(np.array(cyl.idx), g[f"cyl{i}_idx"]), i
        assert np.array_equal(cyl.element_positions, g[f"cyl{i}_pos"])
        i += 1
    assert i >= 4


def test_c1_has_half_integer_ties_resolved_half_even():
    """C1 element 0 sits at x/h = 237.5: Python's round goes to the even 238."""
    from openusct_amd.geometry import RingArray
    ring = RingArray(64, 0.022, 0.051, 2e-4)
    assert ring.xy[0][0] / 2e-4 == pytest.approx(237.5)
    assert ring.idx[0][1] == 238


def test_remap_matches_fwi_from_dataset():
    from openusct_amd.geometry import RingArray, remap_positions
    rm = golden("remap.npz")
    i = 0
    while f"remap{i}_params" in rm:
        ne, r, d, h = rm[f"remap{i}_params"]
        ring = RingArray(int(ne), float(r), float(d), float(h))
        idx = remap_positions(ring.element_positions, ring.n, float(h))
        assert np.array_equal(np.array(idx), rm[f"remap{i}_idx"])
        if f"remap{i}_fwi_idx" in rm:
            assert np.array_equal(np.array(idx), rm[f"remap{i}_fwi_idx"])
        i += 1


def test_ricker_matches_reference():
    from openusct_amd.sources import ricker
    s = golden("sources.npz")
    i = 0
    while f"ricker{i}_params" in s:
        nt, dt, f0, t0 = s[f"ricker{i}_params"]
        w = ricker(int(nt), float(dt), float(f0), None if np.isnan(t0) else float(t0))
        assert np.array_equal(w, s[f"ricker{i}"])
        i += 1


def test_phantoms_and_sponge_match_reference():
    from openusct_amd import phantom
    from openusct_amd.solver import make_sponge
    p = golden("phantom.npz")
    c = phantom.coupling_background((61, 61), 3000.0, 1480.0, 0.02, 1e-3)
    assert np.array_equal(c, p["coupling"])
    assert np.array_equal(phantom.add_inclusion(c, (0.58, 0.46), 0.006, 2650.0, 1e-3),
                          p["inclusion"])
    c3 = phantom.cylinder_background((21, 21, 21), 3000.0, 1480.0, 0.008, 1.5e-3)
    assert np.array_equal(c3, p["cylinder"])
    assert np.array_equal(phantom.add_sphere(c3, (0.56, 0.5, 0.45), 0.004, 2700.0, 1.5e-3),
                          p["sphere"])
    assert np.array_equal(phantom.velocity_to_m(p["inclusion"]), p["m_of_c"])
    assert np.array_equal(phantom.m_to_velocity(p["m_of_c"]), p["c_of_m"])
    assert np.array_equal(make_sponge((61, 61)), p["sponge_61"])
    assert np.array_equal(make_sponge((40, 70), width=10, strength=0.1), p["sponge_40x70"])


def test_seeded_phantoms_are_reproducible_and_in_range():
    from openusct_amd import configs
    a, b = configs.c1(nt_scale=0.01), configs.c1(nt_scale=0.01)
    assert np.array_equal(a.c_true, b.c_true)
    vals = np.unique(a.c_true)
    assert vals.min() >= 1400.0 and vals.max() <= 1800.0 and 1500.0 in vals
    assert len(vals) >= 3
    c4 = configs.c4(n=64, nt_scale=0.01)
    assert 1400.0 <= c4.c_true.min() and c4.c_true.max() <= 1800.0


def test_configs_satisfy_cfl_and_shapes():
    from openusct_amd import configs
    for name, shape, nshots in [("C1", (256, 256), 8), ("C2", (512, 512), 128),
                                ("C3", (1024, 1024), 256)]:
        p = configs.BUILDERS[name](nt_scale=0.001)
        assert p.shape == shape and len(p.src_list) == nshots
        assert 1800.0 * p.dt / p.h <= 1 / np.sqrt(2)
    p4 = configs.c4(nt_scale=0.001)
    assert p4.shape == (256, 256, 256) and len(p4.src_list) == 64 and p4.geom.n_elements == 512
    assert 1800.0 * p4.dt / p4.h <= 1 / np.sqrt(3)


def test_linear_indices_row_major_and_bounds():
    from openusct_amd.plan import linear_indices
    assert list(linear_indices([(1, 2), (0, 0), (3, 4)], (4, 5))) == [7, 0, 19]
    assert list(linear_indices([(1, 2, 3)], (4, 5, 6))) == [1 * 30 + 2 * 6 + 3]
    with pytest.raises(ValueError):
        linear_indices([(4, 0)], (4, 5))


def test_cfl_check_rejects_unstable_dt():
    from openusct_amd.fwi import check_cfl
    m = np.full((8, 8), 1.0 / 1500.0 ** 2)
    check_cfl(m, 5e-8, 2e-4)  # Courant 0.375
    with pytest.raises(ValueError):
        check_cfl(m, 1e-7, 2e-4)  # Courant 0.75 > 0.707
    with pytest.raises(ValueError):
        check_cfl(np.zeros((4, 4)), 1e-8, 1e-3)


def test_fwi_rejects_unsupported_options_before_touching_the_device():
    from openusct_amd import fwi
    g = IdxGeom([(1, 1)])
    m = np.full((8, 8), 1.0 / 1500.0 ** 2)
    with pytest.raises(ValueError):
        fwi.misfit_and_gradient(m, g, np.zeros(4), 1e-8, 1e-3, 4, np.zeros((1, 4, 1)),
                                misfit_type="wasserstein", time_weights=[np.ones((4, 1))])
    with pytest.raises(ValueError):
        fwi.misfit(m, g, np.zeros(4), 1e-8, 1e-3, 4, np.zeros((1, 4, 1)), misfit_type="l1")
    with pytest.raises(ValueError):  # one footprint per element
        fwi._element_groups(g, (8, 8), None, [])
    with pytest.raises(ValueError):  # a weight per point
        fwi._element_groups(g, (8, 8), None, [([(1, 1), (1, 2)], np.array([1.0]))])
    with pytest.raises(ValueError):
        fwi.invert(m, g, np.zeros(4), 1e-8, 1e-3, 4, np.zeros((1, 4, 1)),
                   optimizer="gauss-newton", misfit_type="gcn")


def test_plugin_registry():
    from openusct_amd import plugins
    assert "fwi" in plugins.available()
    assert plugins.info("fwi")["output"] == "model"
    with pytest.raises(KeyError):
        plugins.run("nope", None)

    @plugins.register("echo", description="test")
    def _echo(ds, **kw):
        return ds, kw

    assert plugins.run("echo", 3, a=1) == (3, {"a": 1})


def test_library_loads_and_exports_every_declared_symbol():
    from openusct_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "usct.h")).read()
    declared = set(re.findall(r"^(?:int|int64_t|double|const char \*)\s*(usct_\w+)\(", header, re.M))
    assert declared == set(_lib.EXPORTS)
    assert lib.usct_gsot_workspace_bytes(2, 10, 3) == 6 * 11 * (5 * 8 + 3 * 4)
    assert lib.usct_gsot_workspace_bytes(1, 0, 1) == -1
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.usct_version() == 1
    # an error path that needs no device: bad dims are rejected before any HIP call
    plan = ctypes.c_void_p()
    dims = (ctypes.c_int * 2)(0, 5)
    rc = lib.usct_plan_create(ctypes.byref(plan), 2, dims, 1e-3, 1e-7, 10, None, 0, None, 0,
                              2, 0, 0, 0)
    assert rc == _lib.USCT_EINVAL
    assert b"dims" in lib.usct_last_error()


def test_bandwidth_library_exports_every_declared_symbol():
    """libusct_bw.so (the HBM streaming kernels bench.py measures the device's
    ceiling with) loads and exports what include/usct_bw.h declares; a bad
    argument is refused before any HIP call."""
    lib = ctypes.CDLL(os.path.join(ROOT, "openusct_amd", "libusct_bw.so"))
    header = open(os.path.join(ROOT, "include", "usct_bw.h")).read()
    declared = set(re.findall(r"^(?:int|int64_t)\s+(usct_\w+)\(", header, re.M))
    assert declared == {"usct_bw_run", "usct_bw_run_pass", "usct_bw_pass_groups",
                        "usct_bw_device_cus"}
    for name in declared:
        assert hasattr(lib, name), name
    ms = ctypes.c_float()
    args = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
            ctypes.c_void_p]
    lib.usct_bw_run.argtypes = args
    lib.usct_bw_run_pass.argtypes = args
    assert lib.usct_bw_run(2, None, None, None, 1000, 8, 0, 1, ctypes.byref(ms), None) == 1  # % 16
    assert lib.usct_bw_run(7, None, None, None, 1024, 8, 0, 1, ctypes.byref(ms), None) == 1
    assert lib.usct_bw_run_pass(2, None, None, None, 1000, 1, 0, 1, ctypes.byref(ms), None) == 1
    assert lib.usct_bw_run_pass(2, None, None, None, 1024, 3, 0, 1, ctypes.byref(ms), None) == 1
    lib.usct_bw_pass_groups.restype = ctypes.c_int64
    lib.usct_bw_pass_groups.argtypes = [ctypes.c_int64, ctypes.c_int]
    # 4 GiB in 16-byte elements, 256 lanes per work-group, one access per lane
    assert lib.usct_bw_pass_groups(4 << 30, 1) == (4 << 30) // 16 // 256
    assert lib.usct_bw_pass_groups(4096 * 2 + 16, 2) == 2


def test_kernel_min_bytes_model():
    """bench.py's per-launch minimum bytes: per shot-cell 24 B for both two-level 2D
    kernels, 18 B for the 3D forward, 30 B for the reversing 3D adjoint; per cell
    the S' pair and, for adjoints, the gradient accumulators of each lane group."""
    import bench
    N, B = 1000, 64
    assert bench.kernel_min_bytes("k_adj2d_x2", N, B, 1) == (16 * N * B + 8 * N + 8 * N, 8 * N * B + 8 * N)
    assert bench.kernel_min_bytes("k_fwd2d_x2", N, B, 1) == (12 * N * B + 8 * N, 12 * N * B)
    assert sum(bench.kernel_min_bytes("k_fwd3d", N, B, 1)) == 18 * N * B + 8 * N
    assert sum(bench.kernel_min_bytes("k_radj3d", N, B, 2)) == 30 * N * B + 8 * N + 32 * N


def test_footprint_model_orders():
    from openusct_amd.plan import footprint_bytes
    small = footprint_bytes((256, 256), 1000, 8, 64)
    big = footprint_bytes((512, 512), 1000, 128, 128)
    assert 0 < small < big
    assert footprint_bytes((512, 512), 1000, 128, 128, ckpt_every=50) < \
        footprint_bytes((512, 512), 1000, 128, 128, ckpt_every=999)


def test_build_footprints_match_reference():
    """geometry.py:212-278 point lists (order, rounding, clipping, disc trimming)."""
    from openusct_amd.geometry import CylinderArray, RingArray, build_footprints
    f = golden("footprints.npz")
    i = 0
    while f"case{i}_kind" in f:
        prm = f[f"case{i}_params"]
        if str(f[f"case{i}_kind"]) == "ring":
            g = RingArray(int(prm[0]), float(prm[1]), float(prm[2]), float(prm[3]))
        else:
            g = CylinderArray(int(prm[0]), int(prm[1]), *[float(x) for x in prm[2:]])
        width, height, mp = f[f"case{i}_args"]
        fps = build_footprints(g, float(width), str(f[f"case{i}_shape"]),
                               height_m=None if np.isnan(height) else float(height),
                               max_pts=int(mp))
        cells = np.array([ix for idxs, _ in fps for ix in idxs], dtype=np.int64)
        counts = np.array([len(idxs) for idxs, _ in fps])
        w = np.concatenate([wt for _, wt in fps])
        assert np.array_equal(counts, f[f"case{i}_counts"]), i
        assert np.array_equal(cells, f[f"case{i}_cells"]), i
        assert np.array_equal(w, f[f"case{i}_w"]), i
        i += 1
    assert i >= 10


def test_p_window_weights_match_reference():
    from openusct_amd import fwi
    f = golden("aperture_2d.npz")
    tw = fwi.p_window_weights(f["pt_dobs"], float(f["dt"]), float(f["f0"]))
    assert np.max(np.abs(np.stack(tw) - f["tw"])) <= 1e-12
    twf = fwi.p_window_weights(f["nosponge_dobs"], float(f["dt"]), float(f["f0"]),
                               onset_frac=0.15, keep_cycles=2.0)
    assert np.max(np.abs(np.stack(twf) - f["twf"])) <= 1e-12
    assert all(((w >= 0) & (w <= 1)).all() for w in tw)


def test_element_groups_to_csr():
    from openusct_amd import fwi
    from openusct_amd.geometry import RingArray, build_footprints
    from openusct_amd.plan import _csr
    g = RingArray(n_elements=8, radius_m=0.024, domain_m=0.06, h=1e-3)
    fps = build_footprints(g, 0.003, "flat")
    tx, rec = fwi._element_groups(g, (g.n, g.n), [0, 4], fps)
    assert len(tx) == 2 and len(rec) == 8
    assert np.array_equal(tx[1][0], rec[4][0])
    off, cells, w = _csr(rec)
    assert off[0] == 0 and off[-1] == len(cells) == sum(len(c) for c, _ in fps)
    for j, (c, wt) in enumerate(rec):
        assert np.array_equal(cells[off[j]:off[j + 1]], c)
        assert np.array_equal(w[off[j]:off[j + 1]], wt)
    with pytest.raises(ValueError):
        _csr([(np.array([3]), np.array([]))])


def test_group_plan_rejects_bad_groups_before_any_device_call():
    """usct_plan_create_groups checks its CSR groups on the host: empty groups,
    cells outside the grid, non-finite weights and a bad first offset are
    USCT_EINVAL with a message, before any HIP call."""
    from openusct_amd import _lib
    lib = _lib.load()
    dims = (ctypes.c_int * 2)(8, 8)
    i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731

    def create(tx_off, tx_cells, tx_w, rec_off, rec_cells, rec_w):
        plan = ctypes.c_void_p()
        P = ctypes.POINTER(ctypes.c_int64)
        rc = lib.usct_plan_create_groups(
            ctypes.byref(plan), 2, dims, 1e-3, 1e-7, 10, len(tx_off) - 1,
            tx_off.ctypes.data_as(P), tx_cells.ctypes.data_as(P), tx_w.ctypes.data,
            len(rec_off) - 1, rec_off.ctypes.data_as(P), rec_cells.ctypes.data_as(P),
            rec_w.ctypes.data, 2, 0, 0, 0)
        return rc, lib.usct_last_error().decode()

    good = (i64([0, 2]), i64([9, 10]), f64([0.5, 0.5]))
    cases = {
        "empty": ((i64([0, 0]), i64([0]), f64([1.0])), "empty"),
        "outside": ((i64([0, 1]), i64([64]), f64([1.0])), "outside the grid"),
        "weight": ((i64([0, 1]), i64([3]), f64([np.nan])), "not finite"),
        "offset": ((i64([1, 2]), i64([3, 4]), f64([1.0, 1.0])), "off[0]"),
    }
    for name, (bad, msg) in cases.items():
        rc, err = create(*bad, *good)
        assert rc == _lib.USCT_EINVAL and msg in err, (name, rc, err)
        rc, err = create(*good, *bad)
        assert rc == _lib.USCT_EINVAL and msg in err, (name, rc, err)


@pytest.mark.parametrize("mt", ["envelope", "egcn", "traveltime"])
def test_device_misfit_values_on_reference_traces(mt):
    """openusct_amd.misfits on the reference's own traces at m0 gives the
    reference's misfit (fwi.py:134-263; tests/golden/misfits_2d.npz), with time
    windows and with a sponge run (torch on the CPU here; the same code runs on
    the device)."""
    import torch
    from openusct_amd import misfits
    f = golden("misfits_2d.npz")
    base = golden("fmc_2d.npz")
    d0, dobs = torch.tensor(f["dsyn0"]), torch.tensor(base["nosponge_dobs"])
    J, adj = misfits.misfit_and_source(d0, dobs, mt)
    assert abs(J.item() - float(f[f"{mt}_J"])) <= 1e-10 * abs(float(f[f"{mt}_J"]))
    assert adj.shape == d0.shape and torch.isfinite(adj).all()
    if mt == "envelope":
        J, _ = misfits.misfit_and_source(d0, dobs, mt, torch.tensor(f["tw"]))
        assert abs(J.item() - float(f["tw_envelope_J"])) <= 1e-10 * float(f["tw_envelope_J"])
    if mt == "traveltime":
        J, _ = misfits.misfit_and_source(torch.tensor(f["sponge_dsyn0"]),
                                         torch.tensor(base["sponge_dobs"]), mt)
        assert abs(J.item() - float(f["sponge_traveltime_J"])) <= 1e-10 * float(f["sponge_traveltime_J"])


def test_unknown_misfits_raise():
    from openusct_amd import fwi
    with pytest.raises(ValueError):
        fwi._check_misfit("l3")
    for mt in ("l2", "gcn", "envelope", "egcn", "traveltime", "gsot"):
        fwi._check_misfit(mt)


def test_other_wavelets_phantoms_and_shell_arrays_match_reference():
    """sources.gabor / hann_tone_burst, phantom.smooth / sphere_background /
    voronoi_polycrystal(_3d) from seeded generators, Shell/Hemisphere/Sphere
    arrays and footprints on them: bit-exact against the reference
    (tests/golden/builders.npz)."""
    from openusct_amd import phantom
    from openusct_amd.geometry import HemisphereArray, ShellArray, SphereArray, build_footprints
    from openusct_amd.sources import gabor, hann_tone_burst
    f = golden("builders.npz")
    assert np.array_equal(gabor(400, 5e-8, 5e5), f["gabor0"])
    assert np.array_equal(gabor(333, 2.5e-8, 1.2e6, frac_bw=0.9, t0=2e-6), f["gabor1"])
    assert np.array_equal(hann_tone_burst(300, 1e-7, 4e5), f["hann0"])
    assert np.array_equal(hann_tone_burst(500, 5e-8, 7.5e5, n_cycles=3, t0=1.3e-6), f["hann1"])
    assert np.array_equal(phantom.smooth(f["smooth_in"], 2.0), f["smooth_out"])
    assert np.array_equal(phantom.sphere_background((17, 19, 21), 3000.0, 1480.0, 0.006, 1e-3),
                          f["sphere_bg"])
    lab, ang, th = phantom.voronoi_polycrystal((61, 61), 12, 0.02, 1e-3, rng=np.random.default_rng(7),
                                               relax=2)
    assert np.array_equal(lab, f["vor_labels"]) and np.array_equal(ang, f["vor_angles"])
    assert np.array_equal(th, f["vor_theta"])
    lab, axes, col = phantom.voronoi_polycrystal_3d((15, 21, 21), 9, 0.008, 1e-3,
                                                    rng=np.random.default_rng(8), relax=1)
    assert np.array_equal(lab, f["vor3_labels"]) and np.array_equal(axes, f["vor3_axes"])
    assert np.array_equal(col, f["vor3_colat"])
    with pytest.raises(ValueError):
        phantom.voronoi_polycrystal((5, 5, 5), 2, 0.001, 1e-3)
    arrays = [SphereArray(40, 0.012, 0.03, 1e-3), HemisphereArray(33, 0.011, 0.03, 7.5e-4),
              ShellArray(64, 0.02, 0.0512, 2e-4)]
    for i, a in enumerate(arrays):
        assert np.array_equal(np.array(a.idx), f[f"shell{i}_idx"]), i
        assert np.array_equal(a.element_positions, f[f"shell{i}_pos"]), i
        assert a.n == int(f[f"shell{i}_n"]) and a.array_type == str(f[f"shell{i}_type"])
    fps = build_footprints(arrays[0], 0.004, "disc", height_m=0.003, max_pts=7)
    assert np.array_equal(np.array([ix for idxs, _ in fps for ix in idxs]), f["shell_fp_cells"])
    assert np.array_equal(np.array([len(idxs) for idxs, _ in fps]), f["shell_fp_counts"])
    assert np.array_equal(np.concatenate([w for _, w in fps]), f["shell_fp_w"])


def test_reduce_pmc_scales_counters_and_keeps_each_runs_tuned_grid(tmp_path):
    """bench.py --reduce-pmc: each counter is divided by what it reads for the
    calibration copy (FETCH_SIZE 0.5, WRITE_SIZE 1.0 of the bytes moved here), each
    kernel keeps the median over its run's most frequent grid (the two runs may tune
    to different grids), and runtime copies and fills are left out."""
    import csv
    import bench
    cols = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]

    def write(path, counter, rows):
        with open(path, "w", newline="") as fh:
            w = csv.writer(fh)
            w.writerow(cols)
            for i, (grid, name, kib) in enumerate(rows):
                w.writerow([i, grid, name, counter, kib])

    mib = 1 << 20
    fetch, wr = tmp_path / "f.csv", tmp_path / "w.csv"
    write(fetch, "FETCH_SIZE", [(4096, "k_bw_copy", 0.5 * 1024 * 1024),
                                (10, "k_adj2d_x2", 300.0), (10, "k_adj2d_x2", 310.0),
                                (20, "k_adj2d_x2", 999.0), (10, "k_adj2d_x2", 305.0),
                                (64, "__amd_rocclr_copyBuffer", 8.0)])
    write(wr, "WRITE_SIZE", [(4096, "k_bw_copy", 1.0 * 1024 * 1024),
                             (20, "k_adj2d_x2", 100.0), (20, "k_adj2d_x2", 110.0),
                             (20, "k_adj2d_x2", 90.0), (10, "k_adj2d_x2", 999.0),
                             (64, "__amd_rocclr_copyBuffer", 8.0)])
    out = bench.reduce_pmc(str(fetch), str(wr), str(tmp_path / "o.json"), "test", 1024 * mib)
    assert out["fetch_scale"] == 0.5 and out["write_scale"] == 1.0
    k = out["kernels"]
    assert set(k) == {"k_adj2d_x2"}
    assert k["k_adj2d_x2"]["grid_size"] == 10 and k["k_adj2d_x2"]["grid_size_write_run"] == 20
    assert k["k_adj2d_x2"]["read_bytes"] == 305 * 1024 * 2
    assert k["k_adj2d_x2"]["write_bytes"] == 100 * 1024
    assert json.load(open(tmp_path / "o.json"))["kernels"] == k
    # runs restricted to one tile kind add that kind's entry and keep the rest
    out4 = bench.reduce_pmc(str(fetch), str(wr), str(tmp_path / "o.json"), "test", 1024 * mib,
                            tile_kind=4, only=["k_adj2d_x2", "k_other"])
    saved = json.load(open(tmp_path / "o.json"))
    assert saved["kernels"] == k and out4["kernels"] == k
    assert saved["kernels_tile_kind"]["4"] == {"k_adj2d_x2": k["k_adj2d_x2"]}
