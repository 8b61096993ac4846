"""Write the golden fixtures under tests/golden/ from the reference's Python path.

Run in a container that holds the reference checkout (never on a GPU host):

    python tests/golden/make_golden.py [/path/to/reference] [--only=base,apertures]

It imports ``ringfwi`` from ``<reference>/simulation`` and records inputs and
outputs of the functions whose behaviour only the reference itself pins:

* the 2D sponge path (simulation/ringfwi/solver.py:22-42, fwi.py:68-69,313-314),
  whose gradient is the reference's inexact formula (fmc_2d.npz holds it on the
  test_gradient.py grid, where the sponge covers almost every cell, and
  sponge_2d.npz on a grid where it is a border layer);
* element-index rounding of RingArray / CylinderArray (geometry.py:47-50,118-122)
  and the dataset re-mapping of fwi_from_dataset (fwi.py:744-746);
* phantom builders (phantom.py:13-48,166-204) and the Ricker wavelet
  (sources.py:8-26);
* forward / gradient on the reference's own test problems
  (simulation/tests/test_gradient.py:20-40, test_gradient_3d.py:18-37), with
  duplicate receivers (fwi.py:321-322 accumulates them) and the GCN misfit,
  plus a better-resolved 3D cylinder problem (fmc_3d_fine.npz);
* finite-aperture footprints (geometry.py:212-278) and forward / gradient with
  them in 2D and 3D (fwi.py:74-89, 316-319), P-window time weights
  (fwi.py:670-699) and windowed misfits (fwi.py:255-262, 298-301), and
  Gauss-Newton products with a sponge and with footprints (fwi.py:410-431),
  whose sponge part is the reference's own formula;
* the envelope, normalised-envelope (egcn), traveltime and graph-space optimal
  transport (gsot) misfits and their gradients (fwi.py:134-243) on the
  fmc_2d.npz problem, and the envelope misfit with the footprints of
  aperture_2d.npz (misfits_2d.npz);
* the other wavelets, phantoms and array geometries of the reference
  (builders.npz);
* short optimiser trajectories of invert (fwi.py:535-667): steepest descent
  with the masked, Gaussian-smoothed (scipy) gradient, bounds and the halving
  line search, and truncated Gauss-Newton, in 2D (with and without a sponge)
  and 3D (invert.npz).

The fixtures are data only (numpy .npz, no pickles).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    sys.path.insert(0, os.path.join(ref_root, "simulation"))
    from ringfwi import fwi, phantom  # noqa: E402
    from ringfwi.geometry import CylinderArray, GridArray, RingArray  # noqa: E402
    from ringfwi.solver import make_sponge  # noqa: E402
    from ringfwi.sources import ricker  # noqa: E402
    from ringfwi.acquire import simulate_dataset  # noqa: E402

    # -- geometry: element indices (tie rounding) -----------------------------
    rings = [(8, 0.024, 0.06, 1e-3), (16, 0.024, 0.060, 6e-4), (12, 0.016, 0.045, 1.25e-3),
             (64, 0.022, 0.051, 2e-4), (128, 0.022, 0.1022, 2e-4),
             (256, 0.044, 0.1023, 1e-4), (128, 0.044, 0.1023, 1e-4)]
    cyls = [(2, 6, 0.010, 0.012, 0.03, 1.5e-3), (3, 12, 0.015, 0.014, 0.036, 1.25e-3),
            (8, 64, 0.028, 0.028, 0.06375, 2.5e-4), (4, 32, 0.014, 0.014, 0.032, 2.5e-4)]
    geo = {}
    for i, (ne, r, d, h) in enumerate(rings):
        g = RingArray(n_elements=ne, radius_m=r, domain_m=d, h=h)
        geo[f"ring{i}_params"] = np.array([ne, r, d, h])
        geo[f"ring{i}_n"] = np.array(g.n)
        geo[f"ring{i}_idx"] = np.array(g.idx, dtype=np.int64)
        geo[f"ring{i}_pos"] = g.element_positions
    for i, (nr, pr, r, ht, d, h) in enumerate(cyls):
        g = CylinderArray(n_rings=nr, per_ring=pr, radius_m=r, height_m=ht, domain_m=d, h=h)
        geo[f"cyl{i}_params"] = np.array([nr, pr, r, ht, d, h])
        geo[f"cyl{i}_n"] = np.array(g.n)
        geo[f"cyl{i}_idx"] = np.array(g.idx, dtype=np.int64)
        geo[f"cyl{i}_pos"] = g.element_positions
    np.savez_compressed(os.path.join(HERE, "geometry.npz"), **geo)

    # -- sources --------------------------------------------------------------
    src = {}
    for i, (nt, dt, f0, t0) in enumerate([(350, 1.2e-7, 0.4e6, None), (1000, 5e-8, 5e5, None),
                                          (4000, 2.5e-8, 5e5, None), (300, 1e-7, 0.3e6, 2e-6)]):
        src[f"ricker{i}_params"] = np.array([nt, dt, f0, np.nan if t0 is None else t0])
        src[f"ricker{i}"] = ricker(nt, dt, f0, t0)
    np.savez_compressed(os.path.join(HERE, "sources.npz"), **src)

    # -- phantoms and sponge --------------------------------------------------
    ph = {}
    n, h = 61, 1e-3
    c = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h)
    ph["coupling"] = c
    ph["inclusion"] = phantom.add_inclusion(c, (0.58, 0.46), 0.006, 2650.0, h)
    n3, h3 = 21, 1.5e-3
    c3 = phantom.cylinder_background((n3, n3, n3), 3000.0, 1480.0, 0.008, h3)
    ph["cylinder"] = c3
    ph["sphere"] = phantom.add_sphere(c3, (0.56, 0.5, 0.45), 0.004, 2700.0, h3)
    ph["m_of_c"] = phantom.velocity_to_m(ph["inclusion"])
    ph["c_of_m"] = phantom.m_to_velocity(ph["m_of_c"])
    ph["sponge_61"] = make_sponge((61, 61))
    ph["sponge_40x70"] = make_sponge((40, 70), width=10, strength=0.1)
    np.savez_compressed(os.path.join(HERE, "phantom.npz"), **ph)

    # -- 2D forward / gradient (test_gradient.py:20-40), with and without sponge
    h = 1.0e-3
    geom = RingArray(n_elements=8, radius_m=0.024, domain_m=0.06, h=h)
    n = geom.n
    dt, nt, f0 = 1.2e-7, 350, 0.4e6
    wav = ricker(nt, dt, f0)
    c_true = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h)
    c_true = phantom.add_inclusion(c_true, (0.5, 0.5), 0.006, 3400.0, h)
    m_true = phantom.velocity_to_m(c_true)
    m0 = phantom.velocity_to_m(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h))
    src_list = [0, 4]
    out = dict(idx=np.array(geom.idx, dtype=np.int64), h=h, dt=dt, nt=nt, wavelet=wav,
               m_true=m_true, m0=m0, src_list=np.array(src_list))
    for tag, sp in (("nosponge", None), ("sponge", make_sponge((n, n)))):
        dobs = fwi.forward_fmc(m_true, geom, wav, dt, h, nt, sp, src_list)
        J, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, dobs, sp, src_list)
        out[f"{tag}_dobs"] = dobs
        out[f"{tag}_J"] = np.array(J)
        out[f"{tag}_g"] = g
        out[f"{tag}_misfit"] = np.array(fwi.misfit(m0, geom, wav, dt, h, nt, dobs, sp, src_list))
    Jg, gg = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, out["nosponge_dobs"], None,
                                     src_list, misfit_type="gcn")
    out["gcn_J"] = np.array(Jg)
    out["gcn_g"] = gg
    np.savez_compressed(os.path.join(HERE, "fmc_2d.npz"), **out)

    # -- duplicate receivers: two elements on one grid cell --------------------
    idx = list(geom.idx) + [geom.idx[2], geom.idx[2]]
    dup = GridArray(idx=idx, n=n, h=h, radius_m=0.024, domain_m=0.06)
    dobs = fwi.forward_fmc(m_true, dup, wav, dt, h, nt, None, [2, 9])
    J, g = fwi.misfit_and_gradient(m0, dup, wav, dt, h, nt, dobs, None, [2, 9])
    np.savez_compressed(os.path.join(HERE, "dup_2d.npz"), idx=np.array(idx, dtype=np.int64),
                        src_list=np.array([2, 9]), dobs=dobs, J=np.array(J), g=g)

    # -- 3D forward / gradient (test_gradient_3d.py:18-37) --------------------
    h = 1.5e-3
    cyl = CylinderArray(n_rings=2, per_ring=6, radius_m=0.010, height_m=0.012,
                        domain_m=0.03, h=h)
    n = cyl.n
    dt, nt = 1.5e-7, 140
    wav = ricker(nt, dt, 0.35e6)
    c_true = phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.008, h)
    c_true = phantom.add_sphere(c_true, (0.5, 0.5, 0.5), 0.004, 2700.0, h)
    m_true = phantom.velocity_to_m(c_true)
    m0 = phantom.velocity_to_m(phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.008, h))
    src_list = [0, 6]
    dobs = fwi.forward_fmc(m_true, cyl, wav, dt, h, nt, src_list=src_list)
    J, g = fwi.misfit_and_gradient(m0, cyl, wav, dt, h, nt, dobs, src_list=src_list)
    np.savez_compressed(os.path.join(HERE, "fmc_3d.npz"), idx=np.array(cyl.idx, dtype=np.int64),
                        h=h, dt=dt, nt=nt, wavelet=wav, m_true=m_true, m0=m0,
                        src_list=np.array(src_list), dobs=dobs, J=np.array(J), g=g)

    # -- 2D sponge on a grid where the sponge is a border layer (width 25 of 121)
    h = 5e-4
    geom = RingArray(n_elements=8, radius_m=0.012, domain_m=0.06, h=h)
    n = geom.n
    dt, nt = 1e-7, 500
    wav = ricker(nt, dt, 0.25e6)
    c0 = phantom.coupling_background((n, n), 2000.0, 1480.0, 0.008, h)
    m0 = phantom.velocity_to_m(c0)
    m_true = phantom.velocity_to_m(phantom.add_inclusion(c0, (0.52, 0.5), 0.003, 2300.0, h))
    sp = make_sponge((n, n))
    src_list = [0, 3]
    dobs = fwi.forward_fmc(m_true, geom, wav, dt, h, nt, sp, src_list)
    J, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, dobs, sp, src_list)
    np.savez_compressed(os.path.join(HERE, "sponge_2d.npz"), idx=np.array(geom.idx, dtype=np.int64),
                        h=h, dt=dt, nt=nt, wavelet=wav, m_true=m_true, m0=m0, sponge=sp,
                        src_list=np.array(src_list), dobs=dobs, J=np.array(J), g=g)

    # -- 3D at about 10 points per wavelength (test_gradient_3d.py's problem is ~3)
    h = 5e-4
    cyl = CylinderArray(n_rings=2, per_ring=8, radius_m=0.007, height_m=0.006, domain_m=0.02, h=h)
    n = cyl.n
    dt, nt = 1e-7, 200
    wav = ricker(nt, dt, 0.3e6)
    c0 = phantom.cylinder_background((n, n, n), 1500.0, 1480.0, 0.006, h)
    m0 = phantom.velocity_to_m(c0)
    m_true = phantom.velocity_to_m(phantom.add_sphere(c0, (0.55, 0.5, 0.5), 0.002, 1650.0, h))
    src_list = [0, 9]
    dobs = fwi.forward_fmc(m_true, cyl, wav, dt, h, nt, src_list=src_list)
    J, g = fwi.misfit_and_gradient(m0, cyl, wav, dt, h, nt, dobs, src_list=src_list)
    np.savez_compressed(os.path.join(HERE, "fmc_3d_fine.npz"), idx=np.array(cyl.idx, dtype=np.int64),
                        h=h, dt=dt, nt=nt, wavelet=wav, m_true=m_true, m0=m0,
                        src_list=np.array(src_list), dobs=dobs, J=np.array(J), g=g)

    # -- dataset re-mapping in fwi_from_dataset (fwi.py:732-746) --------------
    rm = {}
    for i, (ne, r, d, hh) in enumerate([(12, 0.016, 0.045, 1.25e-3), (64, 0.022, 0.051, 2e-4)]):
        ring = RingArray(n_elements=ne, radius_m=r, domain_m=d, h=hh)
        nn = ring.n
        wv = ricker(60, 2e-7, 0.4e6)
        ds = simulate_dataset(np.full((nn, nn), 1500.0), ring, wv[:2], 2e-7, 1500.0,
                              src_list=[0]) if nn < 100 else None
        # Re-map positions exactly as fwi_from_dataset does, from stored positions.
        domain = (nn - 1) * hh
        abspos = ring.element_positions + domain / 2.0
        rm[f"remap{i}_params"] = np.array([ne, r, d, hh])
        rm[f"remap{i}_idx"] = np.array([tuple(int(round(p[a] / hh)) for a in reversed(range(2)))
                                        for p in abspos], dtype=np.int64)
        if ds is not None:
            res = fwi.fwi_from_dataset(ds, n_iter=0)
            rm[f"remap{i}_fwi_idx"] = np.array(res["geom"].idx, dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "remap.npz"), **rm)
    print("golden fixtures written to", HERE)


def main_apertures(ref_root):
    """Footprints, P-window time weights and Gauss-Newton products."""
    sys.path.insert(0, os.path.join(ref_root, "simulation"))
    from ringfwi import fwi, phantom  # noqa: E402
    from ringfwi.geometry import CylinderArray, RingArray, build_footprints  # noqa: E402
    from ringfwi.solver import make_sponge  # noqa: E402
    from ringfwi.sources import ricker  # noqa: E402

    def flat(fps):
        cells = np.array([ix for idxs, _ in fps for ix in idxs], dtype=np.int64)
        counts = np.array([len(idxs) for idxs, _ in fps], dtype=np.int64)
        w = np.concatenate([np.asarray(wt, dtype=np.float64) for _, wt in fps])
        return cells, counts, w

    # -- build_footprints (geometry.py:212-278) ------------------------------
    fp = {}
    cases = [("ring", (8, 0.024, 0.06, 1e-3), 0.003, "flat", None, 9),
             ("ring", (8, 0.024, 0.06, 1e-3), 0.003, "rect", 0.001, 9),
             ("ring", (8, 0.024, 0.06, 1e-3), 0.0, "point", None, 9),
             ("ring", (8, 0.024, 0.06, 1e-3), 0.005, "disc", None, 9),
             ("ring", (64, 0.022, 0.051, 2e-4), 0.0011, "rect", None, 9),
             ("ring", (128, 0.022, 0.1022, 2e-4), 0.003, "flat", None, 5),
             ("cyl", (2, 8, 0.010, 0.012, 0.028, 0.028 / 33.0), 0.004, "rect", 0.004, 9),
             ("cyl", (2, 8, 0.010, 0.012, 0.028, 0.028 / 33.0), 0.004, "disc", 0.004, 9),
             ("cyl", (2, 8, 0.010, 0.012, 0.028, 0.028 / 33.0), 0.002, "rect", 0.006, 9),
             ("cyl", (2, 6, 0.010, 0.012, 0.03, 1.5e-3), 0.003, "flat", None, 9),
             ("cyl", (8, 64, 0.028, 0.028, 0.06375, 2.5e-4), 0.001, "disc", 0.0015, 9)]
    for i, (kind, prm, width, shape, height, mp) in enumerate(cases):
        g = RingArray(*prm) if kind == "ring" else CylinderArray(*prm)
        cells, counts, w = flat(build_footprints(g, width, shape, height_m=height, max_pts=mp))
        fp[f"case{i}_kind"] = np.array(kind)
        fp[f"case{i}_params"] = np.array(prm, dtype=np.float64)
        fp[f"case{i}_args"] = np.array([width, np.nan if height is None else height, mp])
        fp[f"case{i}_shape"] = np.array(shape)
        fp[f"case{i}_cells"] = cells
        fp[f"case{i}_counts"] = counts
        fp[f"case{i}_w"] = w
    np.savez_compressed(os.path.join(HERE, "footprints.npz"), **fp)

    # -- apertures and time windows on test_gradient_aperture.py's problem -----
    h = 1.0e-3
    geom = RingArray(n_elements=8, radius_m=0.024, domain_m=0.06, h=h)
    n = geom.n
    dt, nt, f0 = 1.2e-7, 350, 0.4e6
    wav = ricker(nt, dt, f0)
    c_true = phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h)
    c_true = phantom.add_inclusion(c_true, (0.5, 0.5), 0.006, 3400.0, h)
    m_true = phantom.velocity_to_m(c_true)
    m0 = phantom.velocity_to_m(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.02, h))
    fps = build_footprints(geom, width_m=0.003, shape="flat")
    cells, counts, w = flat(fps)
    src_list = [0, 4]
    sp = make_sponge((n, n), width=8)
    out = dict(idx=np.array(geom.idx, dtype=np.int64), h=h, dt=dt, nt=nt, f0=f0, wavelet=wav,
               m_true=m_true, m0=m0, src_list=np.array(src_list), fp_cells=cells,
               fp_counts=counts, fp_w=w, sponge=sp)
    for tag, s_ in (("nosponge", None), ("sponge", sp)):
        dobs = fwi.forward_fmc(m_true, geom, wav, dt, h, nt, s_, src_list, footprints=fps)
        out[f"{tag}_dobs"] = dobs
        for mt in ("l2", "gcn"):
            J, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, dobs, s_, src_list, mt,
                                           footprints=fps)
            out[f"{tag}_{mt}_J"] = np.array(J)
            out[f"{tag}_{mt}_g"] = g
    # time windows (fwi.py:670-699) on the point-element data, then L2 / GCN with them
    dobs_pt = fwi.forward_fmc(m_true, geom, wav, dt, h, nt, None, src_list)
    tw = fwi.p_window_weights(dobs_pt, dt, f0)
    out["pt_dobs"] = dobs_pt
    out["tw"] = np.stack(tw)
    for mt in ("l2", "gcn"):
        J, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, dobs_pt, None, src_list, mt,
                                       time_weights=tw)
        out[f"tw_{mt}_J"] = np.array(J)
        out[f"tw_{mt}_g"] = g
        out[f"tw_{mt}_misfit"] = np.array(fwi.misfit(m0, geom, wav, dt, h, nt, dobs_pt, None,
                                                     src_list, mt, time_weights=tw))
    # windows and footprints together
    twf = fwi.p_window_weights(out["nosponge_dobs"], dt, f0, onset_frac=0.15, keep_cycles=2.0)
    out["twf"] = np.stack(twf)
    J, g = fwi.misfit_and_gradient(m0, geom, wav, dt, h, nt, out["nosponge_dobs"], None,
                                   src_list, "l2", footprints=fps, time_weights=twf)
    out["twf_J"] = np.array(J)
    out["twf_g"] = g
    # Gauss-Newton products (fwi.py:410-431): v over the whole grid
    rng = np.random.default_rng(8)
    v = rng.standard_normal((n, n)) * float(np.mean(m0))
    out["gn_v"] = v
    nt_gn = 250
    out["gn_nt"] = nt_gn
    out["gn_sponge_hv"] = fwi.gn_hessian_vector(m0, geom, wav[:nt_gn], dt, h, nt_gn, v, sp,
                                                src_list)
    out["gn_fp_hv"] = fwi.gn_hessian_vector(m0, geom, wav[:nt_gn], dt, h, nt_gn, v, None,
                                            src_list, footprints=fps)
    np.savez_compressed(os.path.join(HERE, "aperture_2d.npz"), **out)

    # -- 3D apertures (test_gradient_3d.py's cylinder, rect elements) ---------
    h = 1.5e-3
    cyl = CylinderArray(n_rings=2, per_ring=6, radius_m=0.010, height_m=0.012,
                        domain_m=0.03, h=h)
    n = cyl.n
    dt, nt = 1.5e-7, 140
    wav = ricker(nt, dt, 0.35e6)
    c_true = phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.008, h)
    c_true = phantom.add_sphere(c_true, (0.5, 0.5, 0.5), 0.004, 2700.0, h)
    m_true = phantom.velocity_to_m(c_true)
    m0 = phantom.velocity_to_m(phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.008, h))
    fps = build_footprints(cyl, 0.0045, "rect", height_m=0.003)
    cells, counts, w = flat(fps)
    src_list = [0, 6]
    dobs = fwi.forward_fmc(m_true, cyl, wav, dt, h, nt, None, src_list, footprints=fps)
    J, g = fwi.misfit_and_gradient(m0, cyl, wav, dt, h, nt, dobs, None, src_list,
                                   footprints=fps)
    np.savez_compressed(os.path.join(HERE, "aperture_3d.npz"),
                        idx=np.array(cyl.idx, dtype=np.int64), h=h, dt=dt, nt=nt, wavelet=wav,
                        m_true=m_true, m0=m0, src_list=np.array(src_list), fp_cells=cells,
                        fp_counts=counts, fp_w=w, dobs=dobs, J=np.array(J), g=g)
    print("aperture / window / Gauss-Newton fixtures written to", HERE)


def main_misfits(ref_root):
    """Envelope, normalised-envelope and traveltime misfits (fwi.py:134-243)
    through the reference's misfit / misfit_and_gradient, on the problem of
    fmc_2d.npz (test_gradient.py:20-40), with time windows and with a sponge."""
    sys.path.insert(0, os.path.join(ref_root, "simulation"))
    from ringfwi import fwi  # noqa: E402
    from ringfwi.solver import make_sponge  # noqa: E402

    class Geom:  # the element indices of fmc_2d.npz, as the reference reads a geometry
        def __init__(self, idx):
            self.idx = [tuple(int(v) for v in r) for r in idx]
            self.n_elements = len(self.idx)

        def element_index(self, k):
            return self.idx[k]

    f = np.load(os.path.join(HERE, "fmc_2d.npz"))
    geom = Geom(f["idx"])
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    m0, src = f["m0"], [int(s) for s in f["src_list"]]
    dobs = f["nosponge_dobs"]
    t = np.arange(nt, dtype=np.float64)[:, None]
    win = np.clip((0.75 * nt - t) / (0.1 * nt), 0.0, 1.0) * np.ones((1, geom.n_elements))
    tw = [win, win[::-1].copy()]
    sp = make_sponge(m0.shape, width=8)
    out = {"dsyn0": fwi.forward_fmc(m0, geom, w, dt, h, nt, None, src),
           "sponge_dsyn0": fwi.forward_fmc(m0, geom, w, dt, h, nt, sp, src)}
    for mt in ("envelope", "egcn", "traveltime", "gsot"):
        J, g = fwi.misfit_and_gradient(m0, geom, w, dt, h, nt, dobs, None, src, mt)
        out[f"{mt}_J"], out[f"{mt}_g"] = np.array(J), g
        out[f"{mt}_misfit"] = np.array(fwi.misfit(m0, geom, w, dt, h, nt, dobs, None, src, mt))
    J, g = fwi.misfit_and_gradient(m0, geom, w, dt, h, nt, dobs, None, src, "envelope",
                                   time_weights=tw)
    out["tw_envelope_J"], out["tw_envelope_g"] = np.array(J), g
    J, g = fwi.misfit_and_gradient(m0, geom, w, dt, h, nt, f["sponge_dobs"], sp, src, "traveltime")
    out["sponge_traveltime_J"], out["sponge_traveltime_g"] = np.array(J), g
    # finite-aperture elements (footprints of aperture_2d.npz) with the envelope misfit
    a = np.load(os.path.join(HERE, "aperture_2d.npz"))
    ends = np.cumsum(a["fp_counts"])
    fps = [([tuple(int(v) for v in c) for c in a["fp_cells"][e - k:e]], a["fp_w"][e - k:e])
           for k, e in zip(a["fp_counts"], ends)]
    ag = Geom(a["idx"])
    asrc = [int(s) for s in a["src_list"]]
    J, g = fwi.misfit_and_gradient(a["m0"], ag, a["wavelet"], float(a["dt"]), float(a["h"]),
                                   int(a["nt"]), a["nosponge_dobs"], None, asrc, "envelope",
                                   footprints=fps)
    out["fp_envelope_J"], out["fp_envelope_g"] = np.array(J), g
    np.savez_compressed(os.path.join(HERE, "misfits_2d.npz"), tw=np.stack(tw), sponge=sp, **out)
    print("misfit fixtures written to", HERE)


def main_builders(ref_root):
    """The other input builders: Gabor and Hann tone-burst wavelets
    (sources.py:29-62), Gaussian smoothing, the sphere background and the 2D /
    3D Voronoi polycrystals from seeded generators (phantom.py:51-188), shell,
    hemisphere and sphere arrays (geometry.py:140-196) and footprints on them."""
    sys.path.insert(0, os.path.join(ref_root, "simulation"))
    from ringfwi import phantom  # noqa: E402
    from ringfwi.geometry import HemisphereArray, ShellArray, SphereArray, build_footprints  # noqa: E402
    from ringfwi.sources import gabor, hann_tone_burst  # noqa: E402

    out = {}
    out["gabor0"] = gabor(400, 5e-8, 5e5)
    out["gabor1"] = gabor(333, 2.5e-8, 1.2e6, frac_bw=0.9, t0=2e-6)
    out["hann0"] = hann_tone_burst(300, 1e-7, 4e5)
    out["hann1"] = hann_tone_burst(500, 5e-8, 7.5e5, n_cycles=3, t0=1.3e-6)
    c = phantom.coupling_background((41, 41), 3000.0, 1480.0, 0.015, 1e-3)
    out["smooth_in"] = c
    out["smooth_out"] = phantom.smooth(c, 2.0)
    out["sphere_bg"] = phantom.sphere_background((17, 19, 21), 3000.0, 1480.0, 0.006, 1e-3)
    lab, ang, th = phantom.voronoi_polycrystal((61, 61), 12, 0.02, 1e-3, rng=np.random.default_rng(7),
                                                relax=2)
    out.update(vor_labels=lab, vor_angles=ang, vor_theta=th)
    lab, axes, col = phantom.voronoi_polycrystal_3d((15, 21, 21), 9, 0.008, 1e-3,
                                                     rng=np.random.default_rng(8), relax=1)
    out.update(vor3_labels=lab, vor3_axes=axes, vor3_colat=col)
    arrays = [SphereArray(40, 0.012, 0.03, 1e-3), HemisphereArray(33, 0.011, 0.03, 7.5e-4),
              ShellArray(64, 0.02, 0.0512, 2e-4)]
    for i, a in enumerate(arrays):
        out[f"shell{i}_idx"] = np.array(a.idx, dtype=np.int64)
        out[f"shell{i}_pos"] = a.element_positions
        out[f"shell{i}_n"] = np.array(a.n)
        out[f"shell{i}_type"] = np.array(a.array_type)
    fps = build_footprints(arrays[0], 0.004, "disc", height_m=0.003, max_pts=7)
    out["shell_fp_cells"] = np.array([ix for idxs, _ in fps for ix in idxs], dtype=np.int64)
    out["shell_fp_counts"] = np.array([len(idxs) for idxs, _ in fps], dtype=np.int64)
    out["shell_fp_w"] = np.concatenate([np.asarray(w, dtype=np.float64) for _, w in fps])
    np.savez_compressed(os.path.join(HERE, "builders.npz"), **out)
    print("builder fixtures written to", HERE)


def main_invert(ref_root):
    """Trajectories of the reference's invert (fwi.py:535-667) on the problems
    of fmc_2d.npz and fmc_3d.npz: iterates and misfit histories."""
    sys.path.insert(0, os.path.join(ref_root, "simulation"))
    from ringfwi import fwi, phantom  # noqa: E402
    from ringfwi.solver import make_sponge  # noqa: E402

    class Geom:  # element indices of a fixture, as the reference reads a geometry
        def __init__(self, idx):
            self.idx = [tuple(int(v) for v in r) for r in idx]
            self.n_elements = len(self.idx)

        def element_index(self, k):
            return self.idx[k]

    out = {}
    f = np.load(os.path.join(HERE, "fmc_2d.npz"))
    geom = Geom(f["idx"])
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    m0, src = f["m0"], [int(s) for s in f["src_list"]]
    n = m0.shape[0]
    yy, xx = np.mgrid[0:n, 0:n]
    mask = (np.hypot(xx - (n - 1) / 2.0, yy - (n - 1) / 2.0) <= 0.4 * (n - 1)).astype(float)
    bounds = (phantom.velocity_to_m(3600.0), phantom.velocity_to_m(2500.0))
    sp = make_sponge(m0.shape, width=8)
    cases = {
        "gd": dict(dobs=f["nosponge_dobs"], sponge=None, kw=dict(n_iter=3, step_frac=0.02)),
        "gd_mask": dict(dobs=f["nosponge_dobs"], sponge=None,
                        kw=dict(n_iter=3, step_frac=0.03, update_mask=mask, m_bounds=bounds)),
        "gd_gcn": dict(dobs=f["nosponge_dobs"], sponge=None,
                       kw=dict(n_iter=3, step_frac=0.02, misfit_type="gcn")),
        "gn": dict(dobs=f["nosponge_dobs"], sponge=None,
                   kw=dict(n_iter=2, optimizer="gn", n_cg=4, update_mask=mask, m_bounds=bounds)),
    }
    for name, c in cases.items():
        m, hist = fwi.invert(m0, geom, w, dt, h, nt, c["dobs"], c["sponge"], src, **c["kw"])
        out[f"{name}_m"], out[f"{name}_history"] = m, np.array(hist)
        print(name, len(hist), hist[-1] / hist[0])
    out["mask"] = mask
    # the sponge as a border layer (sponge_2d.npz): the reference's inexact
    # sponge gradient still descends there
    s2 = np.load(os.path.join(HERE, "sponge_2d.npz"))
    m, hist = fwi.invert(s2["m0"], Geom(s2["idx"]), s2["wavelet"], float(s2["dt"]), float(s2["h"]),
                         int(s2["nt"]), s2["dobs"], s2["sponge"], [int(s) for s in s2["src_list"]],
                         n_iter=3, step_frac=0.02)
    out["gd_sponge_m"], out["gd_sponge_history"] = m, np.array(hist)
    print("gd_sponge", len(hist), hist[-1] / hist[0])
    g3 = np.load(os.path.join(HERE, "fmc_3d.npz"))
    m, hist = fwi.invert(g3["m0"], Geom(g3["idx"]), g3["wavelet"], float(g3["dt"]), float(g3["h"]),
                         int(g3["nt"]), g3["dobs"], None, [int(s) for s in g3["src_list"]],
                         n_iter=3, step_frac=0.005)
    out["gd3d_m"], out["gd3d_history"] = m, np.array(hist)
    print("gd3d", len(hist), hist[-1] / hist[0])
    np.savez_compressed(os.path.join(HERE, "invert.npz"), **out)
    print("invert fixtures written to", HERE)


SECTIONS = {"base": main, "apertures": main_apertures, "misfits": main_misfits,
            "builders": main_builders, "invert": main_invert}

if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--only=")]
    only = [a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--only=")]
    names = only[0].split(",") if only else list(SECTIONS)
    for name in names:
        SECTIONS[name](args[0] if args else "/root/reference")
