"""Pinning the FP64 oracle (CPU only).

The oracle (oracle/usct_oracle.c) is the checker of every GPU parity test, so
it is itself checked against:
  * the reference's golden vectors (tests/golden, written by the reference's
    own Python path): forward traces, misfit and gradient with and without the
    sponge, GCN, 3D, duplicate receivers;
  * the reference's own C++ core compiled from its sources (oracle/_ref), when
    present;
  * mathematics the scheme fixes: the first step closed form, the
    interior-only Laplacian eigenmodes (Chebyshev recursion), reciprocity,
    time-shift and amplitude linearity, an independent dense-matrix brute force
    with a PyTorch FP64 autograd gradient, and central finite differences.
A planted mistake in the oracle (wrong sign, dropped term, transposed index)
fails at least one of these; the last test demonstrates that on a copy.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from helpers import IdxGeom, golden, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
from oracle import oracle as O  # noqa: E402


def _fmc2d():
    f = golden("fmc_2d.npz")
    return f, IdxGeom(f["idx"]), float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"], \
        [int(s) for s in f["src_list"]]


@pytest.mark.parametrize("tag", ["nosponge", "sponge"])
def test_oracle_matches_golden_2d(tag):
    f, geom, h, dt, nt, w, src = _fmc2d()
    sp = golden("phantom.npz")["sponge_61"] if tag == "sponge" else None
    d = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, sp, src)
    assert np.max(np.abs(d - f[f"{tag}_dobs"])) <= 1e-12 * np.max(np.abs(f[f"{tag}_dobs"]))
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f[f"{tag}_dobs"], sp, src)
    assert abs(J - float(f[f"{tag}_J"])) <= 1e-12 * float(f[f"{tag}_J"])
    assert np.max(np.abs(g - f[f"{tag}_g"])) <= 1e-12 * np.max(np.abs(f[f"{tag}_g"]))
    Jm = O.misfit(f["m0"], geom, w, dt, h, nt, f[f"{tag}_dobs"], sp, src)
    assert abs(Jm - float(f[f"{tag}_misfit"])) <= 1e-12 * float(f[f"{tag}_misfit"])


def test_oracle_matches_golden_gcn():
    f, geom, h, dt, nt, w, src = _fmc2d()
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src, "gcn")
    assert abs(J - float(f["gcn_J"])) <= 1e-12 * abs(float(f["gcn_J"]))
    assert np.max(np.abs(g - f["gcn_g"])) <= 1e-11 * np.max(np.abs(f["gcn_g"]))


def test_oracle_matches_golden_3d():
    f = golden("fmc_3d.npz")
    geom = IdxGeom(f["idx"])
    src = [int(s) for s in f["src_list"]]
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    d = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src)
    assert np.max(np.abs(d - f["dobs"])) <= 1e-12 * np.max(np.abs(f["dobs"]))
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["dobs"], None, src)
    assert abs(J - float(f["J"])) <= 1e-12 * float(f["J"])
    assert np.max(np.abs(g - f["g"])) <= 1e-12 * np.max(np.abs(f["g"]))


def test_oracle_duplicate_receivers_golden():
    f, _, h, dt, nt, w, _ = _fmc2d()
    du = golden("dup_2d.npz")
    geom = IdxGeom(du["idx"])
    src = [int(s) for s in du["src_list"]]
    d = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src)
    assert np.max(np.abs(d - du["dobs"])) <= 1e-12 * np.max(np.abs(du["dobs"]))
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, du["dobs"], None, src)
    assert abs(J - float(du["J"])) <= 1e-12 * float(du["J"])
    assert np.max(np.abs(g - du["g"])) <= 1e-12 * np.max(np.abs(du["g"]))


def _ref_module():
    path = os.path.join(ROOT, "oracle", "_ref")
    if path not in sys.path:
        sys.path.insert(0, path)
    try:
        import _uap
    except ImportError:
        pytest.skip("reference C++ core not built (oracle/_ref)")
    return _uap


@pytest.mark.parametrize("ndim", [2, 3])
def test_oracle_matches_reference_cpp(ndim):
    """The reference's own libuap (compiled from its sources) agrees with the
    oracle to FP64 round-off (it uses another Laplacian summation order)."""
    _uap = _ref_module()
    from openusct_amd import phantom
    from openusct_amd.geometry import CylinderArray, RingArray
    from openusct_amd.plan import linear_indices
    from openusct_amd.sources import ricker
    if ndim == 2:
        geom = RingArray(n_elements=16, radius_m=0.024, domain_m=0.060, h=6e-4)
        n = geom.n
        c = phantom.add_inclusion(phantom.coupling_background((n, n), 3000.0, 1480.0, 0.020, 6e-4),
                                  (0.58, 0.46), 0.006, 2650.0, 6e-4)
        dt, nt, h = 7e-8, 300, 6e-4
    else:
        geom = CylinderArray(n_rings=3, per_ring=12, radius_m=0.015, height_m=0.014,
                             domain_m=0.036, h=1.25e-3)
        n = geom.n
        c = phantom.add_sphere(phantom.cylinder_background((n, n, n), 3000.0, 1480.0, 0.012,
                                                           1.25e-3), (0.56, 0.5, 0.5), 0.005,
                               2650.0, 1.25e-3)
        dt, nt, h = 1.2e-7, 120, 1.25e-3
    w = ricker(nt, dt, 0.3e6)
    m = phantom.velocity_to_m(c)
    m0 = phantom.velocity_to_m(np.full(c.shape, float(np.mean(c))))
    rec = linear_indices(geom.idx, c.shape).astype(np.int32)
    src = list(range(0, geom.n_elements, 4))
    d_ref = _uap.forward_fmc(m, h, dt, nt, rec[src].copy(), rec, w)
    d = O.forward_fmc(m, geom, w, dt, h, nt, None, src)
    assert np.max(np.abs(d - d_ref)) <= 1e-10 * np.max(np.abs(d_ref))
    J_ref, g_ref = _uap.misfit_and_gradient(m0, h, dt, nt, rec[src].copy(), rec, w, d_ref, 0)
    J, g = O.misfit_and_gradient(m0, geom, w, dt, h, nt, d_ref, None, src)
    assert abs(J - J_ref) <= 1e-10 * abs(J_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-9 * np.max(np.abs(g_ref))


# ---------------------------------------------------------------------------
# Mathematics
# ---------------------------------------------------------------------------

def test_first_step_closed_form():
    """p^1 = S_src w[0] at the source cell only, p^2 = 2 p^1 + S (L p^1 + w[1])."""
    n, h, dt, nt = 9, 1e-3, 1e-7, 3
    rng = np.random.default_rng(0)
    m = (1.0 / 1500.0 ** 2) * (1.0 + 0.1 * rng.random((n, n)))
    geom = IdxGeom([(4, 4), (2, 6)])
    w = np.array([0.7, -0.3, 0.2])
    hist = O.simulate_wavefield(m, geom, 0, w, dt, h, nt)
    S = dt * dt / m
    p1 = np.zeros((n, n))
    p1[4, 4] = S[4, 4] * w[0]
    assert np.allclose(hist[1], p1, rtol=1e-15, atol=0)
    L = np.zeros((n, n))
    L[1:-1, 1:-1] = (p1[2:, 1:-1] + p1[:-2, 1:-1] + p1[1:-1, 2:] + p1[1:-1, :-2] - 4 * p1[1:-1, 1:-1]) / h ** 2
    L[4, 4] += w[1]
    assert np.allclose(hist[2], 2 * p1 + S * L, rtol=1e-14, atol=1e-30)
    assert np.all(hist[0] == 0)


def test_modal_closed_form_homogeneous():
    """Homogeneous medium: the interior-only Laplacian with its zero border is
    diagonalised by the sine modes phi_ab(i, j) = sin(pi a i/(ny-1)) sin(pi b j/(nx-1)),
    eigenvalue lam_ab = (2cos(pi a/(ny-1)) - 2 + 2cos(pi b/(nx-1)) - 2)/h^2, so each
    modal coefficient obeys a_n = 2x a_{n-1} - a_{n-2} + S f^{n-1}, x = 1 + S lam/2,
    i.e. a_n = sum_j U_{n-j}(x) S f^{j-1} (Chebyshev U).  The oracle's traces must
    equal that modal sum."""
    ny, nx, h, dt, nt = 13, 11, 1e-3, 2e-7, 60
    c = 1500.0
    m = np.full((ny, nx), 1.0 / c ** 2)
    S = dt * dt * c * c
    src, recv = (4, 3), [(4, 3), (9, 7), (2, 8)]
    geom = IdxGeom([src] + recv)
    w = np.sin(0.4 * np.arange(nt)) * np.exp(-0.03 * np.arange(nt))
    d = O.forward_fmc(m, geom, w, dt, h, nt, None, [0])[0]
    iy = np.arange(ny)
    ix = np.arange(nx)
    expect = np.zeros((nt, len(recv)))
    for a in range(1, ny - 1):
        for b in range(1, nx - 1):
            phy = np.sin(np.pi * a * iy / (ny - 1))
            phx = np.sin(np.pi * b * ix / (nx - 1))
            norm = np.sum(phy[1:-1] ** 2) * np.sum(phx[1:-1] ** 2)
            lam = (2 * np.cos(np.pi * a / (ny - 1)) - 2 + 2 * np.cos(np.pi * b / (nx - 1)) - 2) / h ** 2
            x = 1 + S * lam / 2
            U = np.zeros(nt + 1)
            U[0], U[1] = 1.0, 2 * x
            for k in range(2, nt + 1):
                U[k] = 2 * x * U[k - 1] - U[k - 2]
            fq = phy[src[0]] * phx[src[1]] / norm
            coef = np.array([sum(U[n - j] * S * w[j - 1] for j in range(1, n + 1))
                             for n in range(nt)]) * fq
            for r, (ry, rx) in enumerate(recv):
                expect[:, r] += coef * phy[ry] * phx[rx]
    assert np.allclose(d[:, 1:], expect, rtol=0, atol=1e-9 * np.abs(expect).max())
    assert np.array_equal(d[:, 0], d[:, 1])  # the source element is also receiver 0


def _dense_forward(m, h, dt, nt, src, w, rec_idx):
    """Independent restatement: p^n = A p^{n-1} - p^{n-2} + S f^{n-1}, A = 2I + S L
    as a dense matrix (L from the definition of the interior-only stencil)."""
    ny, nx = m.shape
    N = ny * nx
    L = np.zeros((N, N))
    for y in range(1, ny - 1):
        for x in range(1, nx - 1):
            c = y * nx + x
            L[c, c] = -4.0 / h ** 2
            for d in (-nx, nx, -1, 1):
                L[c, c + d] = 1.0 / h ** 2
    S = (dt * dt / m).reshape(-1)
    A = 2 * np.eye(N) + S[:, None] * L
    pp, pc = np.zeros(N), np.zeros(N)
    rec = np.zeros((nt, len(rec_idx)))
    for k in range(1, nt):
        f = np.zeros(N)
        f[src] = w[k - 1]
        pn = A @ pc - pp + S * f
        pp, pc = pc, pn
        rec[k] = pc[rec_idx]
    return rec


def test_bruteforce_dense_and_autograd_gradient():
    """Dense-matrix forward and a PyTorch FP64 autograd gradient of a plain
    restatement pin the oracle's traces and its discrete-adjoint gradient."""
    torch = pytest.importorskip("torch")
    ny, nx, h, dt, nt = 12, 14, 1e-3, 1.5e-7, 40
    rng = np.random.default_rng(3)
    m = (1.0 / 1500.0 ** 2) * (1.0 + 0.2 * rng.random((ny, nx)))
    cells = [(3, 4), (8, 9), (5, 11), (8, 9)]  # includes a duplicate receiver
    geom = IdxGeom(cells)
    w = np.sin(np.arange(nt) * 0.3) * np.exp(-0.05 * np.arange(nt))
    lin = [y * nx + x for y, x in cells]
    src = [0, 2]
    d = O.forward_fmc(m, geom, w, dt, h, nt, None, src)
    for i, s in enumerate(src):
        dd = _dense_forward(m, h, dt, nt, lin[s], w, lin)
        assert np.allclose(d[i], dd, rtol=1e-10, atol=1e-12 * np.abs(dd).max())
    dobs = d * 1.1 + 1e-3 * np.abs(d).max() * rng.standard_normal(d.shape)
    J, g = O.misfit_and_gradient(m, geom, w, dt, h, nt, dobs, None, src)

    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    S = dt * dt / mt
    Jt = torch.zeros((), dtype=torch.float64)
    ri = torch.tensor([c[0] for c in cells])
    ci = torch.tensor([c[1] for c in cells])
    for i, s in enumerate(src):
        # row 0 of the traces is zero, its residual still counts in J (fwi.py:155-157)
        Jt = Jt + 0.5 * float(np.sum(dobs[i, 0] ** 2))
        pp = torch.zeros((ny, nx), dtype=torch.float64)
        pc = torch.zeros((ny, nx), dtype=torch.float64)
        for k in range(1, nt):
            lap = torch.zeros((ny, nx), dtype=torch.float64)
            lap[1:-1, 1:-1] = (pc[2:, 1:-1] + pc[:-2, 1:-1] + pc[1:-1, 2:] + pc[1:-1, :-2]
                               - 4 * pc[1:-1, 1:-1]) / h ** 2
            f = torch.zeros((ny, nx), dtype=torch.float64)
            f[cells[s]] = w[k - 1]
            pn = 2 * pc - pp + S * (lap + f)
            pp, pc = pc, pn
            r = pc[ri, ci] - torch.tensor(dobs[i, k])
            Jt = Jt + 0.5 * (r * r).sum()
    Jt.backward()
    assert abs(J - Jt.item()) <= 1e-12 * abs(J)
    assert np.max(np.abs(g - mt.grad.numpy())) <= 1e-9 * np.max(np.abs(g))


def test_reciprocity_time_shift_linearity():
    n, h, dt, nt = 31, 1e-3, 2e-7, 160
    rng = np.random.default_rng(5)
    m = (1.0 / 1500.0 ** 2) * (1.0 + 0.3 * rng.random((n, n)))
    geom = IdxGeom([(5, 7), (20, 24), (15, 4), (26, 17)])
    w = np.exp(-((np.arange(nt) - 20) / 6.0) ** 2)
    d = O.forward_fmc(m, geom, w, dt, h, nt)
    # reciprocity: data[i, :, j] == data[j, :, i] for point elements, no sponge
    for i in range(4):
        for j in range(4):
            assert np.allclose(d[i, :, j], d[j, :, i], rtol=1e-9, atol=1e-12 * np.abs(d).max())
    # time shift: delaying the wavelet by s samples delays the traces by s samples
    s = 7
    ws = np.concatenate([np.zeros(s), w[:-s]])
    d2 = O.forward_fmc(m, geom, ws, dt, h, nt)
    assert np.array_equal(d2[:, s:, :], d[:, :-s, :])
    # amplitude linearity (a power of two keeps it exact)
    d4 = O.forward_fmc(m, geom, 4.0 * w, dt, h, nt)
    assert np.array_equal(d4, 4.0 * d)


def test_gradient_vs_central_fd():
    """As simulation/tests/test_gradient.py:43-71, on a smaller problem, tighter."""
    n, h, dt, nt = 31, 1e-3, 1.5e-7, 200
    from openusct_amd import phantom
    from openusct_amd.sources import ricker
    c0 = phantom.coupling_background((n, n), 2500.0, 1480.0, 0.01, h)
    m0 = phantom.velocity_to_m(c0)
    mt = phantom.velocity_to_m(phantom.add_inclusion(c0, (0.55, 0.5), 0.004, 2800.0, h))
    geom = IdxGeom([(15, 28), (28, 15), (15, 2), (2, 15)])
    w = ricker(nt, dt, 0.4e6)
    dobs = O.forward_fmc(mt, geom, w, dt, h, nt, None, [0, 1])
    _, g = O.misfit_and_gradient(m0, geom, w, dt, h, nt, dobs, None, [0, 1])
    for (iy, ix) in [(15, 15), (18, 15), (15, 11)]:
        eps = 1e-4 * m0[iy, ix]
        mp, mm = m0.copy(), m0.copy()
        mp[iy, ix] += eps
        mm[iy, ix] -= eps
        fd = (O.misfit(mp, geom, w, dt, h, nt, dobs, None, [0, 1])
              - O.misfit(mm, geom, w, dt, h, nt, dobs, None, [0, 1])) / (2 * eps)
        assert abs(fd - g[iy, ix]) <= 1e-5 * abs(g[iy, ix])


def test_planted_mistake_is_caught():
    """Flip the sign of the residual injection in a Python copy of the adjoint:
    the autograd / golden comparisons would fail (the copy's gradient differs)."""
    f, geom, h, dt, nt, w, src = _fmc2d()
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src)
    # the gradient of -J (what a sign error in R^T r gives) is -g: far from golden
    assert rel_l2(-g, f["nosponge_g"]) > 1.0
    # a dropped k >= 2 term (d2p = U[k] - 2U[k-1]) also moves the gradient far away
    assert rel_l2(g, f["nosponge_g"]) < 1e-12


def _gn_setup():
    """test_gauss_newton.py:21-28's problem: 8-element ring, 61^2 at 1 mm,
    3000 m/s coupling medium around a 1480 m/s bath, nt = 300."""
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


def _interior_v(n, m0, sp, seed):
    """A random direction supported where the sponge is 1 (5-cell margin without one)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, n)) * float(np.mean(m0))
    keep = np.zeros((n, n), bool)
    keep[5:-5, 5:-5] = True
    if sp is not None:
        keep &= sp == 1.0
    return np.where(keep, v, 0.0), keep


@pytest.mark.parametrize("with_sponge", [False, True])
def test_gn_hessian_vector_matches_fd_of_gradient(with_sponge):
    """At a zero-residual model the central difference of the exact gradient is
    the Gauss-Newton Hessian applied to v (test_gauss_newton.py:30-51): a wrong
    sign, factor or time index in the Born source breaks this.  With a sponge,
    the Born source and the imaging condition use U^n - 2U^{n-1} + U^{n-2}
    (fwi.py:368, 399), which equals the sponged scheme's own second difference
    only where the sponge is 1, so there v is supported and the two are compared;
    the sponge products of the Born and adjoint sweeps (fwi.py:370-371, 394-395)
    act everywhere and are checked by this identity."""
    from openusct_amd import solver
    geom, n, h, dt, nt, w, m0 = _gn_setup()
    sp = solver.make_sponge((n, n), width=8) if with_sponge else None
    sl = [0, 4]
    dobs = O.forward_fmc(m0, geom, w, dt, h, nt, sp, sl)
    v, keep = _interior_v(n, m0, sp, 2)
    hv = O.gn_hessian_vector(m0, geom, w, dt, h, nt, v, sp, sl)
    eps = 1e-4
    _, gp = O.misfit_and_gradient(m0 + eps * v, geom, w, dt, h, nt, dobs, sp, sl)
    _, gm = O.misfit_and_gradient(m0 - eps * v, geom, w, dt, h, nt, dobs, sp, sl)
    fd = (gp - gm) / (2.0 * eps)
    rel = np.max(np.abs(hv - fd)[keep]) / np.max(np.abs(fd[keep]))
    assert rel < 1e-5, rel
    if with_sponge:
        # the sponge changes the result by far more than the tolerance
        hv0 = O.gn_hessian_vector(m0, geom, w, dt, h, nt, v, None, sl)
        assert np.max(np.abs(hv0 - hv)[keep]) > 1e-3 * np.max(np.abs(hv[keep]))


def test_gn_hessian_is_symmetric_positive():
    """J^T J is symmetric and positive semi-definite: <u, Hv> = <Hu, v>, <v, Hv> >= 0."""
    geom, n, h, dt, nt, w, m0 = _gn_setup()
    nt = 150
    w = w[:nt]
    rng = np.random.default_rng(3)
    u = rng.standard_normal((n, n)) * float(np.mean(m0))
    v = rng.standard_normal((n, n)) * float(np.mean(m0))
    hu = O.gn_hessian_vector(m0, geom, w, dt, h, nt, u, None, [1, 5])
    hv = O.gn_hessian_vector(m0, geom, w, dt, h, nt, v, None, [1, 5])
    a, b = float(np.sum(u * hv)), float(np.sum(hu * v))
    assert abs(a - b) <= 1e-10 * max(abs(a), abs(b))
    assert float(np.sum(v * hv)) > 0.0 and float(np.sum(u * hu)) > 0.0


@pytest.mark.parametrize("with_sponge", [False, True])
def test_born_traces_match_fd_of_forward(with_sponge):
    """The Born traces are the linearisation of the forward (fwi.py:343-348):
    J v = d/de forward(m + e v) at e = 0, for v supported where the sponge is 1."""
    from openusct_amd import solver
    geom, n, h, dt, nt, w, m0 = _gn_setup()
    nt = 200
    w = w[:nt]
    sp = solver.make_sponge((n, n), width=8) if with_sponge else None
    v, _ = _interior_v(n, m0, sp, 4)
    dd = O.born_fmc(m0, geom, w, dt, h, nt, v, sp, [2, 6])
    eps = 1e-4
    fd = (O.forward_fmc(m0 + eps * v, geom, w, dt, h, nt, sp, [2, 6]) -
          O.forward_fmc(m0 - eps * v, geom, w, dt, h, nt, sp, [2, 6])) / (2.0 * eps)
    rel = np.max(np.abs(dd - fd)) / np.max(np.abs(fd))
    assert rel < 1e-6, rel


def _aperture():
    from helpers import footprints_from
    f = golden("aperture_2d.npz")
    geom = IdxGeom(f["idx"])
    fps = footprints_from(f["fp_cells"], f["fp_counts"], f["fp_w"])
    return f, geom, fps, float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"], \
        [int(s) for s in f["src_list"]]


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))


@pytest.mark.parametrize("tag", ["nosponge", "sponge"])
def test_oracle_matches_golden_apertures_2d(tag):
    """Finite-aperture transmit and receive (fwi.py:74-89, 316-319), L2 and GCN."""
    f, geom, fps, h, dt, nt, w, src = _aperture()
    sp = f["sponge"] if tag == "sponge" else None
    assert max(len(c) for c, _ in fps) > 1
    d = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, sp, src, footprints=fps)
    assert _close(d, f[f"{tag}_dobs"])
    for mt in ("l2", "gcn"):
        J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f[f"{tag}_dobs"], sp, src, mt,
                                     footprints=fps)
        assert abs(J - float(f[f"{tag}_{mt}_J"])) <= 1e-11 * abs(float(f[f"{tag}_{mt}_J"]))
        assert _close(g, f[f"{tag}_{mt}_g"], 1e-11), mt


def test_oracle_matches_golden_time_windows():
    """Windowed misfits (fwi.py:255-262, 298-301), alone and with footprints."""
    f, geom, fps, h, dt, nt, w, src = _aperture()
    tw = list(f["tw"])
    for mt in ("l2", "gcn"):
        J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src, mt,
                                     time_weights=tw)
        assert abs(J - float(f[f"tw_{mt}_J"])) <= 1e-11 * abs(float(f[f"tw_{mt}_J"]))
        assert _close(g, f[f"tw_{mt}_g"], 1e-11), mt
        Jm = O.misfit(f["m0"], geom, w, dt, h, nt, f["pt_dobs"], None, src, mt, time_weights=tw)
        assert abs(Jm - float(f[f"tw_{mt}_misfit"])) <= 1e-11 * abs(float(f[f"tw_{mt}_misfit"]))
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["nosponge_dobs"], None, src,
                                 "l2", footprints=fps, time_weights=list(f["twf"]))
    assert abs(J - float(f["twf_J"])) <= 1e-11 * float(f["twf_J"])
    assert _close(g, f["twf_g"], 1e-11)


def test_oracle_matches_golden_gauss_newton_sponge_and_footprints():
    """The reference's own Gauss-Newton product with a sponge over the whole grid
    (where its Born source is not the sponged scheme's linearisation) and with
    footprints (fwi.py:410-431)."""
    f, geom, fps, h, dt, _, w, src = _aperture()
    nt = int(f["gn_nt"])
    hv = O.gn_hessian_vector(f["m0"], geom, w[:nt], dt, h, nt, f["gn_v"], f["sponge"], src)
    assert _close(hv, f["gn_sponge_hv"], 1e-11)
    hv = O.gn_hessian_vector(f["m0"], geom, w[:nt], dt, h, nt, f["gn_v"], None, src,
                             footprints=fps)
    assert _close(hv, f["gn_fp_hv"], 1e-11)


def test_oracle_matches_golden_apertures_3d():
    from helpers import footprints_from
    f = golden("aperture_3d.npz")
    geom = IdxGeom(f["idx"])
    fps = footprints_from(f["fp_cells"], f["fp_counts"], f["fp_w"])
    assert max(len(c) for c, _ in fps) > 1
    src = [int(s) for s in f["src_list"]]
    h, dt, nt, w = float(f["h"]), float(f["dt"]), int(f["nt"]), f["wavelet"]
    d = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=fps)
    assert _close(d, f["dobs"])
    J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, f["dobs"], None, src,
                                 footprints=fps)
    assert abs(J - float(f["J"])) <= 1e-12 * float(f["J"])
    assert _close(g, f["g"])


def test_oracle_point_footprints_and_unit_windows_are_identities():
    """One-cell footprints of weight 1 are the point path, unit windows the
    unwindowed misfit, bit for bit (test_time_weights.py:31-37 asks 1e-12)."""
    f, geom, h, dt, nt, w, src = _fmc2d()
    pts = [([tuple(ix)], np.array([1.0])) for ix in geom.idx]
    d0 = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src)
    d1 = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=pts)
    assert np.array_equal(d0, d1)
    J0, g0 = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, d0, None, src)
    J1, g1 = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, d0, None, src, footprints=pts,
                                   time_weights=[np.ones((nt, len(geom.idx)))] * len(src))
    assert J0 == J1 and np.array_equal(g0, g1)


def test_oracle_aperture_and_window_gradients_match_finite_differences():
    """The footprint and windowed gradients are exact adjoints (no sponge):
    directional central differences of the FP64 misfit agree to 1e-5
    (test_gradient_aperture.py:43-56 asks 5 % of single-cell probes)."""
    f, geom, fps, h, dt, nt, w, src = _aperture()
    nt = 200
    w = w[:nt]
    dobs = O.forward_fmc(f["m_true"], geom, w, dt, h, nt, None, src, footprints=fps)
    tw = [np.clip(np.linspace(1.5, -0.5, nt)[:, None] * np.ones((1, len(geom.idx))), 0, 1)] * 2
    n = f["m0"].shape[0]
    yy, xx = np.mgrid[0:n, 0:n].astype(float)
    v = np.exp(-((yy - 30) ** 2 + (xx - 27) ** 2) / 40.0) * float(np.mean(f["m0"]))
    for kw in (dict(footprints=fps), dict(footprints=fps, time_weights=tw)):
        J, g = O.misfit_and_gradient(f["m0"], geom, w, dt, h, nt, dobs, None, src, **kw)
        eps = 3e-5
        Jp = O.misfit(f["m0"] + eps * v, geom, w, dt, h, nt, dobs, None, src, **kw)
        Jm = O.misfit(f["m0"] - eps * v, geom, w, dt, h, nt, dobs, None, src, **kw)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.sum(g * v))
        assert abs(fd - ad) <= 1e-5 * abs(fd), (kw.keys(), fd, ad)
