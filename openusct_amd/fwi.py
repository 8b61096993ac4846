"""Drop-in solver module with the reference's function signatures.

``import openusct_amd.fwi as fwi`` replaces ``from ringfwi import fwi`` (and the
C++ ``uap`` module, software/python/uap.py:1-67) for the acoustic path; the
module itself is also a ``backend=`` for the reference's ``invert``
(simulation/ringfwi/fwi.py:594-603), which calls
``backend.misfit_and_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge,
src_list, misfit_type=...)`` and ``backend.misfit(...)`` positionally.

Inputs are NumPy arrays (float64 models, as in the reference) or torch
tensors; results come back as NumPy float64 like ``ringfwi.fwi``.  Every
call runs on the HIP device through libusct; there is no CPU path.

Semantics (fwi.py:34-119, 246-332): the source sample used to advance to level
n is ``wavelet[n-1]``; receivers record ``p^n`` for n >= 1 with row 0 = 0;
``src_list`` indexes elements and every element receives; the sponge is
applied (unlike the reference's C++ and CuPy backends, which ignore it,
uap.py:32-33); duplicate receivers accumulate in the adjoint; ``g`` is the
gradient with respect to m (squared slowness).
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch

from . import misfits as trace_misfits
from .misfits import DEVICE_MISFITS
from .plan import Plan, linear_indices

# Backend selection for the drop-in functions: "batched" (default) or "plain".
METHOD = "batched"
_CACHE_SIZE = 2
_plans: "OrderedDict[tuple, Plan]" = OrderedDict()


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("openusct_amd needs a HIP device; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def check_cfl(m, dt, h):
    """Reject time steps beyond the leapfrog stability limit
    c_max dt / h <= 1/sqrt(ndim) (simulation/docs/theory.md:24)."""
    m = m.detach().cpu().numpy() if isinstance(m, torch.Tensor) else np.asarray(m)
    if m.size == 0 or not np.all(np.isfinite(m)) or np.min(m) <= 0:
        raise ValueError("m must be finite and positive (squared slowness)")
    c_max = float(1.0 / np.sqrt(np.min(m)))
    lim = 1.0 / np.sqrt(m.ndim)
    if c_max * dt / h > lim * (1.0 + 1e-12):
        raise ValueError(f"CFL violated: c_max*dt/h = {c_max * dt / h:.4f} > {lim:.4f}")


def _tx_rec(geom, shape, src_list):
    if not hasattr(geom, "idx"):
        raise TypeError("geom must provide .idx, .n_elements and .element_index(k)")
    rec = linear_indices(geom.idx, shape)
    src = list(range(geom.n_elements)) if src_list is None else list(src_list)
    for s in src:
        if not 0 <= int(s) < geom.n_elements:
            raise ValueError(f"src_list entry {s} is not an element of the array")
    tx = np.array([linear_indices([geom.element_index(int(s))], shape)[0] for s in src],
                  dtype=np.int64)
    return tx, rec


def _groups_key(groups):
    if groups is None:
        return None
    return tuple((np.asarray(c, np.int64).tobytes(), np.asarray(w, np.float64).tobytes())
                 for c, w in groups)


def get_plan(shape, h, dt, nt, tx, rec, sponge, method=None, groups=None):
    """Plan for this problem, reused across calls (line searches, iterations).
    ``groups`` = (transmit groups, receiver groups) for finite-aperture elements."""
    method = method or METHOD
    gk = None if groups is None else (_groups_key(groups[0]), _groups_key(groups[1]))
    key = (tuple(shape), float(h), float(dt), int(nt), tx.tobytes(), rec.tobytes(), bool(sponge),
           method, gk, torch.cuda.current_device())
    plan = _plans.get(key)
    if plan is not None:
        _plans.move_to_end(key)
        return plan
    while len(_plans) >= _CACHE_SIZE:
        _plans.popitem(last=False)[1].close()
    if groups is None:
        plan = Plan(shape, h, dt, nt, tx, rec, method=method, sponge=sponge)
    else:
        plan = Plan(shape, h, dt, nt, tx, rec, method=method, sponge=sponge,
                    tx_groups=groups[0], rec_groups=groups[1])
    _plans[key] = plan
    return plan


def clear_plans():
    while _plans:
        _plans.popitem()[1].close()


def _to_dev(a, dtype):
    if isinstance(a, torch.Tensor):
        return a.to(device=_device(), dtype=dtype).contiguous()
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:  # read-only arrays (np.load, broadcasts): torch wants its own copy
        a = a.copy()
    return torch.as_tensor(a, dtype=dtype, device=_device())


def _shape_of(m):
    return tuple(m.shape)


def _element_groups(geom, shape, src_list, footprints):
    """Transmit and receiver groups (linear cells, weights) from footprints
    (geometry.py:212-278 format: per element a list of index tuples and weights)."""
    if len(footprints) != geom.n_elements:
        raise ValueError(f"footprints has {len(footprints)} entries, the array "
                         f"{geom.n_elements} elements")
    rec = []
    for idxs, w in footprints:
        w = np.asarray(w, dtype=np.float64).reshape(-1)
        if len(idxs) != len(w) or len(w) == 0:
            raise ValueError("each footprint needs one weight per grid point")
        rec.append((linear_indices(list(idxs), shape), w))
    src = list(range(geom.n_elements)) if src_list is None else [int(s) for s in src_list]
    for s in src:
        if not 0 <= s < geom.n_elements:
            raise ValueError(f"src_list entry {s} is not an element of the array")
    return [rec[s] for s in src], rec


def _setup(m, geom, wavelet, dt, h, nt, sponge, src_list, method=None, footprints=None):
    shape = _shape_of(m)
    if len(shape) not in (2, 3):
        raise ValueError(f"m must be 2D or 3D, got shape {shape}")
    nt = int(nt)
    check_cfl(m, dt, h)
    if len(wavelet) < nt:
        raise ValueError(f"wavelet has {len(wavelet)} samples, nt = {nt}")
    if sponge is not None and len(shape) != 2:
        raise NotImplementedError("the sponge is defined for 2D grids only")
    tx, rec = _tx_rec(geom, shape, src_list)
    groups = None if footprints is None else _element_groups(geom, shape, src_list, footprints)
    plan = get_plan(shape, h, dt, nt, tx, rec, sponge is not None, method, groups)
    md = _to_dev(m, torch.float64)
    wd = _to_dev(np.asarray(wavelet, dtype=np.float64)[:nt], torch.float64)
    sd = None if sponge is None else _to_dev(sponge, torch.float64)
    return plan, md, wd, sd


def forward_fmc(m, geom, wavelet, dt, h, nt, sponge=None, src_list=None, footprints=None):
    """Full-matrix capture (fwi.py:102-119): data (n_src, nt, n_elements) float64.
    ``footprints`` (geometry.build_footprints) models finite-aperture elements."""
    plan, md, wd, sd = _setup(m, geom, wavelet, dt, h, nt, sponge, src_list,
                              footprints=footprints)
    d = plan.forward_fmc(md, wd, sd)
    return d.double().cpu().numpy()


def _dobs_dev(dobs, plan):
    want = (plan.n_tx, plan.nt, plan.n_rec)
    if tuple(np.shape(dobs)) != want:
        raise ValueError(f"dobs has shape {tuple(np.shape(dobs))}, expected {want}")
    return _to_dev(dobs, torch.float64)


def _windows_dev(time_weights, plan):
    """time_weights: a list over transmits of (nt, n_rx) windows (fwi.py:250-253)."""
    if time_weights is None:
        return None
    tw = [np.asarray(w, dtype=np.float64) for w in time_weights]
    if len(tw) != plan.n_tx or any(w.shape != (plan.nt, plan.n_rec) for w in tw):
        raise ValueError(f"time_weights must be {plan.n_tx} arrays of shape "
                         f"{(plan.nt, plan.n_rec)}")
    return _to_dev(np.stack(tw) if tw else np.zeros((0, plan.nt, plan.n_rec)), torch.float64)


def _check_misfit(misfit_type):
    if misfit_type not in ("l2", "gcn") + DEVICE_MISFITS:
        raise ValueError(f"unknown misfit_type {misfit_type!r}")


def misfit(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None, misfit_type="l2",
           footprints=None, time_weights=None):
    """Misfit only (fwi.py:246-263), optionally with footprints and time windows;
    envelope / egcn / traveltime are evaluated on the device traces."""
    _check_misfit(misfit_type)
    plan, md, wd, sd = _setup(m, geom, wavelet, dt, h, nt, sponge, src_list,
                              footprints=footprints)
    if misfit_type in DEVICE_MISFITS:
        d = plan.forward_fmc(md, wd, sd)
        J, _ = trace_misfits.misfit_and_source(d, _dobs_dev(dobs, plan), misfit_type,
                                               _windows_dev(time_weights, plan))
        return float(J.item())
    J = plan.misfit(md, wd, _dobs_dev(dobs, plan), sd, misfit_type,
                    time_weights=_windows_dev(time_weights, plan))
    return float(J.item())


def misfit_and_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
                        misfit_type="l2", footprints=None, time_weights=None):
    """Misfit and exact discrete-adjoint gradient w.r.t. m (fwi.py:266-332),
    optionally with finite-aperture footprints and time windows.  L2 and GCN
    run fused in the library; envelope / egcn / traveltime take the device
    traces, form their adjoint source with torch on the device and run the
    library's adjoint sweep on it (Plan.gradient_from_source)."""
    _check_misfit(misfit_type)
    plan, md, wd, sd = _setup(m, geom, wavelet, dt, h, nt, sponge, src_list,
                              footprints=footprints)
    if misfit_type in DEVICE_MISFITS:
        d = plan.forward_fmc(md, wd, sd)
        J, adj = trace_misfits.misfit_and_source(d, _dobs_dev(dobs, plan), misfit_type,
                                                 _windows_dev(time_weights, plan))
        g = plan.gradient_from_source(md, wd, adj, sd)
        return float(J.item()), g.cpu().numpy()
    J, g = plan.misfit_and_gradient(md, wd, _dobs_dev(dobs, plan), sd, misfit_type,
                                    time_weights=_windows_dev(time_weights, plan))
    return float(J.item()), g.cpu().numpy()


def simulate_wavefield(m, geom, src_element, wavelet, dt, h, nt, sponge=None):
    """Full pressure history of one transmit (fwi.py:92-99), computed level by
    level on the device with one-level forward runs of growing length."""
    shape = _shape_of(m)
    nt = int(nt)
    check_cfl(m, dt, h)
    cells = linear_indices([geom.element_index(int(src_element))], shape)
    # every grid cell is a receiver: the traces are the history
    rec = np.arange(int(np.prod(shape)), dtype=np.int64)
    plan = Plan(shape, h, dt, nt, cells, rec, method="plain", sponge=sponge is not None)
    try:
        d = plan.forward_fmc(_to_dev(m, torch.float64),
                             _to_dev(np.asarray(wavelet, np.float64)[:nt], torch.float64),
                             None if sponge is None else _to_dev(sponge, torch.float64))
    finally:
        plan.close()
    return d[0].double().cpu().numpy().reshape((nt,) + shape)


def gn_hessian_vector(m, geom, wavelet, dt, h, nt, v, sponge=None, src_list=None,
                      footprints=None, wavefields=None):
    """Gauss-Newton Hessian-vector product (J^T J) v summed over transmits
    (fwi.py:410-431): exact discrete Born forward of v on each transmit's
    background wavefield, then the exact adjoint of the Born traces, / m.

    ``wavefields`` (precomputed background histories in the reference) is
    accepted and not needed: the device recomputes the background wavefield
    from checkpoints inside each call, which gives the same histories.
    """
    plan, md, wd, sd = _setup(m, geom, wavelet, dt, h, nt, sponge, src_list,
                              footprints=footprints)
    if tuple(np.shape(v)) != plan.shape:
        raise ValueError(f"v has shape {tuple(np.shape(v))}, expected {plan.shape}")
    return plan.gn_hessian_vector(md, wd, _to_dev(v, torch.float64), sd).cpu().numpy()


def _dot(a, b):
    return float(torch.sum(a * b).item())


def _invert_gauss_newton(m0, geom, wavelet, dt, h, nt, dobs, sponge, src_list, n_iter,
                         update_mask, m_bounds, n_linesearch, verbose, progress, n_cg, gn_damp,
                         footprints=None):
    """Truncated Gauss-Newton (fwi.py:434-531): conjugate gradients on
    (J^T J + mu I) dm = -g with mu = gn_damp <g, Hg> / <g, g>, then a halving
    line search from step 1.  Model, gradient and CG vectors stay on the device."""
    plan, m, wd, sd = _setup(m0, geom, wavelet, dt, h, nt, sponge, src_list,
                             footprints=footprints)
    dd = _dobs_dev(dobs, plan)
    m = m.clone()
    lo = hi = None
    if m_bounds is not None:
        m0n = np.asarray(m0, dtype=np.float64)
        lo = _to_dev(np.broadcast_to(np.minimum(m_bounds[0], m0n), m0n.shape), torch.float64)
        hi = _to_dev(np.broadcast_to(np.maximum(m_bounds[1], m0n), m0n.shape), torch.float64)
    mask = 1.0 if update_mask is None else _to_dev(update_mask, torch.float64)
    history = []
    for it in range(n_iter):
        J, g = plan.misfit_and_gradient(m, wd, dd, sd, "l2")
        J = float(J.item())
        history.append(J)
        if verbose:
            print(f"  GN iter {it:2d}  misfit = {J:.6e}")
        gm = g * mask
        if float(torch.max(torch.abs(gm)).item()) == 0.0:
            break

        def A(x):
            return plan.gn_hessian_vector(m, wd, x * mask, sd) * mask

        Hg = A(gm)
        mu = gn_damp * _dot(gm, Hg) / (_dot(gm, gm) + 1e-300)
        x = torch.zeros_like(m)
        r = -gm
        d = r.clone()
        rs = rs0 = _dot(r, r)
        for _ in range(n_cg):
            Ad = A(d) + mu * d
            alpha = rs / (_dot(d, Ad) + 1e-300)
            x = x + alpha * d
            r = r - alpha * Ad
            rs_new = _dot(r, r)
            if rs_new < 1e-10 * rs0:
                break
            d = r + (rs_new / rs) * d
            rs = rs_new
        step = 1.0
        accepted = False
        for _ in range(n_linesearch):
            m_try = m + step * x
            if m_bounds is not None:
                clipped = torch.minimum(torch.maximum(m_try, lo), hi)
                m_try = torch.where(mask > 0, clipped, m_try) if update_mask is not None \
                    else clipped
            if float(plan.misfit(m_try, wd, dd, sd, "l2").item()) < J:
                accepted = True
                break
            step *= 0.5
        if not accepted:
            if verbose:
                print("  GN line search failed to reduce misfit; stopping")
            break
        m = m_try
        if progress is not None:
            progress(it + 1, n_iter)
    history.append(float(plan.misfit(m, wd, dd, sd, "l2").item()))
    return m.cpu().numpy(), history


def invert(m0, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None, n_iter=15,
           step_frac=0.02, update_mask=None, m_bounds=None, smooth_sigma=1.0, n_linesearch=8,
           verbose=False, backend=None, misfit_type="l2", footprints=None, progress=None,
           time_weights=None, optimizer="gd", n_cg=8, gn_damp=1e-2):
    """Preconditioned steepest descent with backtracking line search
    (fwi.py:535-667), every misfit and gradient evaluated on the device.

    Steps per iteration: mask, Gaussian-smooth (scipy, as the reference), step
    = step_frac * mean(m) / max|g|, halve up to n_linesearch times until the
    misfit drops, clip to the effective bounds inside the update mask.
    ``optimizer="gauss-newton"`` (or "gn") runs truncated Gauss-Newton
    (fwi.py:434-531, L2 only, no time weights) with device Hessian-vector
    products; as in the reference it ignores ``backend``.
    """
    from scipy.ndimage import gaussian_filter

    if optimizer in ("gauss-newton", "gn"):
        if misfit_type != "l2":
            raise ValueError("Gauss-Newton is defined for the L2 misfit; select misfit_type='l2'.")
        if time_weights is not None:
            raise ValueError("Gauss-Newton does not support time_weights yet.")
        return _invert_gauss_newton(m0, geom, wavelet, dt, h, nt, dobs, sponge, src_list,
                                    n_iter, update_mask, m_bounds, n_linesearch, verbose,
                                    progress, n_cg, gn_damp, footprints)
    if footprints is not None or time_weights is not None:
        # the reference evaluates these on its Python path whatever the backend
        # (fwi.py:604-615); here the device runs them
        mg = lambda mm: misfit_and_gradient(mm, geom, wavelet, dt, h, nt, dobs, sponge,  # noqa
                                            src_list, misfit_type, footprints, time_weights)
        mf = lambda mm: misfit(mm, geom, wavelet, dt, h, nt, dobs, sponge, src_list,  # noqa
                               misfit_type, footprints, time_weights)
    else:
        # as fwi.py:594-615: a backend serves L2 and GCN; other misfits run here
        be = backend if backend is not None and misfit_type in ("l2", "gcn") else _THIS
        mg = lambda mm: be.misfit_and_gradient(mm, geom, wavelet, dt, h, nt, dobs, sponge,  # noqa
                                               src_list, misfit_type=misfit_type)
        mf = lambda mm: be.misfit(mm, geom, wavelet, dt, h, nt, dobs, sponge, src_list,  # noqa
                                  misfit_type=misfit_type)
    m = np.array(m0, dtype=np.float64, copy=True)
    lo = hi = None
    if m_bounds is not None:
        lo = np.minimum(m_bounds[0], m0)
        hi = np.maximum(m_bounds[1], m0)
    history = []
    J, g = mg(m)
    for it in range(n_iter):
        history.append(J)
        if verbose:
            print(f"  iter {it:2d}  misfit = {J:.6e}")
        d = g * update_mask if update_mask is not None else g.copy()
        if smooth_sigma:
            d = gaussian_filter(d, smooth_sigma)
            if update_mask is not None:
                d = d * update_mask
        dmax = float(np.max(np.abs(d)))
        if dmax == 0.0:
            break
        step = step_frac * float(np.mean(m)) / dmax
        accepted = False
        for _ in range(n_linesearch):
            trial = m - step * d
            if m_bounds is not None:
                clipped = np.clip(trial, lo, hi)
                trial = np.where(update_mask > 0, clipped, trial) if update_mask is not None \
                    else clipped
            Jt = mf(trial)
            if Jt < J:
                accepted = True
                break
            step *= 0.5
        if not accepted:
            if verbose:
                print("  line search failed to reduce misfit; stopping")
            break
        m = trial
        J, g = mg(m)
        if progress is not None:
            progress(it + 1, n_iter)
    history.append(J)
    return m, history


def p_window_weights(dobs, dt, f0, onset_frac=0.2, keep_cycles=2.5, taper_cycles=0.75):
    """Per-trace P-arrival mute windows from observed data (fwi.py:670-699).

    Onset = first sample whose Hilbert envelope exceeds ``onset_frac`` of the
    trace's envelope maximum; the window is 1 for ``keep_cycles`` periods of f0
    after the onset, then a raised-cosine roll-off over ``taper_cycles`` periods
    (at least one sample).  Returns a list over transmits of (nt, n_rx) weights,
    the ``time_weights`` argument of misfit / misfit_and_gradient.
    """
    from scipy.signal import hilbert

    dobs = np.asarray(dobs)
    n_tx, nt, _ = dobs.shape
    keep = keep_cycles / f0 / dt
    taper = max(1.0, taper_cycles / f0 / dt)
    t = np.arange(nt)[:, None]
    out = []
    for i in range(n_tx):
        env = np.abs(hilbert(dobs[i], axis=0))
        onset = np.argmax(env > onset_frac * (env.max(axis=0, keepdims=True) + 1e-30), axis=0)
        x = np.clip((onset + keep + taper - t) / taper, 0.0, 1.0)
        out.append(0.5 - 0.5 * np.cos(np.pi * x))
    return out


def fwi_from_dataset(dataset, start_c=None, update_frac=0.95, m_bounds=None, n_iter=12,
                     step_frac=0.03, verbose=False, **invert_kw):
    """FWI on a Dataset-like object (fwi.py:702-774): grid from its ground truth
    (else 5 points per wavelength), element positions re-mapped to grid indices,
    update region a disc (2D) or cylinder (3D) of update_frac x array radius."""
    from .geometry import GridArray, remap_positions
    from .phantom import m_to_velocity, velocity_to_m

    ds = dataset
    dim = ds.geometry.dim
    if ds.ground_truth is not None:
        h = ds.ground_truth["h_m"]
        shape = ds.ground_truth["c"].shape
    else:
        h = ds.nominal_speed_m_s / (ds.tx_centre_freq_hz * 5.0)
        npix = int(round(2.2 * ds.geometry.radius_m / h)) + 1
        shape = (npix,) * dim
    n = shape[0]
    domain = (n - 1) * h
    idx = remap_positions(ds.geometry.element_pos, n, h)
    geom = GridArray(idx=idx, n=n, h=h, radius_m=ds.geometry.radius_m, domain_m=domain)
    if start_c is None:
        start_c = np.full(shape, ds.nominal_speed_m_s)
    m0 = velocity_to_m(start_c)
    centre = domain / 2.0
    coords = np.mgrid[tuple(slice(0, n) for _ in range(dim))].astype(float) * h
    axes = range(dim) if dim == 2 else range(1, dim)
    r = np.sqrt(sum((coords[a] - centre) ** 2 for a in axes))
    update_mask = (r <= ds.geometry.radius_m * update_frac).astype(float)
    m_rec, history = invert(m0, geom, ds.tx_wavelet, ds.dt, h, ds.n_samples, ds.data,
                            src_list=list(ds.tx_elements), n_iter=n_iter, step_frac=step_frac,
                            update_mask=update_mask, m_bounds=m_bounds, verbose=verbose,
                            **invert_kw)
    return {"c": m_to_velocity(m_rec), "history": history, "extent": (-centre, domain - centre),
            "h": h, "geom": geom}


import sys as _sys  # noqa: E402

_THIS = _sys.modules[__name__]
