"""Trace misfits beyond L2 / GCN, on the device (simulation/ringfwi/fwi.py:134-243).

``envelope`` (L2 of Hilbert envelopes), ``egcn`` (global correlation norm of
the envelopes) and ``traveltime`` (cross-correlation lag per trace) are
evaluated here on torch tensors of all transmits at once, ``(n_tx, nt,
n_rx)``; ``gsot`` (graph-space optimal transport) solves its per-trace
assignment in the library's kernel; with the reference's per-transmit definitions: every reduction the
reference makes over one transmit's ``(nt, n_rx)`` block (the envelope
regulariser's maximum included) is made per transmit.  Each returns the misfit
J (summed over transmits) and the adjoint source dJ/d(traces) in the
reference's form, which :meth:`openusct_amd.plan.Plan.gradient_from_source`
turns into the exact discrete-adjoint gradient.

The Hilbert transform is the FFT analytic signal of scipy.signal.hilbert (no
padding, weights 1, 2, ..., 2, 1 (even length) or 1, 2, ..., 2 (odd)); the
cross-correlation is computed by FFT with room for every lag, so it equals
numpy.correlate(..., "full") up to rounding.
"""

from __future__ import annotations

import torch

DEVICE_MISFITS = ("envelope", "egcn", "traveltime", "gsot")
GSOT_ETA = 1.0  # fwi.py:125


def analytic(x):
    """Analytic signal along the time axis (dim 1) of (n_tx, nt, n_rx) traces."""
    n = x.shape[1]
    h = torch.zeros(n, dtype=x.dtype, device=x.device)
    if n % 2 == 0:
        h[0] = h[n // 2] = 1.0
        h[1:n // 2] = 2.0
    else:
        h[0] = 1.0
        h[1:(n + 1) // 2] = 2.0
    return torch.fft.ifft(torch.fft.fft(x, dim=1) * h[None, :, None], dim=1)


def _per_tx_max(e):
    return e.amax(dim=(1, 2), keepdim=True)


def envelope(dsyn, dobs):
    """fwi.py:168-177."""
    sa = analytic(dsyn)
    es = sa.abs()
    eo = analytic(dobs).abs()
    de = es - eo
    J = 0.5 * torch.sum(de * de, dim=(1, 2))
    er = es + 1e-12 * (_per_tx_max(es) + 1e-300)
    adj = de * dsyn / er - analytic(de * sa.imag / er).imag
    return J.sum(), adj


def egcn(dsyn, dobs):
    """fwi.py:178-192."""
    sa = analytic(dsyn)
    es = sa.abs()
    eo = analytic(dobs).abs()
    eps = 1e-12
    ns = torch.sqrt(torch.sum(es * es, dim=1, keepdim=True)) + eps
    no = torch.sqrt(torch.sum(eo * eo, dim=1, keepdim=True)) + eps
    shat = es / ns
    ohat = eo / no
    c = torch.sum(shat * ohat, dim=1, keepdim=True)
    J = torch.sum(1.0 - c)
    re = -(ohat - c * shat) / ns
    er = es + 1e-12 * (_per_tx_max(es) + 1e-300)
    adj = re * dsyn / er - analytic(re * sa.imag / er).imag
    return J, adj


def _gradient_t(s):
    """numpy.gradient along time (unit spacing, first-order one-sided edges)."""
    g = torch.empty_like(s)
    g[:, 1:-1] = (s[:, 2:] - s[:, :-2]) / 2.0
    g[:, 0] = s[:, 1] - s[:, 0]
    g[:, -1] = s[:, -1] - s[:, -2]
    return g


def traveltime(dsyn, dobs):
    """fwi.py:218-242: per trace the lag of the cross-correlation maximum
    (first maximum, parabolic sub-sample refinement clipped to half a sample),
    J = sum 0.5 lag^2, adjoint source -lag * s' / (s' . s'); traces that are all
    zero on either side contribute nothing."""
    n_tx, nt, n_rx = dsyn.shape
    if nt < 2:
        raise ValueError("the traveltime misfit needs at least 2 time samples")
    s = dsyn.permute(0, 2, 1).reshape(-1, nt)
    o = dobs.permute(0, 2, 1).reshape(-1, nt)
    L = 2 * nt - 1
    c = torch.fft.irfft(torch.fft.rfft(s, n=L, dim=1) * torch.conj(torch.fft.rfft(o, n=L, dim=1)),
                        n=L, dim=1)
    # lags -(nt-1) .. nt-1 in numpy.correlate's "full" order
    xc = torch.cat([c[:, L - (nt - 1):], c[:, :nt]], dim=1)
    k = torch.argmax(xc, dim=1)
    inner = (k > 0) & (k < 2 * nt - 2)
    kc = k.clamp(1, 2 * nt - 3)
    y0 = xc.gather(1, (kc - 1)[:, None])[:, 0]
    y1 = xc.gather(1, kc[:, None])[:, 0]
    y2 = xc.gather(1, (kc + 1)[:, None])[:, 0]
    den = y0 - 2.0 * y1 + y2
    safe = torch.where(den.abs() > 0, den, torch.ones_like(den))
    delta = torch.where(den.abs() > 0, 0.5 * (y0 - y2) / safe, torch.zeros_like(den))
    delta = torch.where(inner, delta.clamp(-0.5, 0.5), torch.zeros_like(delta))
    lag = (k - (nt - 1)).to(s.dtype) + delta
    live = torch.any(s != 0, dim=1) & torch.any(o != 0, dim=1)
    lag = torch.where(live, lag, torch.zeros_like(lag))
    J = torch.sum(0.5 * lag * lag)
    sdot = _gradient_t(s)
    denom = torch.sum(sdot * sdot, dim=1) + 1e-300
    adj = (-lag)[:, None] * sdot / denom[:, None]
    adj = torch.where(live[:, None], adj, torch.zeros_like(adj))
    return J, adj.reshape(n_tx, n_rx, nt).permute(0, 2, 1).contiguous()


def gsot(dsyn, dobs, eta=GSOT_ETA):
    """fwi.py:193-217: graph-space optimal transport per trace, the assignment
    by the library's Hungarian kernel (usct_gsot), one work-group per trace.
    Device tensors only."""
    from . import _lib
    n_tx, nt, n_rx = dsyn.shape
    if not dsyn.is_cuda:
        raise ValueError("the GSOT misfit runs on the device: pass device traces")
    dsyn = dsyn.to(torch.float64).contiguous()
    dobs = dobs.to(device=dsyn.device, dtype=torch.float64).contiguous()
    lib = _lib.load()
    nbytes = int(lib.usct_gsot_workspace_bytes(n_tx, nt, n_rx))
    if nbytes < 0:
        raise ValueError(f"bad trace shape {tuple(dsyn.shape)}")
    ws = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=dsyn.device)
    jt = torch.empty(n_tx * n_rx, dtype=torch.float64, device=dsyn.device)
    adj = torch.empty_like(dsyn)
    status = torch.zeros(1, dtype=torch.int32, device=dsyn.device)
    stream = torch.cuda.current_stream(dsyn.device).cuda_stream
    _lib.check(lib.usct_gsot(dsyn.data_ptr(), dobs.data_ptr(), n_tx, nt, n_rx, float(eta), ws.data_ptr(),
                             nbytes, jt.data_ptr(), adj.data_ptr(), status.data_ptr(), stream))
    if int(status.item()) != 0:
        raise ValueError("GSOT misfit: the traces hold non-finite samples")
    return jt.sum(), adj


_FUNCS = {"envelope": envelope, "egcn": egcn, "traveltime": traveltime, "gsot": gsot}


def misfit_and_source(dsyn, dobs, misfit_type, weights=None):
    """(J, adjoint source) of device traces (n_tx, nt, n_rx).  With ``weights``
    the misfit is of the windowed traces and the source is windowed again
    (fwi.py:295-301, exact chain rule)."""
    try:
        fn = _FUNCS[misfit_type]
    except KeyError:
        raise ValueError(f"unknown device misfit {misfit_type!r}") from None
    if weights is not None:
        J, adj = fn(dsyn * weights, dobs * weights)
        return J, adj * weights
    return fn(dsyn, dobs)
