"""Torch-tensor front end of a libusct plan.

A :class:`Plan` fixes the grid, time axis, transmit cells and receiver cells of
a full-matrix capture and holds the device workspace (fields, checkpoints,
traces, the captured hipGraph).  Models, wavelets and observed data vary per
call.  Tensors are device tensors; torch is used only for memory and streams.

The three operations are the reference's drop-in trio
(simulation/ringfwi/fwi.py:102-119 forward_fmc, fwi.py:246-263 misfit,
fwi.py:266-332 misfit_and_gradient), with the same index semantics as the C++
binding (software/python/uap.py:19-30: row-major linear cells, one shot per
transmit cell, every receiver cell records).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_METHODS = {"auto": _lib.METHOD_AUTO, "plain": _lib.METHOD_PLAIN, "batched": _lib.METHOD_BATCHED}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def linear_indices(idx, shape):
    """Row-major linear cell of each grid index tuple (software/python/uap.py:19-22)."""
    idx = np.asarray(idx, dtype=np.int64).reshape(-1, len(shape))
    if idx.size and ((idx < 0).any() or (idx >= np.asarray(shape)).any()):
        raise ValueError(f"element index outside the grid {tuple(shape)}")
    return np.ravel_multi_index(tuple(idx.T), shape).astype(np.int64)


class Plan:
    """A reusable solver plan on the current HIP device."""

    def __init__(self, shape, h, dt, nt, tx_lin, rec_lin, method="batched", ckpt_every=0,
                 mem_budget=0, sponge=False, autotune=True, graph=True, forward_only=False,
                 fuse=True, tx_groups=None, rec_groups=None, reverse=None, fp32_lo=False,
                 levels=None):
        """``tx_lin`` / ``rec_lin``: one linear cell per transmit / receiver.  With
        ``tx_groups`` and ``rec_groups`` (lists of (linear cells, weights) per
        element, as finite-aperture footprints) those replace the point cells and
        ``tx_lin`` / ``rec_lin`` are ignored.  ``reverse`` (3D gradients): None lets
        the plan choose between keeping forward levels and running the forward
        backwards alongside the adjoint; True / False force one.  ``fp32_lo`` carries
        the forward state as FP32 hi + FP32 lo instead of FP32 hi + bfloat16 lo (8
        bytes per shot-cell instead of 6), for time axes so short that the traces'
        rounding is a visible share of the residual; no Gauss-Newton products.
        ``levels`` (batched 2D without sponge): time levels per fused launch, 2 (the
        default) or 3 (fewer bytes per level, more registers and arithmetic: slower
        on MI355X at 512^2 x 128 shots).  The results do not depend on it."""
        lib = _lib.load()
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.h, self.dt, self.nt = float(h), float(dt), int(nt)
        self.groups = tx_groups is not None or rec_groups is not None
        if self.groups:
            if tx_groups is None or rec_groups is None:
                raise ValueError("tx_groups and rec_groups go together")
            tx_csr, rec_csr = _csr(tx_groups), _csr(rec_groups)
            self.tx = tx_csr[1][tx_csr[0][:-1]] if len(tx_groups) else np.zeros(0, np.int64)
            self.rec = rec_csr[1][rec_csr[0][:-1]]
        else:
            self.tx = np.ascontiguousarray(tx_lin, dtype=np.int64).reshape(-1)
            self.rec = np.ascontiguousarray(rec_lin, dtype=np.int64).reshape(-1)
        self.n_tx, self.n_rec = len(self.tx), len(self.rec)
        self.N = int(np.prod(self.shape))
        self.sponge = bool(sponge)
        if method not in _METHODS:
            raise ValueError(f"unknown method {method!r}; expected one of {sorted(_METHODS)}")
        flags = (_lib.FLAG_SPONGE if sponge else 0) | (_lib.FLAG_AUTOTUNE if autotune else 0) | \
            (0 if graph else _lib.FLAG_NO_GRAPH) | (_lib.FLAG_FORWARD_ONLY if forward_only else 0) | \
            (0 if fuse else _lib.FLAG_NO_FUSE) | \
            {None: 0, True: _lib.FLAG_REVERSE, False: _lib.FLAG_NO_REVERSE}[reverse] | \
            (_lib.FLAG_FP32_LO if fp32_lo else 0) | \
            {None: 0, 2: _lib.FLAG_LEVELS_2, 3: _lib.FLAG_LEVELS_3}[levels]
        dims = (ctypes.c_int * self.ndim)(*self.shape)
        handle = ctypes.c_void_p()
        torch.cuda.init()
        i64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))  # noqa: E731
        if self.groups:
            rc = lib.usct_plan_create_groups(
                ctypes.byref(handle), self.ndim, dims, self.h, self.dt, self.nt, self.n_tx,
                i64(tx_csr[0]), i64(tx_csr[1]), tx_csr[2].ctypes.data, self.n_rec,
                i64(rec_csr[0]), i64(rec_csr[1]), rec_csr[2].ctypes.data, _METHODS[method],
                int(ckpt_every), int(mem_budget), flags)
        else:
            rc = lib.usct_plan_create(
                ctypes.byref(handle), self.ndim, dims, self.h, self.dt, self.nt, i64(self.tx),
                self.n_tx, i64(self.rec), self.n_rec, _METHODS[method], int(ckpt_every),
                int(mem_budget), flags)
        _lib.check(rc)
        self._h = handle
        self.device = torch.device("cuda", torch.cuda.current_device())

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().usct_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def info(self):
        inf = _lib.PlanInfo()
        _lib.check(_lib.load().usct_plan_get_info(self._h, ctypes.byref(inf)))
        return inf.as_dict()

    # -- argument checks --------------------------------------------------------
    def _model(self, m):
        if not isinstance(m, torch.Tensor) or m.device.type != "cuda":
            raise TypeError("m must be a CUDA (HIP) torch tensor")
        if tuple(m.shape) != self.shape:
            raise ValueError(f"m has shape {tuple(m.shape)}, plan expects {self.shape}")
        return m.to(torch.float64).contiguous()

    def _wavelet(self, w):
        w = w.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        if w.numel() < self.nt:
            raise ValueError(f"wavelet has {w.numel()} samples, nt = {self.nt}")
        return w

    def _sponge(self, sponge):
        if self.sponge and sponge is None:
            raise ValueError("plan was created for a sponge; pass the sponge mask")
        if not self.sponge and sponge is not None:
            raise ValueError("plan was created without a sponge")
        if sponge is None:
            return None
        sponge = sponge.to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(sponge.shape) != self.shape:
            raise ValueError(f"sponge has shape {tuple(sponge.shape)}, expected {self.shape}")
        return sponge

    def _dobs(self, dobs, what="dobs"):
        want = (self.n_tx, self.nt, self.n_rec)
        if tuple(dobs.shape) != want:
            raise ValueError(f"{what} has shape {tuple(dobs.shape)}, expected {want}")
        return dobs.to(device=self.device, dtype=torch.float64).contiguous()

    def _windows(self, time_weights):
        return None if time_weights is None else self._dobs(time_weights, "time_weights")

    # -- operations -------------------------------------------------------------
    def forward_fmc(self, m, wavelet, sponge=None, out=None, stream=None):
        """Traces (n_tx, nt, n_rec) float64 on the device (row 0 is zero)."""
        m, w, sp = self._model(m), self._wavelet(wavelet), self._sponge(sponge)
        if out is None:
            out = torch.empty((self.n_tx, self.nt, self.n_rec), dtype=torch.float64,
                              device=self.device)
        _lib.check(_lib.load().usct_forward_fmc(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(out),
                                                _stream_handle(stream)))
        return out

    def misfit(self, m, wavelet, dobs, sponge=None, misfit_type="l2", stream=None,
               time_weights=None):
        """Misfit as a 1-element float64 device tensor; ``time_weights`` (the
        shape of dobs) windows every trace (fwi.py:255-262)."""
        if misfit_type not in _lib.MISFIT_CODES:
            raise ValueError(f"unknown misfit_type {misfit_type!r}")
        m, w, sp, d = self._model(m), self._wavelet(wavelet), self._sponge(sponge), self._dobs(dobs)
        tw = self._windows(time_weights)
        J = torch.empty(1, dtype=torch.float64, device=self.device)
        lib = _lib.load()
        if tw is None:
            rc = lib.usct_misfit(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(d),
                                 _lib.MISFIT_CODES[misfit_type], _ptr(J), _stream_handle(stream))
        else:
            rc = lib.usct_misfit_tw(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(d), _ptr(tw),
                                    _lib.MISFIT_CODES[misfit_type], _ptr(J),
                                    _stream_handle(stream))
        _lib.check(rc)
        return J

    def misfit_and_gradient(self, m, wavelet, dobs, sponge=None, misfit_type="l2", stream=None,
                            time_weights=None):
        """(J, g): J a 1-element float64 tensor, g float64 with the model's shape;
        ``time_weights`` as for misfit (fwi.py:298-301)."""
        if misfit_type not in _lib.MISFIT_CODES:
            raise ValueError(f"unknown misfit_type {misfit_type!r}")
        m, w, sp, d = self._model(m), self._wavelet(wavelet), self._sponge(sponge), self._dobs(dobs)
        tw = self._windows(time_weights)
        J = torch.empty(1, dtype=torch.float64, device=self.device)
        g = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        lib = _lib.load()
        if tw is None:
            rc = lib.usct_misfit_and_gradient(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(d),
                                              _lib.MISFIT_CODES[misfit_type], _ptr(J), _ptr(g),
                                              _stream_handle(stream))
        else:
            rc = lib.usct_misfit_and_gradient_tw(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(d),
                                                 _ptr(tw), _lib.MISFIT_CODES[misfit_type],
                                                 _ptr(J), _ptr(g), _stream_handle(stream))
        _lib.check(rc)
        return J, g

    def gradient_from_source(self, m, wavelet, adj, sponge=None, stream=None):
        """dJ/dm for an adjoint source ``adj`` = dJ/d(traces) of shape
        (n_tx, nt, n_rec): the device forward (history kept) and the exact
        adjoint sweep injecting -adj (fwi.py:295-331 with a given residual)."""
        m, w, sp, a = self._model(m), self._wavelet(wavelet), self._sponge(sponge), self._dobs(adj, "adj")
        g = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().usct_gradient_from_source(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(a),
                                                         _ptr(g), _stream_handle(stream)))
        return g

    def gn_hessian_vector(self, m, wavelet, v, sponge=None, out=None, stream=None):
        """Gauss-Newton Hessian-vector product (J^T J) v (fwi.py:410-431), float64
        with the model's shape; the background wavefield is recomputed per call."""
        m, w, sp = self._model(m), self._wavelet(wavelet), self._sponge(sponge)
        if not isinstance(v, torch.Tensor) or tuple(v.shape) != self.shape:
            raise ValueError(f"v must be a tensor of shape {self.shape}")
        v = v.to(device=self.device, dtype=torch.float64).contiguous()
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().usct_gn_hessian_vector(self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(v),
                                                      _ptr(out), _stream_handle(stream)))
        return out

    def profile_gradient(self, m, wavelet, dobs, sponge=None, stream=None):
        """Kernel timing of one gradient: mean duration (ms) of the forward and of the
        adjoint launches that advance ``fwd_levels`` / ``adj_levels`` time levels each,
        and the number of time levels advanced in all (re-computed levels count as
        forward)."""
        m, w, sp, d = self._model(m), self._wavelet(wavelet), self._sponge(sponge), self._dobs(dobs)
        fwd, adj = ctypes.c_double(), ctypes.c_double()
        nf, na, lv = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 2)()
        _lib.check(_lib.load().usct_profile_gradient(
            self._h, _ptr(m), _ptr(w), _ptr(sp), _ptr(d), ctypes.byref(fwd), ctypes.byref(adj),
            ctypes.byref(nf), ctypes.byref(na), lv, _stream_handle(stream)))
        return {"fwd_ms": fwd.value, "adj_ms": adj.value, "n_fwd": nf.value, "n_adj": na.value,
                "fwd_levels": lv[0], "adj_levels": lv[1]}


def _csr(groups):
    """(off, cells, weights) arrays of a list of (linear cells, weights) groups."""
    off, cells, w = [0], [], []
    for c, wt in groups:
        c = np.asarray(c, dtype=np.int64).reshape(-1)
        wt = np.asarray(wt, dtype=np.float64).reshape(-1)
        if len(c) != len(wt) or len(c) == 0:
            raise ValueError("each group needs one weight per cell and at least one cell")
        cells.extend(c.tolist())
        w.extend(wt.tolist())
        off.append(len(cells))
    return (np.asarray(off, dtype=np.int64), np.asarray(cells or [0], dtype=np.int64),
            np.asarray(w or [0.0], dtype=np.float64))


def device_cache(max_bytes):
    """Keep up to ``max_bytes`` of device memory from closed plans for later plans
    (0 frees what is kept and keeps nothing, the default); returns the bytes kept.
    Closing a plan then costs no hipFree (which waits for the whole device) and a
    new plan of a similar size takes kept blocks instead of mapping fresh memory."""
    held = ctypes.c_int64(0)
    _lib.check(_lib.load().usct_device_cache(int(max_bytes), ctypes.byref(held)))
    return int(held.value)


def footprint_bytes(shape, nt, n_tx, n_rec, ckpt_every=0, gradient=True):
    """Device bytes of a batched plan in one pass, the same model as the C planner:
    every field is N * Bp floats (bf16 low parts half that); checkpoint interval
    ``ckpt_every`` or, by default, the memory minimiser K ~ sqrt(2 nt)."""
    N = int(np.prod(shape))
    lb = 64
    for c in (64, 32, 16):
        lb = c
        pad = -(-n_tx // c) * c
        if (pad - n_tx) * 8 <= pad:
            break
    Bp = -(-n_tx // lb) * lb
    K = ckpt_every if ckpt_every > 0 else max(2, int(round((2.0 * max(nt - 1, 1)) ** 0.5)))
    K = min(K, max(nt - 1, 2))
    NS = max(1, -(-(nt - 1) // K))
    field = 4 * N * Bp
    fields = 5 + ((K + 2) + 2 * max(NS - 2, 0) if gradient else 0)
    halves = 5 + (2 * max(NS - 2, 0) if gradient else 0)
    traces = 4 * nt * n_rec * Bp
    b = fields * field + halves * (field // 2) + 5 * traces
    b += (4 * fields + 2 * halves) * 2 * 64 * Bp  # guard zones
    b += 8 * N + 4 * N + 8 * N * max(Bp // 16, 1) + 8 * nt * Bp + 24 * n_rec * Bp + 8 * N
    return int(b + (1 << 20))
