"""ctypes front end of the FP64 CPU oracle (oracle/usct_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, as the checker.  The product package never imports
this module.

The Python signatures follow the reference's drop-in protocol
(simulation/ringfwi/fwi.py:102-119 ``forward_fmc``, fwi.py:246-263 ``misfit``,
fwi.py:266-332 ``misfit_and_gradient``): ``geom`` only needs ``idx``,
``n_elements`` and ``element_index(k)``; ``src_list`` indexes elements and every
element receives.  The sponge is applied, as in the Python reference.
``footprints`` (geometry.py:212-278 lists of (cells, weights) per element) and
``time_weights`` (per-transmit (nt, n_rx) windows) follow fwi.py:74-89,
255-262 and 298-319.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so (and the reference core when its sources exist)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lp = ctypes.POINTER(ctypes.c_int64)
        i, d = ctypes.c_int, ctypes.c_double
        grid = [dp, ip, i, d, d, i]
        grp = [i, lp, lp, dp]
        lib.oracle_forward_fmc_g.argtypes = grid + grp + grp + [dp, dp, dp]
        lib.oracle_misfit_and_gradient_g.argtypes = grid + grp + grp + [dp, dp, dp, i, dp, dp,
                                                                         dp]
        lib.oracle_gn_g.argtypes = grid + grp + grp + [dp, dp, dp, dp, dp]
        lib.oracle_simulate_wavefield.argtypes = [dp, ip, i, d, d, i, ctypes.c_int64, dp, dp, dp]
        lib.oracle_num_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _d(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _l(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def num_threads():
    return int(_lib().oracle_num_threads())


def linear_indices(idx, shape):
    """Row-major linear index of each grid index tuple (software/python/uap.py:19-22)."""
    idx = np.asarray(idx, dtype=np.int64).reshape(len(idx), len(shape))
    return np.ravel_multi_index(tuple(idx.T), shape).astype(np.int64)


class _Groups:
    """Elements as CSR groups of cells with weights (footprints, geometry.py:212-278;
    a point element is one cell of weight 1)."""

    def __init__(self, elements, shape):
        cells, w, off = [], [], [0]
        for idxs, wt in elements:
            lin = linear_indices(list(idxs), shape)
            wt = np.asarray(wt, dtype=np.float64).reshape(-1)
            if len(lin) != len(wt):
                raise ValueError("footprint cells and weights differ in length")
            cells.extend(int(c) for c in lin)
            w.extend(float(x) for x in wt)
            off.append(len(cells))
        self.n = len(off) - 1
        self.off = np.asarray(off, dtype=np.int64)
        self.cells = np.asarray(cells if cells else [0], dtype=np.int64)
        self.w = np.asarray(w if w else [0.0], dtype=np.float64)

    def args(self):
        return (self.n, _l(self.off), _l(self.cells), _d(self.w))


def _prep(m, geom, src_list, sponge, footprints=None):
    m = np.ascontiguousarray(m, dtype=np.float64)
    if src_list is None:
        src_list = list(range(geom.n_elements))
    src_list = [int(s) for s in src_list]
    if footprints is None:
        elements = [([tuple(ix)], [1.0]) for ix in geom.idx]
    else:
        elements = list(footprints)
        if len(elements) != geom.n_elements:
            raise ValueError("footprints must hold one entry per element")
    rec = _Groups(elements, m.shape)
    tx = _Groups([elements[s] for s in src_list], m.shape)
    dims = np.asarray(m.shape, dtype=np.int32)
    sp = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    return m, rec, tx, dims, sp


def _grid(m, dims, h, dt, nt):
    return (_d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), m.ndim, float(h),
            float(dt), int(nt))


def forward_fmc(m, geom, wavelet, dt, h, nt, sponge=None, src_list=None, footprints=None):
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge, footprints)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    data = np.zeros((tx.n, nt, rec.n))
    rc = _lib().oracle_forward_fmc_g(*_grid(m, dims, h, dt, nt), *tx.args(), *rec.args(), _d(w),
                                     _d(sp), _d(data))
    if rc:
        raise ValueError(f"oracle_forward_fmc_g failed ({rc})")
    return data


_CODES = {"l2": 0, "gcn": 1}


def _misfit_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge, src_list, misfit_type,
                     footprints, time_weights, want_grad):
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge, footprints)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    dobs = np.ascontiguousarray(dobs, dtype=np.float64)
    if dobs.shape != (tx.n, nt, rec.n):
        raise ValueError(f"dobs shape {dobs.shape} != {(tx.n, nt, rec.n)}")
    tw = None
    if time_weights is not None:
        tw = np.ascontiguousarray(np.stack([np.asarray(x, np.float64) for x in time_weights]))
        if tw.shape != dobs.shape:
            raise ValueError(f"time_weights shape {tw.shape} != {dobs.shape}")
    g = np.zeros(m.shape) if want_grad else None
    J = ctypes.c_double(0.0)
    rc = _lib().oracle_misfit_and_gradient_g(*_grid(m, dims, h, dt, nt), *tx.args(), *rec.args(),
                                             _d(w), _d(dobs), _d(tw), _CODES[misfit_type],
                                             _d(sp), _d(g), ctypes.byref(J))
    if rc:
        raise ValueError(f"oracle_misfit_and_gradient_g failed ({rc})")
    return float(J.value), g


def misfit_and_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
                        misfit_type="l2", footprints=None, time_weights=None):
    """Misfit and exact adjoint gradient (fwi.py:266-332), FP64."""
    return _misfit_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge, src_list, misfit_type,
                            footprints, time_weights, True)


def misfit(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None, misfit_type="l2",
           footprints=None, time_weights=None):
    """Misfit only (fwi.py:246-263), FP64."""
    return _misfit_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge, src_list, misfit_type,
                            footprints, time_weights, False)[0]


def _gn(m, geom, wavelet, dt, h, nt, v, sponge, src_list, footprints, born):
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge, footprints)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    if v.shape != m.shape:
        raise ValueError(f"v shape {v.shape} != {m.shape}")
    hv = None if born else np.zeros(m.shape)
    dd = np.zeros((tx.n, nt, rec.n)) if born else None
    rc = _lib().oracle_gn_g(*_grid(m, dims, h, dt, nt), *tx.args(), *rec.args(), _d(w), _d(v),
                            _d(sp), _d(hv), _d(dd))
    if rc:
        raise ValueError(f"oracle_gn_g failed ({rc})")
    return dd if born else hv


def gn_hessian_vector(m, geom, wavelet, dt, h, nt, v, sponge=None, src_list=None,
                      footprints=None):
    """Gauss-Newton Hessian-vector product (J^T J) v (fwi.py:410-431), FP64."""
    return _gn(m, geom, wavelet, dt, h, nt, v, sponge, src_list, footprints, False)


def born_fmc(m, geom, wavelet, dt, h, nt, v, sponge=None, src_list=None, footprints=None):
    """Born (linearised) traces J v of every transmit (fwi.py:350-378), FP64."""
    return _gn(m, geom, wavelet, dt, h, nt, v, sponge, src_list, footprints, True)


def simulate_wavefield(m, geom, src_element, wavelet, dt, h, nt, sponge=None):
    m = np.ascontiguousarray(m, dtype=np.float64)
    dims = np.asarray(m.shape, dtype=np.int32)
    sp = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    src = int(linear_indices([geom.element_index(int(src_element))], m.shape)[0])
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    hist = np.zeros((nt,) + m.shape)
    rc = _lib().oracle_simulate_wavefield(_d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          m.ndim, float(h), float(dt), int(nt), src, _d(w),
                                          _d(sp), _d(hist))
    if rc:
        raise ValueError(f"oracle_simulate_wavefield failed ({rc})")
    return hist
