"""ctypes front end of the FP64 CPU oracle (oracle/usct_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, as the checker.  The product package never imports
this module.

The Python signatures follow the reference's drop-in protocol
(simulation/ringfwi/fwi.py:102-119 ``forward_fmc``, fwi.py:246-263 ``misfit``,
fwi.py:266-332 ``misfit_and_gradient``): ``geom`` only needs ``idx``,
``n_elements`` and ``element_index(k)``; ``src_list`` indexes elements and every
element receives.  The sponge is applied, as in the Python reference.
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
        lib.oracle_forward_fmc.argtypes = [dp, ip, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, lp, ctypes.c_int,
                                           lp, ctypes.c_int, dp, dp, dp]
        lib.oracle_misfit_and_gradient.argtypes = [dp, ip, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, lp,
                                                   ctypes.c_int, lp, ctypes.c_int, dp, dp,
                                                   ctypes.c_int, dp, dp, dp]
        lib.oracle_simulate_wavefield.argtypes = [dp, ip, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int,
                                                  ctypes.c_int64, dp, dp, dp]
        lib.oracle_gn_hessian_vector.argtypes = [dp, ip, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_int, lp,
                                                 ctypes.c_int, lp, ctypes.c_int, dp, dp, dp, dp]
        lib.oracle_born_fmc.argtypes = lib.oracle_gn_hessian_vector.argtypes
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


def _prep(m, geom, src_list, sponge):
    m = np.ascontiguousarray(m, dtype=np.float64)
    rec = linear_indices(geom.idx, m.shape)
    if src_list is None:
        src_list = list(range(geom.n_elements))
    tx = np.ascontiguousarray(rec[np.asarray(list(src_list), dtype=np.int64)]) \
        if len(src_list) else np.zeros(0, np.int64)
    dims = np.asarray(m.shape, dtype=np.int32)
    sp = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    return m, rec, tx, dims, sp


def forward_fmc(m, geom, wavelet, dt, h, nt, sponge=None, src_list=None):
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    data = np.zeros((len(tx), nt, len(rec)))
    rc = _lib().oracle_forward_fmc(_d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   m.ndim, float(h), float(dt), int(nt), _l(tx), len(tx),
                                   _l(rec), len(rec), _d(w), _d(sp), _d(data))
    if rc:
        raise ValueError(f"oracle_forward_fmc failed ({rc})")
    return data


_CODES = {"l2": 0, "gcn": 1}


def misfit_and_gradient(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
                        misfit_type="l2"):
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    dobs = np.ascontiguousarray(dobs, dtype=np.float64)
    if dobs.shape != (len(tx), nt, len(rec)):
        raise ValueError(f"dobs shape {dobs.shape} != {(len(tx), nt, len(rec))}")
    g = np.zeros(m.shape)
    J = ctypes.c_double(0.0)
    rc = _lib().oracle_misfit_and_gradient(
        _d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), m.ndim, float(h),
        float(dt), int(nt), _l(tx), len(tx), _l(rec), len(rec), _d(w), _d(dobs),
        _CODES[misfit_type], _d(sp), _d(g), ctypes.byref(J))
    if rc:
        raise ValueError(f"oracle_misfit_and_gradient failed ({rc})")
    return float(J.value), g


def misfit(m, geom, wavelet, dt, h, nt, dobs, sponge=None, src_list=None,
           misfit_type="l2"):
    d = forward_fmc(m, geom, wavelet, dt, h, nt, sponge, src_list)
    if misfit_type == "l2":
        r = d - dobs
        return 0.5 * float(np.sum(r * r))
    eps = 1e-12
    ns = np.sqrt(np.sum(d * d, axis=1)) + eps
    no = np.sqrt(np.sum(dobs * dobs, axis=1)) + eps
    c = np.sum((d / ns[:, None]) * (dobs / no[:, None]), axis=1)
    return float(np.sum(1.0 - c))


def gn_hessian_vector(m, geom, wavelet, dt, h, nt, v, sponge=None, src_list=None):
    """Gauss-Newton Hessian-vector product (J^T J) v (fwi.py:406-431), FP64."""
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    if v.shape != m.shape:
        raise ValueError(f"v shape {v.shape} != {m.shape}")
    hv = np.zeros(m.shape)
    rc = _lib().oracle_gn_hessian_vector(
        _d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), m.ndim, float(h), float(dt),
        int(nt), _l(tx), len(tx), _l(rec), len(rec), _d(w), _d(v), _d(sp), _d(hv))
    if rc:
        raise ValueError(f"oracle_gn_hessian_vector failed ({rc})")
    return hv


def born_fmc(m, geom, wavelet, dt, h, nt, v, sponge=None, src_list=None):
    """Born (linearised) traces J v of every transmit (fwi.py:351-378), FP64."""
    m, rec, tx, dims, sp = _prep(m, geom, src_list, sponge)
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    dd = np.zeros((len(tx), nt, len(rec)))
    rc = _lib().oracle_born_fmc(
        _d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), m.ndim, float(h), float(dt),
        int(nt), _l(tx), len(tx), _l(rec), len(rec), _d(w), _d(v), _d(sp), _d(dd))
    if rc:
        raise ValueError(f"oracle_born_fmc failed ({rc})")
    return dd


def simulate_wavefield(m, geom, src_element, wavelet, dt, h, nt, sponge=None):
    m, rec, _, dims, sp = _prep(m, geom, [], sponge)
    src = int(rec[src_element])
    w = np.ascontiguousarray(wavelet, dtype=np.float64)
    hist = np.zeros((nt,) + m.shape)
    rc = _lib().oracle_simulate_wavefield(_d(m), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          m.ndim, float(h), float(dt), int(nt), src, _d(w),
                                          _d(sp), _d(hist))
    if rc:
        raise ValueError(f"oracle_simulate_wavefield failed ({rc})")
    return hist
