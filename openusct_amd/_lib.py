"""Loader of the C-ABI library ``libusct.so`` (declared in include/usct.h).

There is no CPU fallback: if the HIP library is missing, importing the solver
raises.  Build it with ``python -c "import __graft_entry__ as g; g.build()"``
from the repository root (or :func:`openusct_amd.build.build`).
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libusct.so")

USCT_OK, USCT_EINVAL, USCT_ENOMEM, USCT_EHIP, USCT_ENOTSUP = range(5)
METHOD_AUTO, METHOD_PLAIN, METHOD_BATCHED = 0, 1, 2
FLAG_SPONGE, FLAG_AUTOTUNE, FLAG_NO_GRAPH, FLAG_FORWARD_ONLY, FLAG_NO_FUSE = 1, 2, 4, 8, 16
FLAG_NO_REVERSE, FLAG_REVERSE, FLAG_FP32_LO = 32, 64, 128
FLAG_LEVELS_2, FLAG_LEVELS_3 = 256, 512
MISFIT_CODES = {"l2": 0, "gcn": 1}

# Every symbol declared in include/usct.h.
EXPORTS = (
    "usct_plan_create", "usct_plan_create_groups", "usct_plan_destroy", "usct_device_cache",
    "usct_plan_get_info", "usct_forward_fmc",
    "usct_misfit", "usct_misfit_and_gradient", "usct_misfit_tw", "usct_misfit_and_gradient_tw",
    "usct_gn_hessian_vector", "usct_gradient_from_source", "usct_gsot_workspace_bytes", "usct_gsot",
    "usct_profile_gradient",
    "usct_forward_fmc_direct", "usct_misfit_and_gradient_direct", "usct_uap_forward_fmc",
    "usct_uap_misfit_and_gradient", "usct_forward_fmc_noalloc", "usct_last_error",
    "usct_version",
)


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("lanes_per_group", ctypes.c_int),
        ("tile_x", ctypes.c_int),
        ("tile_y", ctypes.c_int),
        ("rows_per_wave", ctypes.c_int),
        ("shots_padded", ctypes.c_int),
        ("shot_passes", ctypes.c_int),
        ("shots_per_pass", ctypes.c_int),
        ("ckpt_every", ctypes.c_int),
        ("n_windows", ctypes.c_int),
        ("workspace_bytes", ctypes.c_int64),
        ("fwd_lanes_per_group", ctypes.c_int),
        ("fwd_tile_x", ctypes.c_int),
        ("fwd_tile_y", ctypes.c_int),
        ("fwd_rows_per_wave", ctypes.c_int),
        ("tile_kind", ctypes.c_int),
        ("fwd_tile_kind", ctypes.c_int),
        ("reverse", ctypes.c_int),
        ("sched", ctypes.c_int),
        ("fwd_sched", ctypes.c_int),
        ("strip_block", ctypes.c_int),
        ("fwd_strip_block", ctypes.c_int),
        ("levels_per_launch", ctypes.c_int),
        ("fwd_levels_per_launch", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class USCTError(RuntimeError):
    """Failure reported by libusct (HIP error or resource exhaustion)."""


_lib = None


def load(path=None):
    """Load libusct.so (``LIB_PATH`` unless ``path`` is given) and declare its
    signatures (raises if it is missing).  torch is imported first: the library
    must bind to the HIP runtime torch ships, not load a second one."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: the HIP extension has not been built "
            "(run __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    vp, ip, dp, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    intp = ctypes.POINTER(ctypes.c_int)
    lib.usct_plan_create.argtypes = [ctypes.POINTER(vp), ip, intp, dp, dp, ip, i64p, ip, i64p, ip,
                                     ip, ip, ctypes.c_int64, ip]
    lib.usct_plan_create_groups.argtypes = [ctypes.POINTER(vp), ip, intp, dp, dp, ip,
                                            ip, i64p, i64p, vp, ip, i64p, i64p, vp,
                                            ip, ip, ctypes.c_int64, ip]
    lib.usct_plan_destroy.argtypes = [vp]
    lib.usct_device_cache.argtypes = [ctypes.c_int64, i64p]
    lib.usct_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    lib.usct_forward_fmc.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.usct_misfit.argtypes = [vp, vp, vp, vp, vp, ip, vp, vp]
    lib.usct_misfit_and_gradient.argtypes = [vp, vp, vp, vp, vp, ip, vp, vp, vp]
    lib.usct_gn_hessian_vector.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.usct_gradient_from_source.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.usct_gsot_workspace_bytes.argtypes = [ip, ip, ip]
    lib.usct_gsot_workspace_bytes.restype = ctypes.c_int64
    lib.usct_gsot.argtypes = [vp, vp, ip, ip, ip, dp, vp, ctypes.c_int64, vp, vp, vp, vp]
    lib.usct_misfit_tw.argtypes = [vp, vp, vp, vp, vp, vp, ip, vp, vp]
    lib.usct_misfit_and_gradient_tw.argtypes = [vp, vp, vp, vp, vp, vp, ip, vp, vp, vp]
    lib.usct_profile_gradient.argtypes = [vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), intp, intp, intp, vp]
    lib.usct_forward_fmc_direct.argtypes = [vp, intp, ip, dp, dp, ip, i64p, ip, i64p, ip, vp, fp,
                                            ip, vp]
    lib.usct_misfit_and_gradient_direct.argtypes = [vp, intp, ip, dp, dp, ip, i64p, ip, i64p, ip,
                                                    vp, vp, vp, ctypes.POINTER(ctypes.c_double),
                                                    ip, ip, vp]
    lib.usct_uap_forward_fmc.argtypes = [vp, intp, ip, dp, dp, ip, intp, ip, intp, ip, vp, vp]
    lib.usct_uap_misfit_and_gradient.argtypes = [vp, intp, ip, dp, dp, ip, intp, ip, intp, ip,
                                                 vp, vp, vp, ip]
    lib.usct_uap_misfit_and_gradient.restype = ctypes.c_double
    lib.usct_forward_fmc_noalloc.argtypes = [vp, intp, ip, dp, dp, ip, i64p, ip, i64p, ip, vp,
                                             vp, vp]
    lib.usct_last_error.restype = ctypes.c_char_p
    lib.usct_version.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc):
    """Raise the Python exception matching a usct_status code."""
    if rc == USCT_OK:
        return
    msg = load().usct_last_error().decode(errors="replace")
    if rc == USCT_EINVAL:
        raise ValueError(msg)
    if rc == USCT_ENOTSUP:
        raise NotImplementedError(msg)
    if rc == USCT_ENOMEM:
        raise MemoryError(msg)
    raise USCTError(f"usct status {rc}: {msg}")
