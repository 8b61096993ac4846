"""GPU: usct_forward_fmc_noalloc, the plain forward whose buffers all live in
the library's static device workspace, against usct_forward_fmc_direct with
the plain method (the entry point it is a variant of), through the C ABI.

Traces must be equal bit for bit: both run the same one-shot kernels on the
same operands.  Two calls in a row, the second on a smaller problem, check
that nothing the first call left in the shared workspace leaks into the
second; a 3D call follows; a problem beyond the workspace is refused with
USCT_ENOMEM before any work is queued, and the call after it still works.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(shape, n_el, nt, seed):
    from openusct_amd import phantom
    from openusct_amd.sources import ricker
    rng = np.random.default_rng(seed)
    h = 1e-3
    dt = 0.3 * h / 3000.0 / np.sqrt(len(shape))
    c = np.full(shape, 1500.0) + rng.uniform(0.0, 1500.0, size=shape) * (rng.random(shape) < 0.1)
    m = phantom.velocity_to_m(c)
    N = int(np.prod(shape))
    cells = np.unique(rng.choice(N, size=n_el, replace=False)).astype(np.int64)
    rec = np.concatenate([cells, cells[:2]])  # duplicate receivers accumulate
    tx = cells[: max(1, n_el // 2)].copy()
    return m, h, dt, nt, tx, rec, ricker(nt, dt, 0.3e6)


def _run(fn_name, m, h, dt, nt, tx, rec, w, method=None):
    import torch
    from openusct_amd import _lib
    lib = _lib.load()
    dev = torch.device("cuda")
    md = torch.tensor(m, device=dev)
    wd = torch.tensor(w, device=dev)
    data = torch.full((len(tx), nt, len(rec)), np.nan, dtype=torch.float64, device=dev)
    dims = (ctypes.c_int * m.ndim)(*m.shape)
    i64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))  # noqa: E731
    args = [ctypes.c_void_p(md.data_ptr()), dims, m.ndim, h, dt, nt, i64(tx), len(tx), i64(rec),
            len(rec), ctypes.c_void_p(wd.data_ptr()), ctypes.c_void_p(data.data_ptr())]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if method is None:
        rc = getattr(lib, fn_name)(*args, stream)
    else:
        rc = getattr(lib, fn_name)(*args, method, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return data.cpu().numpy()


def _compare(shape, n_el, nt, seed):
    from openusct_amd import _lib
    p = _problem(shape, n_el, nt, seed)
    print(f"[case] no-alloc forward {shape}, {len(p[4])} shots, {len(p[5])} receivers, nt {nt}",
          flush=True)
    a = _run("usct_forward_fmc_noalloc", *p)
    b = _run("usct_forward_fmc_direct", *p, method=_lib.METHOD_PLAIN)
    assert np.isfinite(a).all() and np.abs(a).max() > 0.0
    assert np.array_equal(a, b), float(np.abs(a - b).max())
    return a


def test_noalloc_forward_matches_direct_plain_twice_then_smaller():
    big = _compare((61, 67), 12, 220, 1)
    small = _compare((37, 41), 6, 130, 2)
    assert small.shape[1] < big.shape[1]
    _compare((19, 21, 23), 8, 90, 3)


def test_noalloc_forward_rejects_a_problem_beyond_its_workspace():
    import torch
    from openusct_amd import _lib
    lib = _lib.load()
    shape = (1500, 1500)  # 36 B per cell: 81 MB > 64 MiB
    dims = (ctypes.c_int * 2)(*shape)
    tx = np.array([10], dtype=np.int64)
    rec = np.array([10, 20], dtype=np.int64)
    i64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))  # noqa: E731
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    ptr = ctypes.c_void_p(dummy.data_ptr())
    rc = lib.usct_forward_fmc_noalloc(ptr, dims, 2, 1e-3, 1e-7, 100, i64(tx), 1, i64(rec), 2, ptr,
                                      ptr, None)
    assert rc == _lib.USCT_ENOMEM
    assert b"static workspace" in lib.usct_last_error()
    _compare((29, 31), 4, 60, 4)
