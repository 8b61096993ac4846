"""The native C example (tests/c_example/usct_example.c): the problem of the
reference's software/core/example_main.cpp:19-73 through the C ABI from a
plain C program, with no Python on the path.

CPU: the program compiles with gcc as C11 against include/usct.h and links
against the built libusct.so (every symbol it calls resolves).  GPU: it runs
and its checks pass -- reciprocity of the traces, J = 0 and a zero gradient
for the model's own traces, plan and stateless traces bit for bit, an element
outside the grid refused -- and its traces, misfit and gradient equal the
reference's own C++ core (libuap, compiled from its sources into oracle/_ref)
on the inputs the program used, to relative L2 <= 1e-5.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openusct_amd")
ROCM = "/opt/rocm"


def _build(out_dir):
    if shutil.which("gcc") is None or not os.path.exists(os.path.join(ROCM, "include", "hip")):
        pytest.skip("needs gcc and the HIP headers")
    if not os.path.exists(os.path.join(LIBDIR, "libusct.so")):
        pytest.skip("libusct.so is not built")
    exe = os.path.join(str(out_dir), "usct_example")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROCM, "include"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_example", "usct_example.c"), "-L", LIBDIR, "-lusct",
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-lm", f"-Wl,-rpath,{LIBDIR}",
           f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c_example_builds_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_c_example_runs(tmp_path):
    exe = _build(tmp_path)
    print("[case] native C example (41x41, 12-element ring, nt=300)", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "libusct native example OK" in r.stdout
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_c_example_matches_reference_core(tmp_path):
    """The example's arrays against uap::forward_fmc / uap::misfit_and_gradient
    (software/core/libuap.hpp:109-116, 169-224) on the same inputs."""
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    if ref_dir not in sys.path:
        sys.path.insert(0, ref_dir)
    try:
        import _uap
    except ImportError:
        pytest.skip("reference C++ core not built (oracle/_ref)")
    exe = _build(tmp_path)
    out = str(tmp_path / "example.bin")
    print("[case] native C example vs the reference core", flush=True)
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n, nt, ne = 41, 300, 12
    v = np.fromfile(out, dtype=np.float64)
    sizes = [nt, n * n, ne, ne * nt * ne, n * n, 1]
    assert v.size == sum(sizes)
    w, m, el, data, grad, J = np.split(v, np.cumsum(sizes)[:-1])
    m = m.reshape(n, n)
    elem = el.astype(np.int32)
    d_ref = _uap.forward_fmc(m, 1.0e-3, 1.5e-7, nt, elem, elem, w)
    m0 = np.full((n, n), 1.0 / 3000.0 ** 2)
    J_ref, g_ref = _uap.misfit_and_gradient(m0, 1.0e-3, 1.5e-7, nt, elem, elem, w, d_ref, 0)
    e = (np.linalg.norm(data - d_ref.ravel()) / np.linalg.norm(d_ref),
         abs(J[0] - J_ref) / J_ref, np.linalg.norm(grad - g_ref.ravel()) / np.linalg.norm(g_ref))
    print(f"traces {e[0]:.2e}  J {e[1]:.2e}  gradient {e[2]:.2e}")
    assert max(e) < 1e-5
