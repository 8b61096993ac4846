"""CPU run of the batched 2D and 3D kernels (no device needed).

tests/cpu_emu/emu_kernels.cpp compiles openusct_amd/csrc/usct_kernels.hpp with
g++ against a stand-in for the HIP device API (tests/cpu_emu/hip/hip_runtime.h)
and runs the kernels lane by lane under AddressSanitizer.  It checks that the
two-level (temporally blocked) forward and adjoint kernels produce bit for bit
what two single-level launches produce, in every row-chunk order and march
direction they can be launched with -- fields, bfloat16 low parts, traces
and gradient accumulators -- and that the 2D (with and without the sponge)
and 3D forward kernels equal a direct loop over cells of their FP64 update and
the 3D adjoint (patch and LDS-tile kernels) a direct loop of its FP32 update
and imaging sums, the 3D adjoint that runs the forward backwards (k_radj3d) a
direct loop of both updates and its imaging sums, the fused 3D Gauss-Newton
level (k_born3d) k_fwd3d followed by k_born, the Born step a direct loop of its FP64 update, the two-level
3D forwards (LDS tiles, and lane patches marching z in registers) two single-level launches, the
forward kernels and the reversing 3D adjoint with FP32 low parts (USCT_FLAG_FP32_LO) against their direct
loops and two single-level launches, the 3D patch kernels with every strip-block order the tuner may pick, the 2D kernels that advance
two, three or four levels per launch (k_fwd2d_xn, k_adj2d_xn: with bfloat16 and FP32 low parts,
storing every level or only what a gradient keeps) LV single-level launches, the two-level Gauss-Newton launch
(background and Born levels together) two forward and two Born launches, the
GSOT assignment kernel against the optimality certificate of its own dual
potentials and, on short traces, brute force over all permutations,
over odd grids, sources and receivers on borders
and strip edges, duplicate receivers and every lane-group / tile shape the
autotuner may pick, and that no access leaves the guard zones the library
allocates (kGuardCells in usct.hip).
"""

from __future__ import annotations

import os
import platform
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "cpu_emu")


def _guard_cells():
    src = open(os.path.join(ROOT, "openusct_amd", "csrc", "usct.hip")).read()
    return int(re.search(r"constexpr int64_t kGuardCells = (\d+);", src).group(1))


@pytest.fixture(scope="module")
def emu_binary(tmp_path_factory):
    if platform.machine() != "x86_64" or shutil.which("g++") is None:
        pytest.skip("the lane emulator needs g++ on x86-64")
    out = str(tmp_path_factory.mktemp("emu") / "emu_kernels")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-ffp-contract=off", "-Wno-unknown-pragmas", "-I", EMU,
           "-I", os.path.join(ROOT, "openusct_amd", "csrc"), f"-DGUARD_CELLS={_guard_cells()}",
           os.path.join(EMU, "emu_kernels.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("which", ["fwd", "adj", "fwd3d", "lds", "born", "x2-3d", "x2p-3d", "born-x2",
                                   "gsot", "radj3d", "radj3d_ysh", "fwd3d_ysh", "born3d", "lo32", "blk3d",
                                   "xn"])
def test_two_level_kernels_equal_single_level(emu_binary, which, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([emu_binary, str(seed), which], capture_output=True, text=True,
                       timeout=300, env=env)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-20:])
    assert r.returncode == 0, tail
    assert "ALL EQUAL" in r.stdout, tail
    assert r.stdout.count("ok ") == {"fwd": 495, "adj": 485, "fwd3d": 50, "lds": 12,
                                     "born": 5, "x2-3d": 6, "x2p-3d": 10, "born-x2": 52, "gsot": 13,
                                     "radj3d": 28, "radj3d_ysh": 42, "fwd3d_ysh": 13, "born3d": 20, "lo32": 24,
                                     "blk3d": 120, "xn": 524}[which], tail
