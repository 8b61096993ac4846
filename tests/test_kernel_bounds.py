"""CPU check of the batched kernels' addressing (no device needed).

The 2D/3D batched kernels read stencil columns without clamping; the x halo of a
strip may fall before the first or after the last element of a field, into the
guard zone the plan allocates around every stencil operand.  This enumerates
every wavefront, lane and halo access of the kernels' index arithmetic (as
written in openusct_amd/csrc/usct_kernels.hpp) over odd grid sizes and every
tile shape the autotuner may pick, and checks the extremes against the guard
size the library uses (kGuardCells in usct.hip).
"""

from __future__ import annotations

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _guard_cells():
    src = open(os.path.join(ROOT, "openusct_amd", "csrc", "usct.hip")).read()
    return int(re.search(r"constexpr int64_t kGuardCells = (\d+);", src).group(1))


def _cdiv(a, b):
    return -(-a // b)


def _extent_2d(nx, ny, Bp, LB, CX, RY, adjoint):
    W = 64 // LB
    ng, nxs, nyc = Bp // LB, _cdiv(nx, W * CX), _cdiv(ny, RY)
    lo = hi = 0
    for wv in range(ng * nxs * nyc):
        sg, t = wv % ng, wv // ng
        xs, yc = t % nxs, t // nxs
        y0 = yc * RY
        y1 = min(y0 + RY, ny)
        rows = {max(y0 - 1, 0), y0} | {min(y + 1, ny - 1) for y in range(y0, y1)}
        for lane in range(64):
            sb, st = lane % LB, lane // LB
            b, x0 = sg * LB + sb, (xs * W + st) * CX
            if not adjoint and x0 >= nx:
                continue  # the forward kernel returns for such lanes
            for r in rows:
                for i in (-1, CX):  # halo columns; cells lie between them
                    e = (r * nx + x0 + i) * Bp + b
                    lo, hi = min(lo, e), max(hi, e)
    return lo, hi - nx * ny * Bp


def _extent_3d(nx, ny, nz, Bp, LB, CX, CY, adjoint):
    W = 64 // LB
    ng, nxs, nys = Bp // LB, _cdiv(nx, W * CX), _cdiv(ny, CY)
    lo = hi = 0
    for xs in range(nxs):
        for lane in range(64):
            sb, st = lane % LB, lane // LB
            for sg in range(ng):
                b, x0 = sg * LB + sb, (xs * W + st) * CX
                if not adjoint and x0 >= nx:
                    continue
                for z in (0, nz - 1):
                    for ys in (0, nys - 1):
                        for j in (-1, CY):
                            r = min(max(ys * CY + j, 0), ny - 1)
                            for i in (-1, CX):
                                e = ((z * ny + r) * nx + x0 + i) * Bp + b
                                lo, hi = min(lo, e), max(hi, e)
    return lo, hi - nx * ny * nz * Bp


@pytest.mark.parametrize("nx,ny", [(61, 61), (5, 7), (101, 37), (3, 3), (130, 9)])
def test_2d_halo_reads_stay_in_guard(nx, ny):
    g = _guard_cells()
    for LB in (16, 32, 64):
        for Bp in (LB, 2 * LB):
            for CX in (4, 8):
                for RY in (4, 64):
                    for adj in (False, True):
                        lo, over = _extent_2d(nx, ny, Bp, LB, CX, RY, adj)
                        assert lo >= -g * Bp and over < g * Bp, (nx, ny, Bp, LB, CX, lo, over)


@pytest.mark.parametrize("nx,ny,nz", [(21, 21, 21), (5, 6, 7), (33, 17, 9), (64, 64, 8)])
def test_3d_halo_reads_stay_in_guard(nx, ny, nz):
    g = _guard_cells()
    for LB in (16, 32, 64):
        for Bp in (LB, 2 * LB):
            for CX, CY in ((4, 4), (4, 2), (8, 2)):
                if CX * CY > LB:
                    continue
                for adj in (False, True):
                    lo, over = _extent_3d(nx, ny, nz, Bp, LB, CX, CY, adj)
                    assert lo >= -g * Bp and over < g * Bp, (nx, ny, nz, Bp, LB, CX, CY, lo, over)
