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


def _chunk2d(wv, ng, nxs, nyc, sched):
    """(sg, xs, yc, dir) of 2D wavefront wv, as chunk2d in usct_kernels.hpp."""
    if sched & 4:
        sg, t = wv // (nxs * nyc), wv % (nxs * nyc)
    else:
        sg, t = wv % ng, wv // ng
    if sched & 2:
        yc, xs = t % nyc, t // nyc
    else:
        xs, yc = t % nxs, t // nxs
    d = -1 if (sched & 1) and (yc & 1) else 1
    return sg, xs, yc, d


def _extent_2d(nx, ny, Bp, LB, CX, RY, adjoint, sched=0):
    W = 64 // LB
    ng, nxs, nyc = Bp // LB, _cdiv(nx, W * CX), _cdiv(ny, RY)
    lo = hi = 0
    n_waves = _cdiv(ng * nxs * nyc, 4) * 4  # whole blocks of 4 wavefronts
    for wv in range(n_waves):
        sg, xs, yc, d = _chunk2d(wv, ng, nxs, nyc, sched)
        if sg >= ng or xs >= nxs or yc >= nyc:
            continue  # the kernels return for these wavefronts
        y0 = yc * RY
        y1 = min(y0 + RY, ny)
        # every row the march touches: fronts y0-1 .. y1 and up to two rows ahead
        # and behind them, clamped to the grid (two-level kernels)
        fronts = range(y0 - 1, y1 + 1)
        rows = {min(max(f + r * d, 0), ny - 1) for f in fronts for r in (-2, -1, 0, 1, 2)}
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


BLOCKS3D = ((1, 2), (2, 2), (4, 4), (8, 4), (8, 8), (16, 8), (4, 16))  # strip blocks (bx, by)


def _tasks3d(nxs, nys, nzc, bx, by):
    return _cdiv(nxs, bx) * bx * _cdiv(nys, by) * by * nzc


def _strip3d(t, nxs, nys, nzc, bx=1, by=2):
    """(xs, ys, zc, valid) of a 3D wavefront task, as strip3d in usct_kernels.hpp."""
    q = t % (bx * by)
    t //= bx * by
    nbx, nby = _cdiv(nxs, bx), _cdiv(nys, by)
    xs = (t % nbx) * bx + q // by
    t //= nbx
    ys = (t % nby) * by + q % by
    zc = t // nby
    return xs, ys, zc, xs < nxs and ys < nys and zc < nzc


def test_strip_blocks_visit_every_strip_once():
    for nxs, nys, nzc in ((7, 5, 3), (16, 16, 2), (1, 9, 4), (33, 2, 1)):
        for bx, by in BLOCKS3D:
            seen = [_strip3d(t, nxs, nys, nzc, bx, by) for t in range(_tasks3d(nxs, nys, nzc, bx, by))]
            valid = [(x, y, z) for x, y, z, ok in seen if ok]
            assert sorted(valid) == [(x, y, z) for x in range(nxs) for y in range(nys)
                                     for z in range(nzc)], (nxs, nys, nzc, bx, by)


def _ysh_strips(ng, nxs, nys, nzc, waves=4):
    """(sg, xs, ys, zc) of every wavefront of the reversing adjoint's tile kind 4,
    as k_radj3d<..., YSH = true> maps work-group t and wavefront w."""
    nyb = _cdiv(nys, waves)
    out = []
    for t in range(ng * nxs * nyb * nzc):
        sg, t2 = t % ng, t // ng
        xs, t2 = t2 % nxs, t2 // nxs
        for w in range(waves):
            out.append((sg, xs, (t2 % nyb) * waves + w, t2 // nyb))
    return out


def test_y_stacked_work_groups_visit_every_strip_once():
    """Tile kind 4: every strip of every shot group and z chunk once; the strips
    past the last y strip (a part-empty last block) read clamped rows only."""
    for ng, nxs, nys, nzc in ((1, 7, 5, 3), (2, 3, 8, 2), (3, 1, 9, 4), (1, 128, 128, 32)):
        got = _ysh_strips(ng, nxs, nys, nzc)
        valid = sorted(v for v in got if v[2] < nys)
        assert valid == [(g, x, y, z) for g in range(ng) for x in range(nxs) for y in range(nys)
                         for z in range(nzc)], (ng, nxs, nys, nzc)
        assert all(v[3] < nzc for v in got)
        assert len(got) - len(valid) == ng * nxs * nzc * (_cdiv(nys, 4) * 4 - nys)


def _extent_3d(nx, ny, nz, Bp, LB, CX, CY, adjoint, RZ=4, halo=1, blk=(1, 2)):
    W = 64 // LB
    ng, nxs, nys, nzc = Bp // LB, _cdiv(nx, W * CX), _cdiv(ny, CY), _cdiv(nz, RZ)
    lo = hi = 0
    n_waves = _cdiv(ng * _tasks3d(nxs, nys, nzc, *blk), 4) * 4
    for wv in range(n_waves):
        sg, t = wv % ng, wv // ng
        xs, ys, zc, ok = _strip3d(t, nxs, nys, nzc, *blk)
        if not ok:
            continue  # the kernels return for these wavefronts
        z0 = zc * RZ
        planes = {min(max(z, 0), nz - 1) for z in range(z0 - 2, min(z0 + RZ, nz) + 2)}
        zs = (min(planes), max(planes))
        for st in range(W):
            x0 = (xs * W + st) * CX
            if not adjoint and x0 >= nx:
                continue
            for sb in (0, LB - 1):
                b = sg * LB + sb
                for z in zs:
                    for j in (-halo, CY - 1 + halo):
                        r = min(max(ys * CY + j, 0), ny - 1)
                        for i in (-halo, CX - 1 + halo):
                            e = ((z * ny + r) * nx + x0 + i) * Bp + b
                            lo, hi = min(lo, e), max(hi, e)
    return lo, hi - nx * ny * nz * Bp


@pytest.mark.parametrize("nx,ny", [(61, 61), (5, 7), (101, 37), (3, 3), (130, 9)])
def test_2d_halo_reads_stay_in_guard(nx, ny):
    g = _guard_cells()
    for LB in (16, 32, 64):
        for Bp in (LB, 2 * LB):
            for CX in (4, 8):
                for RY in (4, 43, 64):
                    for adj in (False, True):
                        for sched in range(8):
                            lo, over = _extent_2d(nx, ny, Bp, LB, CX, RY, adj, sched)
                            assert lo >= -g * Bp and over < g * Bp, \
                                (nx, ny, Bp, LB, CX, sched, lo, over)


@pytest.mark.parametrize("nx,ny,nz", [(21, 21, 21), (5, 6, 7), (33, 17, 9), (64, 64, 8)])
def test_3d_halo_reads_stay_in_guard(nx, ny, nz):
    g = _guard_cells()
    for LB in (16, 32, 64):
        for Bp in (LB, 2 * LB):
            # every lane patch the tuner may pick (the reversing adjoint's are a subset)
            for CX, CY in ((4, 4), (4, 2), (8, 2), (2, 2), (2, 4), (2, 1), (1, 2), (1, 1)):
                for adj in (False, True):
                    if adj and CX * CY > LB:
                        continue  # the adjoint reduces a lane's cells over its lane group
                    lo, over = _extent_3d(nx, ny, nz, Bp, LB, CX, CY, adj)
                    assert lo >= -g * Bp and over < g * Bp, (nx, ny, nz, Bp, LB, CX, CY, lo, over)
            # the two-level forward on lane patches reads two cells out in x and y
            for CX, CY in ((1, 1), (2, 1), (1, 2)):
                lo, over = _extent_3d(nx, ny, nz, Bp, LB, CX, CY, False, halo=2)
                assert lo >= -g * Bp and over < g * Bp, (nx, ny, nz, Bp, LB, CX, CY, "x2", lo, over)
        # strip blocks change only the order of the strips, not their set
        for blk in BLOCKS3D[1:]:
            lo, over = _extent_3d(nx, ny, nz, 2 * LB, LB, 2, 2, True, blk=blk)
            assert lo >= -g * 2 * LB and over < g * 2 * LB, (nx, ny, nz, LB, blk, lo, over)
