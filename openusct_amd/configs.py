"""The workloads the solver is measured on, as reproducible problem builders.

C1  2D ring 256^2, dx 0.2 mm, dt 50 ns, nt 1000, 64 elements R = 22 mm, shots
    every 8th element (8), 3 seeded inclusions (default_rng(1)).
C1-16  the same with 16 shots (every 4th element).
C2  2D ring 512^2, dx 0.2 mm, dt 50 ns, nt 1000, 128 elements, 128 shots, seed 2.
    (dx = 0.1 mm with dt = 50 ns violates the 2D CFL limit, so the grid keeps
    C1's spacing and doubles the domain.)
C3  2D ring 1024^2, dx 0.1 mm, dt 25 ns, nt 4000, 256 elements at R = 44 mm
    (C1's radius / domain ratio), 256 shots, seed 3.
C4  3D cylinder 256^3, dx 0.25 mm, dt 30 ns, nt 2000, 8 rings x 64 elements at
    R = 28 mm over a 28 mm height, 64 shots (every 8th element), 8 seeded
    ellipsoids (default_rng(4)), checkpoint interval 50.
Every model is 1500 m/s water with inclusions of c in [1400, 1800] m/s, Ricker
500 kHz source; the starting model is homogeneous 1500 m/s.

``scale`` shrinks the time axis (and ``shots`` the shot list) for tests that
need the same geometry at a fraction of the cost.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import phantom
from .geometry import CylinderArray, RingArray
from .sources import ricker


@dataclass
class Problem:
    name: str
    geom: object
    h: float
    dt: float
    nt: int
    wavelet: np.ndarray
    c_true: np.ndarray
    m_true: np.ndarray
    m0: np.ndarray
    src_list: list
    ckpt_every: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.m0.shape

    @property
    def updates_per_gradient(self):
        return 2 * int(np.prod(self.shape)) * (self.nt - 1) * len(self.src_list)


def _ring_problem(name, n_el, radius, domain, h, dt, nt, seed, shots_every, inside_mm,
                  nt_scale=1.0, shots=None):
    geom = RingArray(n_elements=n_el, radius_m=radius, domain_m=domain, h=h)
    n = geom.n
    nt = max(2, int(round(nt * nt_scale)))
    w = ricker(nt, dt, 5e5)
    c = phantom.seeded_discs(n, h, seed, 3, inside_mm=inside_mm)
    src = list(range(0, n_el, shots_every))
    if shots is not None:
        src = src[:shots]
    return Problem(name, geom, h, dt, nt, w, c, phantom.velocity_to_m(c),
                   np.full((n, n), 1.0 / 1500.0 ** 2), src)


def c1(nt_scale=1.0, shots=None):
    return _ring_problem("C1", 64, 0.022, 0.051, 2e-4, 5e-8, 1000, 1, 8, 16.0, nt_scale, shots)


def c1_16(nt_scale=1.0, shots=None):
    """The 256^2 problem with 16 shots (every 4th element): a small request, where
    the per-launch cost of a gradient's ~1000 launches is largest."""
    return _ring_problem("C1-16", 64, 0.022, 0.051, 2e-4, 5e-8, 1000, 1, 4, 16.0, nt_scale, shots)


def c2(nt_scale=1.0, shots=None):
    return _ring_problem("C2", 128, 0.022, 0.1022, 2e-4, 5e-8, 1000, 2, 1, 16.0, nt_scale, shots)


def c2_256(nt_scale=1.0, shots=None):
    """512^2 with a 256-element ring and 256 shots (plain vs batched comparison)."""
    return _ring_problem("C2-256", 256, 0.022, 0.1022, 2e-4, 5e-8, 1000, 2, 1, 16.0, nt_scale,
                         shots)


def c3(nt_scale=1.0, shots=None):
    return _ring_problem("C3", 256, 0.044, 0.1023, 1e-4, 2.5e-8, 4000, 3, 1, 32.0, nt_scale, shots)


def c4(nt_scale=1.0, shots=None, n=256):
    """3D cylinder; ``n`` < 256 keeps spacing and array, shrinking the domain."""
    h = 2.5e-4
    domain = (n - 1) * h
    radius = min(0.028, 0.44 * domain)
    height = min(0.028, 0.44 * domain)
    geom = CylinderArray(n_rings=8, per_ring=64, radius_m=radius, height_m=height,
                         domain_m=domain, h=h)
    nt = max(2, int(round(2000 * nt_scale)))
    dt = 3e-8
    w = ricker(nt, dt, 5e5)
    f = domain / 0.06375  # reduced grids shrink the inclusions with the domain
    c = phantom.seeded_ellipsoids(geom.n, h, 4, 8, a_mm=(3.0 * f, 8.0 * f), inside_mm=20.0 * f)
    src = list(range(0, geom.n_elements, 8))
    if shots is not None:
        src = src[:shots]
    return Problem("C4", geom, h, dt, nt, w, c, phantom.velocity_to_m(c),
                   np.full((geom.n,) * 3, 1.0 / 1500.0 ** 2), src, ckpt_every=50)


BUILDERS = {"C1": c1, "C1-16": c1_16, "C2": c2, "C2-256": c2_256, "C3": c3, "C4": c4}
