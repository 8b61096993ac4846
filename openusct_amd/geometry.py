"""Element-to-grid mapping of ring (2D) and cylinder (3D) arrays.

The index of an element is the nearest grid node of its physical position,
rounded with Python's ``round`` (half to even) on the same float64 expression
as the reference, because several arrays put elements exactly on ties
(simulation/ringfwi/geometry.py:47-50, 118-122).  The attributes used by the
solver are ``idx`` (list of index tuples), ``n_elements`` and
``element_index(k)`` (software/python/uap.py:8-9).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class RingArray:
    """Ring of ``n_elements`` elements of radius ``radius_m`` centred in a square
    domain of side ``domain_m`` with grid spacing ``h`` (geometry.py:18-69)."""

    n_elements: int
    radius_m: float
    domain_m: float
    h: float

    array_type = "ring"

    def __post_init__(self):
        self.n = int(round(self.domain_m / self.h)) + 1
        centre = self.domain_m / 2.0
        self.angles = np.linspace(0.0, 2.0 * np.pi, self.n_elements, endpoint=False)
        self.xy = np.column_stack([centre + self.radius_m * np.cos(self.angles),
                                   centre + self.radius_m * np.sin(self.angles)])
        # (iy, ix): row is y, column is x
        self.idx = [(int(round(y / self.h)), int(round(x / self.h))) for x, y in self.xy]

    @property
    def grid_shape(self):
        return (self.n, self.n)

    def element_index(self, k):
        return self.idx[k]

    def receiver_indices(self, exclude=None):
        return [ij for k, ij in enumerate(self.idx) if k != exclude]

    @property
    def element_positions(self):
        """(x, y) in metres relative to the array centre."""
        return self.xy - self.domain_m / 2.0


@dataclass
class CylinderArray:
    """``n_rings`` rings of ``per_ring`` elements stacked along z over
    ``height_m`` in a cubic domain (geometry.py:72-137).  Element order is ring
    (z) outer, angle inner; indices are (iz, iy, ix)."""

    n_rings: int
    per_ring: int
    radius_m: float
    height_m: float
    domain_m: float
    h: float

    array_type = "cylinder"

    def __post_init__(self):
        self.n = int(round(self.domain_m / self.h)) + 1
        c = self.domain_m / 2.0
        if self.n_rings == 1:
            zs = np.array([c])
        else:
            zs = c + np.linspace(-self.height_m / 2.0, self.height_m / 2.0, self.n_rings)
        ang = np.linspace(0.0, 2.0 * np.pi, self.per_ring, endpoint=False)
        pts = [[c + self.radius_m * np.cos(a), c + self.radius_m * np.sin(a), z]
               for z in zs for a in ang]
        self.xyz = np.asarray(pts)
        self.idx = [(int(round(z / self.h)), int(round(y / self.h)), int(round(x / self.h)))
                    for x, y, z in self.xyz]
        self.n_elements = len(self.idx)

    @property
    def grid_shape(self):
        return (self.n, self.n, self.n)

    def element_index(self, k):
        return self.idx[k]

    @property
    def element_positions(self):
        return self.xyz - self.domain_m / 2.0


@dataclass
class GridArray:
    """Array given directly by grid index tuples (geometry.py:281-301)."""

    idx: list
    n: int
    h: float
    radius_m: float
    domain_m: float

    @property
    def n_elements(self):
        return len(self.idx)

    def element_index(self, k):
        return self.idx[k]


def remap_positions(element_pos, n, h):
    """Grid indices of centred physical positions on an n-point grid of spacing
    h, as fwi_from_dataset rebuilds them (simulation/ringfwi/fwi.py:742-746)."""
    dim = element_pos.shape[1]
    domain = (n - 1) * h
    abspos = element_pos + domain / 2.0
    return [tuple(int(round(p[a] / h)) for a in reversed(range(dim))) for p in abspos]
