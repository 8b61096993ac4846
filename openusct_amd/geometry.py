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


def fibonacci_directions(n, hemisphere=False):
    """n near-uniform unit vectors on the sphere, or on the upper hemisphere
    (geometry.py:140-146): a Fibonacci lattice, cos(polar angle) evenly spaced at
    the midpoints of n intervals and azimuths advancing by the golden angle."""
    k = np.arange(n) + 0.5
    cz = (1.0 - k / n) if hemisphere else (1.0 - 2.0 * k / n)
    rho = np.sqrt(np.clip(1.0 - cz * cz, 0.0, 1.0))
    az = np.pi * (1.0 + 5.0 ** 0.5) * k
    return np.column_stack([rho * np.cos(az), rho * np.sin(az), cz])


@dataclass
class ShellArray:
    """Elements on a spherical shell of radius ``radius_m`` about the centre of a
    cubic domain (geometry.py:149-186), over the whole sphere or its upper half
    (z above the centre); indices (iz, iy, ix), rounded half to even."""

    n_elements: int
    radius_m: float
    domain_m: float
    h: float
    hemisphere: bool = False

    def __post_init__(self):
        self.n = int(round(self.domain_m / self.h)) + 1
        c = self.domain_m / 2.0
        self.xyz = c + self.radius_m * fibonacci_directions(self.n_elements, self.hemisphere)
        self.idx = [(int(round(z / self.h)), int(round(y / self.h)), int(round(x / self.h)))
                    for x, y, z in self.xyz]

    @property
    def grid_shape(self):
        return (self.n, self.n, self.n)

    def element_index(self, k):
        return self.idx[k]

    @property
    def element_positions(self):
        return self.xyz - self.domain_m / 2.0

    @property
    def array_type(self):
        return "hemisphere" if self.hemisphere else "sphere"


def HemisphereArray(n_elements, radius_m, domain_m, h):
    """Shell array over the upper hemisphere (geometry.py:189-191)."""
    return ShellArray(n_elements, radius_m, domain_m, h, hemisphere=True)


def SphereArray(n_elements, radius_m, domain_m, h):
    """Shell array over the whole sphere (geometry.py:194-196)."""
    return ShellArray(n_elements, radius_m, domain_m, h, hemisphere=False)


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


def _tangents(normal):
    """Unit vectors spanning the aperture plane of an element with outward
    ``normal``: one in 2D (the normal turned by +90 degrees), two in 3D (the
    normal crossed with +z, or with +x when the normal is within ~25 degrees of
    z, then the normal crossed with that) (geometry.py:199-209)."""
    normal = np.asarray(normal, dtype=np.float64)
    if normal.size == 2:
        return [np.array([-normal[1], normal[0]])]
    axis = np.array([0.0, 0.0, 1.0])
    if abs(np.dot(normal, axis)) > 0.9:
        axis = np.array([1.0, 0.0, 0.0])
    t1 = np.cross(normal, axis)
    t1 /= np.linalg.norm(t1)
    return [t1, np.cross(normal, t1)]


def build_footprints(geom, width_m, shape="flat", height_m=None, max_pts=9):
    """Finite-aperture element footprints on the grid (geometry.py:212-278).

    Each element is a flat aperture in the plane tangent to the array at the
    element: ``"rect"`` (``width_m`` across, ``height_m`` along the array axis in
    3D; a segment in 2D), ``"disc"`` (ellipse with those axes), ``"flat"`` (a
    square: height = width) or ``"point"`` (also any aperture no wider than one
    cell).  Each extent is sampled at min(round(extent / h) + 1, max_pts)
    evenly spaced offsets, the first tangent's offsets varying fastest; every
    sample is rounded (half to even) to its grid node and clipped to the grid;
    repeated nodes are kept once, in first-seen order.  Cylinder apertures are
    oriented about the radial direction in the ring plane, so height runs
    along z.  Returns, per element, (list of index tuples, weights 1 / count).
    """
    pos = np.asarray(geom.element_positions, dtype=np.float64)
    dim = pos.shape[1]
    half = geom.domain_m / 2.0
    h, n = geom.h, geom.n
    if height_m is None or shape == "flat":
        height_m = width_m
    point = shape == "point" or (width_m <= h and height_m <= h)
    cylinder = getattr(geom, "array_type", "") == "cylinder"

    def offsets(extent):
        k = int(np.clip(round(extent / h) + 1, 1, max_pts))
        return np.linspace(-extent / 2.0, extent / 2.0, k) if k > 1 else np.zeros(1)

    out = []
    for e in range(len(pos)):
        if point:
            out.append(([geom.element_index(e)], np.array([1.0])))
            continue
        p = pos[e]
        radial = p.copy()
        if cylinder:
            radial[2] = 0.0
        r = np.linalg.norm(radial)
        normal = radial / r if r > 0 else np.eye(dim)[0]
        tang = _tangents(normal)
        axes = [offsets(width_m)] + ([offsets(height_m)] if len(tang) > 1 else [])
        if len(axes) == 1:
            samples = [(a,) for a in axes[0]]
        else:
            samples = [(a, b) for b in axes[1] for a in axes[0]]
        nodes = {}
        for s in samples:
            if shape == "disc":
                u = s[0] / (width_m / 2.0 + 1e-30)
                v = s[1] / (height_m / 2.0 + 1e-30) if len(s) > 1 else 0.0
                if u * u + v * v > 1.0 + 1e-9:
                    continue
            shift = 0
            for c, t in zip(s, tang):
                shift = shift + c * t
            xyz = p + half + shift
            key = tuple(int(np.clip(round(xyz[d] / h), 0, n - 1)) for d in range(dim - 1, -1, -1))
            nodes.setdefault(key, True)
        cells = list(nodes)
        out.append((cells, np.full(len(cells), 1.0 / len(cells))))
    return out
