"""Velocity phantoms and the c <-> m conversion (simulation/ringfwi/phantom.py).

``velocity_to_m`` / ``m_to_velocity`` (phantom.py:13-18), the disc and cylinder
backgrounds (phantom.py:21-33, 166-177) and circular / spherical inclusions
(phantom.py:36-48, 191-204) follow the reference's float64 expressions.  The
ellipsoid inclusion and the seeded inclusion sets are this build's workload
generators (the reference has spheres only).
"""

from __future__ import annotations

import numpy as np


def velocity_to_m(c):
    return 1.0 / (c * c)


def m_to_velocity(m):
    return 1.0 / np.sqrt(m)


def coupling_background(shape, c_specimen, c_couplant, radius_m, h):
    n = shape[0]
    centre = (n - 1) * h / 2.0
    y, x = np.mgrid[0:n, 0:n] * h
    r = np.hypot(x - centre, y - centre)
    return np.where(r <= radius_m, c_specimen, c_couplant).astype(float)


def add_inclusion(c, centre_frac, radius_m, c_inclusion, h):
    n = c.shape[0]
    domain = (n - 1) * h
    cx, cy = centre_frac[0] * domain, centre_frac[1] * domain
    y, x = np.mgrid[0:n, 0:n] * h
    r = np.hypot(x - cx, y - cy)
    out = c.copy()
    out[r <= radius_m] = c_inclusion
    return out


def cylinder_background(shape, c_specimen, c_couplant, radius_m, h):
    nz, ny, nx = shape
    cx = (nx - 1) * h / 2.0
    cy = (ny - 1) * h / 2.0
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(float) * h
    r = np.hypot(x - cx, y - cy)
    return np.where(r <= radius_m, c_specimen, c_couplant)


def add_sphere(c, centre_frac, radius_m, c_inclusion, h):
    nz, ny, nx = c.shape
    cx = centre_frac[0] * (nx - 1) * h
    cy = centre_frac[1] * (ny - 1) * h
    cz = centre_frac[2] * (nz - 1) * h
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(float) * h
    r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2)
    out = c.copy()
    out[r <= radius_m] = c_inclusion
    return out


def add_ellipsoid(c, centre_m, semi_axes_m, c_inclusion, h):
    """Axis-aligned ellipsoid inclusion; centre (x, y, z) and semi-axes in metres.
    Only the bounding box of the ellipsoid is visited."""
    nz, ny, nx = c.shape
    out = c.copy()
    lo = [max(0, int(np.floor((c0 - a) / h)) - 1) for c0, a in zip(centre_m, semi_axes_m)]
    hi = [min(n, int(np.ceil((c0 + a) / h)) + 2) for c0, a, n in zip(centre_m, semi_axes_m,
                                                                       (nx, ny, nz))]
    if any(l >= u for l, u in zip(lo, hi)):
        return out
    z, y, x = np.mgrid[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].astype(float) * h
    q = sum(((v - c0) / a) ** 2 for v, c0, a in zip((x, y, z), centre_m, semi_axes_m))
    box = out[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    box[q <= 1.0] = c_inclusion
    return out


def seeded_discs(n, h, seed, count, c_bg=1500.0, r_mm=(3.0, 8.0), c_range=(1400.0, 1800.0),
                 inside_mm=16.0):
    """Background c_bg with ``count`` circular inclusions drawn from default_rng(seed):
    radius ~ U(r_mm) mm, c ~ U(c_range), centre uniform in the disc of radius
    (inside_mm - radius) about the grid centre; applied in order (later wins)."""
    rng = np.random.default_rng(seed)
    c = np.full((n, n), float(c_bg))
    domain = (n - 1) * h
    for _ in range(count):
        r = rng.uniform(*r_mm) * 1e-3
        cv = rng.uniform(*c_range)
        rho = (inside_mm * 1e-3 - r) * np.sqrt(rng.uniform())
        th = rng.uniform(0.0, 2.0 * np.pi)
        fx = (domain / 2.0 + rho * np.cos(th)) / domain
        fy = (domain / 2.0 + rho * np.sin(th)) / domain
        c = add_inclusion(c, (fx, fy), r, cv, h)
    return c


def seeded_ellipsoids(n, h, seed, count, c_bg=1500.0, a_mm=(3.0, 8.0),
                      c_range=(1400.0, 1800.0), inside_mm=20.0):
    """Cubic background with ``count`` ellipsoids from default_rng(seed): semi-axes
    ~ U(a_mm) mm, c ~ U(c_range), centred inside a cylinder of radius
    (inside_mm - max semi-axis) and the same half-height about the grid centre."""
    rng = np.random.default_rng(seed)
    c = np.full((n, n, n), float(c_bg))
    mid = (n - 1) * h / 2.0
    for _ in range(count):
        ax = rng.uniform(*a_mm, size=3) * 1e-3
        cv = rng.uniform(*c_range)
        lim = max(inside_mm * 1e-3 - ax.max(), 0.0)
        rho = lim * np.sqrt(rng.uniform())
        th = rng.uniform(0.0, 2.0 * np.pi)
        zc = rng.uniform(-lim, lim)
        c = add_ellipsoid(c, (mid + rho * np.cos(th), mid + rho * np.sin(th), mid + zc), ax, cv, h)
    return c


def smooth(field, sigma_cells):
    """Gaussian-smoothed copy of a model (a starting model for FWI;
    simulation/ringfwi/phantom.py:51-55, scipy.ndimage.gaussian_filter)."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(field, sigma_cells)


def sphere_background(shape, c_specimen, c_couplant, radius_m, h):
    """3D model: c_specimen within radius_m of the volume centre, c_couplant
    outside (simulation/ringfwi/phantom.py:180-188)."""
    nz, ny, nx = shape
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx].astype(float) * h
    r = np.sqrt((xx - (nx - 1) * h / 2.0) ** 2 + (yy - (ny - 1) * h / 2.0) ** 2
                + (zz - (nz - 1) * h / 2.0) ** 2)
    return np.where(r <= radius_m, c_specimen, c_couplant)


def _grain_seeds(rng, n_grains, radius_m, centre, zmax=None):
    """Seeds uniform in the specimen disc by rejection (2D), with a uniform
    height in [0, zmax] per accepted seed (3D)."""
    seeds = []
    while len(seeds) < n_grains:
        q = rng.uniform(-radius_m, radius_m, size=2)
        if np.hypot(*q) <= radius_m:
            s = [q[0] + centre, q[1] + centre] if zmax is not None else q + centre
            if zmax is not None:
                s.append(rng.uniform(0.0, zmax))
            seeds.append(s)
    return np.asarray(seeds, dtype=np.float64)


def _nearest_seed(coords, seeds):
    """Index of the nearest seed per cell (the Voronoi tessellation on the grid;
    first seed on ties)."""
    d2 = (coords[0][..., None] - seeds[:, 0]) ** 2 + (coords[1][..., None] - seeds[:, 1]) ** 2
    if len(coords) == 3:
        d2 = d2 + (coords[2][..., None] - seeds[:, 2]) ** 2
    return np.argmin(d2, axis=-1)


def _lloyd(coords, seeds, inside, relax):
    """``relax`` Lloyd iterations: each seed moves to the centroid of its cell
    inside the specimen."""
    for _ in range(int(relax)):
        lab = _nearest_seed(coords, seeds)
        for k in range(len(seeds)):
            sel = (lab == k) & inside
            if sel.any():
                seeds[k] = [c[sel].mean() for c in coords]
    return seeds


def voronoi_polycrystal(shape, n_grains, radius_m, h, rng=None, relax=1):
    """2D Voronoi polycrystal in the specimen disc (simulation/ringfwi/phantom.py:58-107):
    (labels: grain per cell, -1 outside; angles: c-axis angle per grain in
    [0, pi); theta_map: the angle per cell, 0 outside)."""
    if len(shape) != 2:
        raise ValueError("voronoi_polycrystal is 2D; use voronoi_polycrystal_3d")
    rng = np.random.default_rng() if rng is None else rng
    n = shape[0]
    c = (n - 1) * h / 2.0
    y, x = np.mgrid[0:n, 0:n].astype(float) * h
    inside = np.hypot(x - c, y - c) <= radius_m
    seeds = _lloyd((x, y), _grain_seeds(rng, n_grains, radius_m, c), inside, relax)
    labels = np.where(inside, _nearest_seed((x, y), seeds), -1)
    angles = rng.uniform(0.0, np.pi, n_grains)
    return labels, angles, np.where(labels >= 0, angles[np.clip(labels, 0, None)], 0.0)


def voronoi_polycrystal_3d(shape, n_grains, radius_m, h, rng=None, relax=1):
    """3D Voronoi polycrystal in the specimen cylinder (simulation/ringfwi/phantom.py:110-161):
    (labels, c-axis unit vector per grain on the upper hemisphere, colatitude
    per cell, 0 outside)."""
    rng = np.random.default_rng() if rng is None else rng
    nz, ny, nx = shape
    c = (nx - 1) * h / 2.0
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(float) * h
    inside = np.hypot(x - c, y - c) <= radius_m
    seeds = _grain_seeds(rng, n_grains, radius_m, c, zmax=(nz - 1) * h)
    seeds = _lloyd((x, y, z), seeds, inside, relax)
    labels = np.where(inside, _nearest_seed((x, y, z), seeds), -1)
    v = rng.standard_normal((n_grains, 3))
    v[:, 2] = np.abs(v[:, 2])
    axes = v / np.linalg.norm(v, axis=1, keepdims=True)
    colat = np.arccos(np.clip(axes[:, 2], -1.0, 1.0))
    return labels, axes, np.where(labels >= 0, colat[np.clip(labels, 0, None)], 0.0)
