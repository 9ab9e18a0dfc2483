"""Synthetic, seeded geometries for the sponge-flow workloads, and the initial state.

Masks are returned as ``uint8`` arrays ``mask[k, j, i]`` of shape (n, m, l) (C order), which is
the memory order of the reference's ``obs(1:l,1:m,1:n)`` (i fastest).  Site (i, j, k) in the
formulas below is 1-based, like the reference's loop indices.  The streamwise axis is x (i), the
cylinder / tube axis is z (k), and every body spans the periodic z extent.

Definitions (frozen; the parity fixtures and benchmarks depend on them):

* ``cylinder``: solid where (i-ic)^2 + (j-jc)^2 < R^2.  The reference's own generator places the
  cylinder at (lx/4, ly/2) (CODE/SRC/build.obs.F90:283-287); the workloads pass that centre.
* ``ridged_cylinder``: the cylinder plus ``n_ridges`` helical ridges of square section ``w x w``
  on its surface: radial extent R <= r < R + w, tangential half-width w/2 measured at the ridge
  mid-radius rc = R + w/2.  Ridge p sits at angle 2*pi*p/n + k*rate with
  rate = round(lz*tan(45 deg)/(rc*2*pi/n)) * (2*pi/n) / lz, i.e. the twist over the periodic
  height is a whole number of ridge spacings (seamless across the z wrap), and the pitch angle
  is the closest such angle to 45 degrees.
* ``sponge``: a hollow tube of radius R (wall r in [R - t/2, R + t/2), t = 3 by default) whose wall
  keeps only struts: ``nv = round(2 pi R / cell)`` vertical struts, rings every ``cell`` in z
  (``lz // cell`` of them) and, in alternating cells (checkerboard), both diagonals of the cell
  ("double" diagonal struts).  Strut half-width t/2 for vertical/ring struts and t/3 for
  diagonals, measured in surface coordinates (arc length s = R*theta, z).  Each vertical strut
  and ring is displaced by a seeded U[-0.5, 0.5] jitter (in cells of the lattice).  The tube is
  open (no sieve-plate cap) and spans the periodic z extent.  Optional helical ridges on the
  outer surface use the ``ridged_cylinder`` rule with the tube's outer radius.
* ``grey_spheres``: solid fraction s in [0, 1]: union of ``n`` seeded spheres of radius U[4, 8],
  Gaussian-blurred with sigma = 1.5 and clipped to [0, 1].
"""
import math

import numpy as np

from . import d3q19


def _grid_ij(l, m):
    i = np.arange(1, l + 1, dtype=np.float64)[None, :]
    j = np.arange(1, m + 1, dtype=np.float64)[:, None]
    return i, j


def cylinder(l, m, n, radius, centre):
    ic, jc = centre
    i, j = _grid_ij(l, m)
    disk = ((i - ic) ** 2 + (j - jc) ** 2) < radius * radius
    return np.broadcast_to(disk[None, :, :], (n, m, l)).astype(np.uint8)


def _helix_rate(n_ridges, rc, lz):
    spacing = 2.0 * math.pi / n_ridges
    turns = max(1, round(lz * math.tan(math.radians(45.0)) / (rc * spacing)))
    return turns * spacing / lz


def _add_ridges(mask, r_in, width, n_ridges, centre, lz):
    n, m, l = mask.shape
    ic, jc = centre
    i, j = _grid_ij(l, m)
    dx, dy = i - ic, j - jc
    r = np.sqrt(dx * dx + dy * dy)
    theta = np.arctan2(dy, dx)
    rc = r_in + 0.5 * width
    shell = (r >= r_in) & (r < r_in + width)
    rate = _helix_rate(n_ridges, rc, lz)
    spacing = 2.0 * math.pi / n_ridges
    for k in range(n):
        # angular distance to the nearest ridge centre line at height k+1
        phase = (theta - (k + 1) * rate) / spacing
        d = np.abs(phase - np.round(phase)) * spacing * rc
        mask[k] |= (shell & (d < 0.5 * width)).astype(np.uint8)
    return mask


def ridged_cylinder(l, m, n, radius, centre, n_ridges=8, width=3):
    mask = np.array(cylinder(l, m, n, radius, centre))
    return _add_ridges(mask, float(radius), float(width), n_ridges, centre, n)


def sponge(l, m, n, radius, centre, cell=16, thickness=3.0, seed=0, ridges=0, ridge_width=3):
    """Reticular sponge tube (see module docstring); returns mask[k, j, i]."""
    rng = np.random.default_rng(seed)
    ic, jc = centre
    R = float(radius)
    t = float(thickness)
    nv = max(3, int(round(2.0 * math.pi * R / cell)))
    nr = max(1, n // cell)
    circ = 2.0 * math.pi * R
    ds = circ / nv                     # vertical strut spacing in arc length
    dz = n / nr                        # ring spacing
    s_pos = (np.arange(nv) + 0.5) * ds + rng.uniform(-0.5, 0.5, nv)
    z_pos = (np.arange(nr) + 0.5) * dz + rng.uniform(-0.5, 0.5, nr)

    i, j = _grid_ij(l, m)
    dx, dy = i - ic, j - jc
    r = np.sqrt(dx * dx + dy * dy)
    wall = (r >= R - 0.5 * t) & (r < R + 0.5 * t)
    s = (np.arctan2(dy, dx) % (2.0 * math.pi)) * R          # arc length in [0, circ)

    def periodic_dist(x, pos, period):
        d = np.abs(x[..., None] - pos)
        return np.minimum(d, period - d).min(axis=-1)

    ws = wall
    s_w = s[ws]
    dv = periodic_dist(s_w, s_pos, circ)
    vert = dv < 0.5 * t
    # cell index along s for each wall site (between jittered vertical struts)
    order = np.argsort(s_pos)
    sp = s_pos[order]
    a_idx = np.searchsorted(sp, s_w) % nv                   # strut to the right
    a_left = (a_idx - 1) % nv
    s0 = sp[a_left]
    s1 = sp[a_idx]
    width_s = (s1 - s0) % circ
    u = ((s_w - s0) % circ) / np.where(width_s > 0, width_s, 1.0)   # 0..1 across the cell

    mask = np.zeros((n, m, l), dtype=np.uint8)
    zsorted = np.sort(z_pos)
    for k in range(n):
        z = k + 1.0
        dr = np.abs(z - zsorted)
        dr = np.minimum(dr, n - dr).min()
        ring = dr < 0.5 * t
        b_idx = np.searchsorted(zsorted, z) % nr
        b_low = (b_idx - 1) % nr
        z0 = zsorted[b_low]
        z1 = zsorted[b_idx]
        hz = (z1 - z0) % n
        v = ((z - z0) % n) / (hz if hz > 0 else 1.0)
        diag_cell = ((a_left + b_low) % 2) == 0
        # distance to the two diagonals of the cell, in surface units
        d1 = np.abs(u - v) * width_s * hz / np.sqrt(width_s ** 2 + hz ** 2)
        d2 = np.abs(u + v - 1.0) * width_s * hz / np.sqrt(width_s ** 2 + hz ** 2)
        diag = diag_cell & ((d1 < t / 3.0) | (d2 < t / 3.0))
        solid_w = vert | ring | diag
        plane = np.zeros((m, l), dtype=bool)
        plane[ws] = solid_w
        mask[k] = plane
    if ridges:
        _add_ridges(mask, R + 0.5 * t, float(ridge_width), ridges, centre, n)
    return mask


def scattered(l, m, n, fraction=0.1, seed=0):
    """Seeded scattered solid sites over the whole lattice (uint8, shape (n, m, l)), the first and
    last x planes and the y/z domain edges included: the boundary cases of bounce-back next to the
    inlet and outlet ghosts and across the periodic seams."""
    rng = np.random.default_rng(seed)
    return (rng.random((n, m, l)) < fraction).astype(np.uint8)


def grey_spheres(l, m, n, n_spheres=12, rmin=4.0, rmax=8.0, sigma=1.5, seed=0, region=None):
    """Grey-lattice solid fraction s in [0, 1] (float64, shape (n, m, l))."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    if region is None:
        region = ((l // 8, l // 2), (m // 4, 3 * m // 4), (n // 4, 3 * n // 4))
    (x0, x1), (y0, y1), (z0, z1) = region
    k = np.arange(1, n + 1, dtype=np.float64)[:, None, None]
    j = np.arange(1, m + 1, dtype=np.float64)[None, :, None]
    i = np.arange(1, l + 1, dtype=np.float64)[None, None, :]
    solid = np.zeros((n, m, l), dtype=np.float64)
    for _ in range(n_spheres):
        cx = rng.uniform(x0, x1)
        cy = rng.uniform(y0, y1)
        cz = rng.uniform(z0, z1)
        rad = rng.uniform(rmin, rmax)
        solid = np.maximum(solid, (((i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2) < rad * rad)
                           .astype(np.float64))
    s = gaussian_filter(solid, sigma=sigma, mode="wrap")
    return np.clip(s, 0.0, 1.0)


def initial_pops(l, m, n, u_in, noise=1e-4, seed=0):
    """Yield (q, ghosted population q - 1 as ``[n+2, m+2, l+2]`` float64) for q = 1..19.

    rho = 1 and u = (u_in, U[-noise, noise], 0) at every site including the ghost layer, each
    site set to its own equilibrium ``w*(1 + (3 cu + 1.5 (3 cu^2 - u^2)))``.  One population
    at a time keeps the host footprint at a few arrays for large lattices.
    """
    rng = np.random.default_rng(seed)
    shape = (n + 2, m + 2, l + 2)
    ux = float(u_in)
    uy = rng.uniform(-noise, noise, size=shape)
    usq = ux * ux + uy * uy
    for q in range(19):
        cx, cy, cz = (float(c) for c in d3q19.C[q])
        cu = cx * ux + cy * uy
        yield q + 1, d3q19.W[q] * (1.0 + (3.0 * cu + 1.5 * (3.0 * cu * cu - usq)))


def initial_state(l, m, n, u_in, noise=1e-4, seed=0):
    """Ghosted initial populations ``[19, n+2, m+2, l+2]`` (float64); see ``initial_pops``."""
    pops = np.empty((19, n + 2, m + 2, l + 2), dtype=np.float64)
    for q, arr in initial_pops(l, m, n, u_in, noise, seed):
        pops[q - 1] = arr
    return pops


def bbox(mask):
    """Obstacle bounding box (imin, imax, jmin, jmax, kmin, kmax), 1-based, as the reference
    computes it while reading the obstacle file (CODE/SRC/build.obs.F90:73-79,155-163); an empty
    box is (l, 0, m, 0, n, 0)."""
    n, m, l = mask.shape
    k, j, i = np.nonzero(mask)
    if len(i) == 0:
        return (l, 0, m, 0, n, 0)
    return (int(i.min()) + 1, int(i.max()) + 1, int(j.min()) + 1, int(j.max()) + 1,
            int(k.min()) + 1, int(k.max()) + 1)
