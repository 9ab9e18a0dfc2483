"""Shared pieces of the open-x pins: a body that touches no periodic seam, and the sites where the
serial reference step and the periodic-wrap step can differ.

In the reference's serial step (CODE/SRC/bcond.bc.F90) the x faces are the inlet (equilibrium at
(u_in, 0, 0) over the whole plane x = 0, :260-321) and the outlet (equilibrium of the momentum of
x = l into x = l+1, :127-164).  The physics and grey steps use the same inlet and outlet values and
wrap y and z by index.  The two agree on every fluid site except where the serial fill differs
from a periodic wrap:
  * the y-z edge lines (i, 0|m+1, 0|n+1), never written after the start (frozen), pulled by the
    sites with j in {1, m} and k in {1, n};
  * the corner copies a03(0,0,k) = a03(l,m,k) (:547) and a10(l+1,m+1,k) = a10(1,1,k) (:612),
    pulled by the columns (1, 1, k) and (l, m, k);
  * the outlet border fix (:171-204), which copies the outlet ghosts of the edge rows from the
    adjacent row instead of the periodic one, pulled by the sites x = l with j in {1, m} or k in
    {1, n};
  * bounce-back across a periodic seam, which the serial step takes from ghosts one step stale
    (:385-625 run after bcond_obs) -- avoided here by a body away from the seams.
After N steps a difference has moved at most N sites (Chebyshev distance), so the sites farther
than N from those sets must agree bit for bit.
"""
import numpy as np


def sphere_mask(l, m, n, centre, radius):
    """mask[n, m, l] of a solid ball (1-based centre (i, j, k))."""
    k, j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, m + 1), np.arange(1, l + 1), indexing="ij")
    ci, cj, ck = centre
    return (((i - ci) ** 2 + (j - cj) ** 2 + (k - ck) ** 2) <= radius * radius).astype(np.uint8)


def differing_sites(l, m, n):
    """Boolean [n, m, l]: the interior sites that pull a value the serial fill and the periodic
    wrap set differently."""
    bad = np.zeros((n, m, l), dtype=bool)
    jb = [0, m - 1]
    kb = [0, n - 1]
    for k in kb:
        for j in jb:
            bad[k, j, :] = True              # y-z edge lines
    bad[:, 0, 0] = True                      # corner column (1, 1, k)
    bad[:, m - 1, l - 1] = True              # corner column (l, m, k)
    bad[:, jb, l - 1] = True                 # outlet border fix, y edges
    bad[kb, :, l - 1] = True                 # outlet border fix, z edges
    return bad


def dilate(bad, steps):
    """Chebyshev dilation by ``steps`` sites, periodic in y and z, bounded in x."""
    out = bad.copy()
    for _ in range(steps):
        # one site along z, then y, then x: the union is the whole 3 x 3 x 3 neighbourhood
        for ax in (0, 1):
            out = out | np.roll(out, 1, axis=ax) | np.roll(out, -1, axis=ax)
        nxt = out.copy()
        nxt[:, :, 1:] |= out[:, :, :-1]
        nxt[:, :, :-1] |= out[:, :, 1:]
        out = nxt
    return out


def comparable(mask, steps):
    """Fluid sites farther than ``steps`` from every differing site."""
    n, m, l = mask.shape
    return (mask == 0) & ~dilate(differing_sites(l, m, n), steps)
