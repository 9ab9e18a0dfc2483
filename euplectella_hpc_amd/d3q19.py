"""D3Q19 lattice conventions of the sponge-flow solver.

The direction numbering, weights and opposite table follow the reference's direction table
(CODE/SRC/bgk3d.F90:31-49, CODE/SRC/alloca.F90:31-89) and the bounce-back pairs of
CODE/SRC/bcond.obs.F90:56-73.  Directions are numbered 1..19 as in the reference; Python arrays
below are indexed 0..18 (index = q - 1).
"""
import numpy as np

#: lattice velocities c_q, q = 1..19 (row q-1)
C = np.array([
    (1, -1, 0), (1, 0, -1), (1, 1, 0), (1, 0, 1), (1, 0, 0),
    (0, 0, 1), (0, 1, 1), (0, 1, 0), (0, 1, -1), (-1, -1, 0),
    (-1, 0, -1), (-1, 1, 0), (-1, 0, 1), (-1, 0, 0), (0, 0, -1),
    (0, -1, -1), (0, -1, 0), (0, -1, 1), (0, 0, 0)], dtype=np.int64)

#: weights: 1/36 for the 12 diagonal links, 1/18 for the 6 axis links, 1/3 for rest
#: (CODE/SRC/storage.F90:77-81; p2 = 1/36, p1 = 1/18, p0 = 1/3)
W = np.array([1 / 36 if int(np.abs(c).sum()) == 2 else (1 / 18 if int(np.abs(c).sum()) == 1 else 1 / 3)
              for c in C], dtype=np.float64)

#: opposite direction index (0-based): C[OPP[q]] == -C[q]
OPP = np.array([int(np.flatnonzero((C == -c).all(axis=1))[0]) for c in C], dtype=np.int64)

Q = 19
NQ_MOVING = 18


def equilibrium(rho, ux, uy, uz):
    """Second-order BGK equilibrium in the reference's arithmetic form.

    e_q = w_q * rho * (1 + 3 c.u + 1.5 (3 (c.u)^2 - u^2)), evaluated as
    ``w*rho*(1 + rf*(c.u) + qf*(3*(c.u)*(c.u) - u^2))`` (CODE/SRC/flux_X.F90:202-220 form).
    Inputs broadcast; returns an array with a leading axis of 19.
    """
    rho = np.asarray(rho, dtype=np.float64)
    ux = np.asarray(ux, dtype=np.float64)
    uy = np.asarray(uy, dtype=np.float64)
    uz = np.asarray(uz, dtype=np.float64)
    usq = ux * ux + uy * uy + uz * uz
    out = []
    for q in range(Q):
        cu = C[q, 0] * ux + C[q, 1] * uy + C[q, 2] * uz
        out.append(rho * W[q] * (1.0 + 3.0 * cu + 1.5 * (3.0 * cu * cu - usq)))
    return np.stack(out)


def moments(f):
    """rho, rho*u from populations with a leading axis of 19 (CODE/SRC/col.F90:127-135)."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(axis=0)
    j = np.tensordot(C.T.astype(np.float64), f, axes=(1, 0))
    return rho, j
