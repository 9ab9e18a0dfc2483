"""Discrete momentum balance of one physics-mode step around a body (test helper).

For post-collision populations P (ghosted ``[19, n+2, m+2, l+2]``) and a box
[i0, i1] x [j0, j1] x all k (1-based, z periodic) away from the body and the x ends:
momentum stored in the box after the next step - before = momentum that streaming carries in
through the box surface - the momentum the body takes by bounce-back (the momentum-exchange
force), because the BGK collision conserves momentum at every site.
"""
import numpy as np


def box_momentum(P, box):
    """(momentum stored in the box, net momentum the next streaming carries into it) for the
    post-collision populations P of a physics-mode state; box = [i0, i1] x [j0, j1] x all k
    (1-based, z periodic)."""
    from euplectella_hpc_amd.d3q19 import C
    i0, i1, j0, j1 = box
    m, l = P.shape[2] - 2, P.shape[3] - 2
    ii = np.arange(1, l + 1)[None, :]
    jj = np.arange(1, m + 1)[:, None]

    def inside(i, j):
        return (i >= i0) & (i <= i1) & (j >= j0) & (j <= j1)

    src_in = inside(ii, jj)
    stored, inflow = np.zeros(3), np.zeros(3)
    for q in range(19):
        c = np.asarray(C[q], dtype=np.float64)
        f = P[q, 1:-1, 1:-1, 1:-1].sum(axis=0)  # summed over the periodic z column
        stored += c * f[src_in].sum()
        if q < 18:
            dst_in = inside(ii + int(c[0]), jj + int(c[1]))
            inflow += c * (f[~src_in & dst_in].sum() - f[src_in & ~dst_in].sum())
    return stored, inflow
