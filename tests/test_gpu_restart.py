"""Restart through the state: N steps, eup_get_state, eup_set_state into a fresh handle, M more
steps must equal N + M uninterrupted steps bit for bit, in every mode and storage: eup_get_state
returns ref + deviation formed in double, eup_set_state rounds f - ref formed in double once, so the
16-bit deviation storages get their stored codes back.  This is what the Fortran host's restore dump
and the drag floor of bench.py (an FP64 state rounded once to FP32 and set back) rely on.

Case: 150 x 28 x 26 with a sphere of radius 5, open x, the start-up noise of the open-x region
tests, 7 + 5 steps.
"""
import sys

import numpy as np
import pytest

from euplectella_hpc_amd import geometry
from open_x_region import sphere_mask

pytestmark = pytest.mark.gpu

SHAPE = (150, 28, 26)
SVISC = (2 * 0.52 - 1) / 6
U_IN = 0.05
N, M = 7, 5


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def make(mode, storage, mask, pops):
    from euplectella_hpc_amd.lib import Lattice
    lat = Lattice(*SHAPE, svisc=SVISC, u_inflow=U_IN, mode=mode, storage=storage)
    if mode == "grey":
        lat.set_grey(mask.astype(np.float64) * 0.7)
    else:
        lat.set_mask(mask)
    lat.set_state(pops)
    return lat


@pytest.mark.parametrize("mode,storage", [("parity", "f64"), ("physics", "f64"), ("physics", "f32"),
                                          ("physics", "f16"), ("physics", "bf16"), ("physics", "e3m12"),
                                          ("grey", "f64"), ("grey", "f32")])
def test_restart_through_the_state_continues_the_run(mode, storage):
    l, m, n = SHAPE
    announce(f"restart {mode} {storage}: {N} steps, get/set state, {M} steps vs {N + M} steps")
    mask = sphere_mask(l, m, n, centre=(40, 14, 13), radius=5)
    pops = geometry.initial_state(l, m, n, U_IN, noise=1e-3, seed=11)
    with make(mode, storage, mask, pops) as a:
        a.step(N + M)
        want = a.get_state()
    with make(mode, storage, mask, pops) as b:
        b.step(N)
        mid = b.get_state()
    with make(mode, storage, mask, mid) as c:
        c.step(M)
        got = c.get_state()
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    bad = got[inner].view(np.uint64) != want[inner].view(np.uint64)
    assert not bad.any(), f"{int(bad.sum())} of {bad.size} values differ"
