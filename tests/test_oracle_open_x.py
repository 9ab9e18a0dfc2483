"""Open-x pins of the oracle's physics and grey steps: the benchmarked configuration (inlet at
x = 0, outlet at x = l+1, periodic y and z) checked against the reference itself.

* A uniform free stream at the inlet velocity is a fixed point of the step with open x (inlet
  equilibrium, outlet equilibrium of the momentum at x = l, BGK) to rounding.
* Away from the edges where the serial reference's ghost fill differs from a periodic wrap
  (open_x_region), the physics step and the grey step with a binary solid fraction equal the
  serial reference step bit for bit -- and that step is pinned to the reference executable's
  saved states (test_oracle_golden.py).  The inlet and outlet values that the physics and grey
  steps select for the pulls from x = 0 and x = l+1 (CODE/SRC/bcond.bc.F90:127-205,260-321) are
  thereby pinned to the reference, not to the oracle's own reading of it.
* Each pin fails on a copy of the oracle with a planted mistake: two inlet populations swapped
  in the physics / grey pull, or two outlet populations swapped.
"""
import os
import subprocess

import numpy as np
import pytest

from euplectella_hpc_amd import d3q19, geometry
from oracle import oracle as orc
from open_x_region import comparable, sphere_mask

L_, M_, N_ = 40, 28, 26
STEPS = 8
U_IN = 0.05
OMEGA = orc.omega_from_svisc((2 * 0.6 - 1) / 6)


def case():
    mask = sphere_mask(L_, M_, N_, centre=(16, 14, 13), radius=5)  # touches no periodic seam
    pops = geometry.initial_state(L_, M_, N_, U_IN, noise=1e-3, seed=11)
    return mask, pops


def interior(p):
    return p[:, 1:-1, 1:-1, 1:-1]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def free_stream(l, m, n):
    e = d3q19.equilibrium(1.0, U_IN, 0.0, 0.0)
    return e, np.ascontiguousarray(np.broadcast_to(e[:, None, None, None], (19, n + 2, m + 2, l + 2)))


# --------------------------------------------------------------------------------------------
# free stream: a fixed point with open x
# --------------------------------------------------------------------------------------------

def free_stream_error_physics(L=None, steps=50):
    l, m, n = 12, 6, 5
    e, p = free_stream(l, m, n)
    orc.physics_steps(p, np.zeros((n, m, l), np.uint8), OMEGA, U_IN, steps, x_open=True, L=L)
    return np.abs(interior(p) - e[:, None, None, None]).max()


def free_stream_error_grey(L=None, steps=50):
    l, m, n = 12, 6, 5
    e, p = free_stream(l, m, n)
    a = np.ascontiguousarray(interior(p))
    orc.grey_steps(a, np.zeros((n, m, l)), OMEGA, U_IN, steps, x_open=True, L=L)
    return np.abs(a - e[:, None, None, None]).max()


def test_physics_free_stream_is_stationary_with_open_x():
    assert free_stream_error_physics() < 1e-15


def test_grey_free_stream_is_stationary_with_open_x():
    assert free_stream_error_grey() < 1e-15


# --------------------------------------------------------------------------------------------
# region test against the serial reference step
# --------------------------------------------------------------------------------------------

def physics_vs_reference(L=None):
    mask, pops = case()
    ref = orc.parity_steps(pops.copy(), mask, OMEGA, U_IN, STEPS)
    got = orc.physics_steps(pops.copy(), mask, OMEGA, U_IN, STEPS, x_open=True, L=L)
    cmp = comparable(mask, STEPS)
    return interior(ref)[:, cmp], interior(got)[:, cmp]


def grey_vs_reference(L=None):
    mask, pops = case()
    ref = orc.parity_steps(pops.copy(), mask, OMEGA, U_IN, STEPS)
    a = np.ascontiguousarray(interior(pops))
    orc.grey_steps(a, mask.astype(np.float64), OMEGA, U_IN, STEPS, x_open=True, L=L)
    cmp = comparable(mask, STEPS)
    return interior(ref)[:, cmp], a[:, cmp]


def test_compared_region_covers_inlet_outlet_and_body():
    mask, _ = case()
    cmp = comparable(mask, STEPS)
    assert cmp[:, :, 0].sum() >= 50          # sites at x = 1, pulling from the inlet
    assert cmp[:, :, -1].sum() >= 50         # sites at x = l, whose outlet ghost they feed
    near_body = np.zeros_like(cmp)
    k, j, i = np.nonzero(mask)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                near_body[k + dk, j + dj, i + di] = True
    assert (cmp & near_body).sum() >= 100    # fluid sites that bounce back from the sphere


def test_physics_open_x_equals_the_reference_step_away_from_the_edges():
    ref, got = physics_vs_reference()
    assert np.array_equal(bits(ref), bits(got))
    _, pops = case()
    assert not np.array_equal(ref, interior(pops)[:, comparable(case()[0], STEPS)])


def test_grey_binary_open_x_equals_the_reference_step_away_from_the_edges():
    ref, got = grey_vs_reference()
    assert np.array_equal(bits(ref), bits(got))


def test_the_excluded_edges_do_differ():
    """The region restriction is needed: the serial step and the periodic wrap disagree there."""
    mask, pops = case()
    ref = orc.parity_steps(pops.copy(), mask, OMEGA, U_IN, STEPS)
    got = orc.physics_steps(pops.copy(), mask, OMEGA, U_IN, STEPS, x_open=True)
    fluid = mask == 0
    assert not np.array_equal(interior(ref)[:, fluid], interior(got)[:, fluid])


# --------------------------------------------------------------------------------------------
# planted mistakes: each pin must fail on them
# --------------------------------------------------------------------------------------------

PLANTS = {
    # swap inlet populations q = 1 (c = (1,-1,0), w = 1/36) and q = 5 (c = (1,0,0), w = 1/18)
    "inlet": [("x[q] = inl[q];", "x[q] = inl[q == 0 ? 4 : (q == 4 ? 0 : q)];"),
              ("src = inl[q];", "src = inl[q == 0 ? 4 : (q == 4 ? 0 : q)];")],
    # swap outlet populations q = 10 (c = (-1,-1,0)) and q = 14 (c = (-1,0,0))
    "outlet": [("x[q] = o[q - 9];", "x[q] = o[q == 9 ? 4 : (q == 13 ? 0 : q - 9)];"),
               ("src = o[q - 9];", "src = o[q == 9 ? 4 : (q == 13 ? 0 : q - 9)];")],
}


@pytest.fixture(scope="module", params=sorted(PLANTS))
def planted(request, tmp_path_factory):
    src = open(os.path.join(os.path.dirname(orc.__file__), "eup_oracle.c")).read()
    for old, new in PLANTS[request.param]:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    d = tmp_path_factory.mktemp(f"planted_{request.param}")
    c, so = d / "eup_oracle.c", d / "liboracle_planted.so"
    c.write_text(src)
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                    str(c), "-o", str(so)], check=True)
    return request.param, orc.load(str(so))


def test_pins_fail_on_a_planted_boundary_mistake(planted):
    kind, L = planted
    ref, got = physics_vs_reference(L)
    assert not np.array_equal(bits(ref), bits(got))
    ref, got = grey_vs_reference(L)
    assert not np.array_equal(bits(ref), bits(got))
    assert free_stream_error_physics(L) > 1e-3
    assert free_stream_error_grey(L) > 1e-3
