"""eup_drag_cv_noalloc (the control-volume drag without device allocation) against eup_drag_cv,
through the C-ABI.

Two calls in a row on the same handle, the second with a smaller box, must give eup_drag_cv's
fluxX / fluxY bit for bit (the scratch keeps the previous call's contents, so a stale value would
show); in parity mode on the reference's drag case (680 x 340 x 4, drag boxes of R = 100) and in
physics mode with FP32 and FP16 storage.  A box whose face exceeds the fixed scratch is refused
with EUP_EINVAL while eup_drag_cv still computes it.
"""
import sys

import numpy as np
import pytest

from golden_cases import CASES, make_inputs

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def lattice(shape, **kw):
    from euplectella_hpc_amd.lib import Lattice
    return Lattice(*shape, **kw)


def bits(x):
    return np.float64(x).view(np.uint64)


def check_two_calls(lat, big, small):
    for box in (big, small):
        ref = lat.drag_cv(box)
        got = lat.drag_cv(box, noalloc=True)
        announce(f"  box {box}: drag_cv {ref}, noalloc {got}")
        assert bits(got[0]) == bits(ref[0]) and bits(got[1]) == bits(ref[1])
        assert ref[0] != 0.0 and ref[1] != 0.0


def test_noalloc_drag_equals_drag_cv_in_parity_mode(golden):
    from euplectella_hpc_amd.lib import drag_boxes
    rec = golden["drag_r100"]
    mask, pops = make_inputs(CASES["drag_r100"])
    announce("eup_drag_cv_noalloc vs eup_drag_cv, parity drag_r100 680x340x4, two calls")
    boxes = [b for _, b in drag_boxes(rec["l"], rec["m"], rec["n"])]
    big = max(boxes, key=lambda b: (b[1] - b[0]) * (b[3] - b[2]))
    small = min(boxes, key=lambda b: (b[1] - b[0]) * (b[3] - b[2]))
    with lattice((rec["l"], rec["m"], rec["n"]), svisc=rec["svisc"], u_inflow=rec["u_in"], mode="parity",
                 storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(3)
        check_two_calls(lat, big, small)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_drag_equals_drag_cv_in_physics_mode(storage):
    from euplectella_hpc_amd import geometry
    l, m, n = 200, 96, 40
    announce(f"eup_drag_cv_noalloc vs eup_drag_cv, physics {storage} {l}x{m}x{n}, two calls")
    mask = geometry.cylinder(l, m, n, 10, (60, 48))
    with lattice((l, m, n), svisc=(2 * 0.6 - 1) / 6, u_inflow=0.05, mode="physics", storage=storage) as lat:
        lat.set_mask(mask)
        lat.init_equilibrium(0.05, 0.0, 0.0, 1e-4, 7)
        lat.step(50)
        check_two_calls(lat, [30, 95, 20, 80, 1, 40], [45, 80, 33, 63, 5, 30])


def test_noalloc_drag_refuses_a_face_beyond_its_scratch():
    from euplectella_hpc_amd.lib import ERRORS, EupError
    l, m, n = 4, 1030, 1030   # an x face of 1030 x 1030 = 1 060 900 sites > 2^20
    box = [1, 4, 1, m, 1, n]
    announce(f"eup_drag_cv_noalloc refuses a {m}x{n} face; eup_drag_cv computes it")
    with lattice((l, m, n), svisc=(2 * 0.6 - 1) / 6, u_inflow=0.05, mode="physics", storage="f16") as lat:
        lat.set_mask(np.zeros((n, m, l), np.uint8))
        lat.init_equilibrium(0.05, 0.0, 0.0, 0.0, 1)
        lat.step(2)
        with pytest.raises(EupError) as err:
            lat.drag_cv(box, noalloc=True)
        assert ERRORS[-1] in str(err.value) and "no-allocation scratch" in str(err.value)
        fx, fy = lat.drag_cv(box)
        assert np.isfinite(fx) and np.isfinite(fy)
        # the handle is still usable for a box that fits
        ref = lat.drag_cv([1, 4, 1, 500, 1, 500])
        got = lat.drag_cv([1, 4, 1, 500, 1, 500], noalloc=True)
        assert bits(got[0]) == bits(ref[0]) and bits(got[1]) == bits(ref[1])
