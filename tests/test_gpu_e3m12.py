"""The e3m12 storage (EUP_STORE_E3M12): deviations (f_q - ref_q) / w_q as 16-bit floats with 1
sign, 3 exponent and 12 mantissa bits (binades 2^-5 .. 4 in units of w_q, a step of 2^-17 w_q
below them, the top code NaN).

The format is restated here in numpy FP32 arithmetic (IEEE, round to nearest even, denormals
kept; the code rounds ties away from zero): a state written through eup_set_state and read back
through eup_get_state must equal the restatement's encode / decode bit for bit, in both deviation
references, with ties among the values; a deviation of 4 w_q or more must come back as NaN, and a
time step must carry a stored NaN into the update of every site that reads it.
"""
import sys

import numpy as np
import pytest

from euplectella_hpc_amd import d3q19

pytestmark = pytest.mark.gpu

F32 = np.float32


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def scaled_bits(d, w):
    """FP32 bits of d 2^-121 / w (the product rounded to nearest even, denormals kept)."""
    return (np.asarray(d, F32) * F32(F32(2.0 ** -121) / F32(w))).astype(F32).view(np.uint32)


def encode(d, w):
    b = scaled_bits(d, w)
    a = np.minimum(b & np.uint32(0x7FFFFFFF), np.uint32(0x03FFFBFF))  # NaN and infinity clamp too
    c = (a + np.uint32(0x400)) >> np.uint32(11)  # nearest, ties away from zero
    return (((b >> np.uint32(16)) & np.uint32(0x8000)) | c).astype(np.uint16)


def decode(x, w):
    b = x.astype(np.uint32)
    v = (((b & np.uint32(0x8000)) << np.uint32(16)) | ((b & np.uint32(0x7FFF)) << np.uint32(11))).view(F32)
    out = (v * F32(F32(w) * F32(2.0 ** 121))).astype(F32)
    return np.where((b & np.uint32(0x7FFF)) == np.uint32(0x7FFF), F32(np.nan), out)


def refs(shifted, u_in):
    """The deviation references of the library: w_q, or the free-stream equilibrium, evaluated in
    double as w (1 + (3 c.u + 1.5 (3 (c.u)^2 - u^2))) and rounded to FP32."""
    u = u_in if shifted else 0.0
    cu = d3q19.C[:, 0].astype(np.float64) * u
    return (d3q19.W * (1.0 + (3.0 * cu + 1.5 * (3.0 * cu * cu - u * u)))).astype(F32)


@pytest.mark.parametrize("shifted", [False, True])
def test_state_round_trip_equals_the_format(shifted):
    from euplectella_hpc_amd.lib import Lattice
    l, m, n, u_in = 70, 5, 4, 0.06
    announce(f"e3m12 set/get state round trip vs the numpy format, shifted={shifted}")
    rng = np.random.default_rng(3)
    w = d3q19.W.astype(F32)
    r = refs(shifted, u_in)
    # deviations spread over every binade of the format, the subnormal range and zero
    mag = 10.0 ** rng.uniform(-8, np.log10(3.9), size=(19, n, m, l))
    dev = np.where(rng.random((19, n, m, l)) < 0.5, -mag, mag) * w[:, None, None, None]
    dev[:, 0, 0, :8] = 0.0
    pops = np.zeros((19, n + 2, m + 2, l + 2))
    pops[:, 1:-1, 1:-1, 1:-1] = r[:, None, None, None].astype(np.float64) + dev
    # values whose scaled FP32 bits end in exactly 0x400 (a tie of the 11 dropped bits), both signs,
    # in the binades and the subnormal range
    for q in range(19):
        cand = r[q].astype(np.float64) + np.where(rng.random(400_000) < 0.5, -1, 1) * \
            10.0 ** rng.uniform(-6, np.log10(3.9), 400_000) * w[q]
        d = (cand - r[q].astype(np.float64)).astype(F32)
        tie = cand[(scaled_bits(d, w[q]) & np.uint32(0x7FF)) == np.uint32(0x400)][:l]
        assert tie.size >= 20, (q, tie.size)
        pops[q, 2, 2, 1:1 + tie.size] = tie
    with Lattice(l, m, n, svisc=0.05, u_inflow=u_in, mode="physics", storage="e3m12", shifted=shifted) as lat:
        lat.set_mask(np.zeros((n, m, l), np.uint8))
        lat.set_state(pops)
        got = lat.get_state()[:, 1:-1, 1:-1, 1:-1]
    want = np.empty_like(got)
    for q in range(19):
        # set_state takes d = f - ref in double, rounded once to FP32; get_state returns ref + the
        # decoded deviation in double
        d = (pops[q, 1:-1, 1:-1, 1:-1] - r[q].astype(np.float64)).astype(F32)
        want[q] = r[q].astype(np.float64) + decode(encode(d, w[q]), w[q]).astype(np.float64)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # and the format's error bounds: 2^-13 relative on the binades, 2^-18 w_q below them
    d = got - pops[:, 1:-1, 1:-1, 1:-1]
    dev = pops[:, 1:-1, 1:-1, 1:-1] - r[:, None, None, None]  # with the tie values in place
    big = np.abs(dev) >= 2.0 ** -5 * w[:, None, None, None]
    assert np.all(np.abs(d[big]) <= 1.3e-4 * np.abs(dev[big]) + 1e-8)


def test_out_of_range_deviation_is_nan():
    from euplectella_hpc_amd.lib import Lattice
    l, m, n = 64, 3, 2
    announce("e3m12: a deviation of 4 w_q or more stores NaN")
    pops = np.broadcast_to(d3q19.W[:, None, None, None], (19, n + 2, m + 2, l + 2)).copy()
    pops[18, 1, 1, 5] += 4.5 * d3q19.W[18]
    pops[18, 1, 1, 6] += 3.9 * d3q19.W[18]
    with Lattice(l, m, n, svisc=0.05, u_inflow=0.0, mode="physics", storage="e3m12") as lat:
        lat.set_mask(np.zeros((n, m, l), np.uint8))
        lat.set_state(pops)
        got = lat.get_state()
    assert np.isnan(got[18, 1, 1, 5]) and np.isfinite(got[18, 1, 1, 6])
    assert np.isfinite(np.delete(got[18, 1, 1, 1:-1], 4)).all()


@pytest.mark.parametrize("q,sign", [(18, 1.0), (18, -1.0), (4, 1.0), (13, -1.0)])
def test_stored_nan_reaches_the_updates_that_read_it(q, sign):
    """The step kernel tests the NaN code once per site (the largest magnitude of its loads): a
    NaN in population q at x = 20 turns the whole update of the site that pulls it NaN (x = 20 for
    the rest population, x = 21 for q = 4 with c_x = +1, x = 19 for q = 13 with c_x = -1), and no
    other site, after one step."""
    from euplectella_hpc_amd.lib import Lattice
    l, m, n = 64, 3, 2
    c = d3q19.C[q]
    announce(f"e3m12: NaN code in population {q} (sign {sign:+.0f}) through one step")
    pops = np.broadcast_to(d3q19.W[:, None, None, None], (19, n + 2, m + 2, l + 2)).copy()
    pops[q, 1, 1, 20] += sign * 4.5 * d3q19.W[q]
    with Lattice(l, m, n, svisc=0.05, u_inflow=0.0, mode="physics", storage="e3m12") as lat:
        lat.set_mask(np.zeros((n, m, l), np.uint8))
        lat.set_state(pops)
        lat.step(1)
        got = lat.get_state()[:, 1:-1, 1:-1, 1:-1]
    # the site that pulls population q from (x, y, z) = (19, 0, 0) sits at (19, 0, 0) + c_q (periodic)
    bad = np.zeros((n, m, l), bool)
    bad[(0 + int(c[2])) % n, (0 + int(c[1])) % m, 19 + int(c[0])] = True
    assert np.isnan(got[:, bad]).all()
    assert np.isfinite(got[:, ~bad]).all()
