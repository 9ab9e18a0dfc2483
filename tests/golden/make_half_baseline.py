"""Reduced-precision baseline of the reference itself: deck A (CODE/RUN/bgk.A.input: the R = 100
cylinder of the in-code generator, u0 = 0.1, svisc = 0.20, i.e. Re = 100) shrunk to 680 x 340 x 4,
500 steps from rest, control-volume drag every 25 steps (-DDRAG, CODE/SRC/drag.F90:48-67), run by
two builds of the reference (oracle/Makefile.ref):

  * "serial": the double-precision build (DOUBLE_P, ORIGINAL) -- the build that parity mode
    reproduces bit for bit;
  * "halforig": the same path with HALF_P and MIXEDPRECISION (CODE/SRC/precision.F90:9-12,23-27):
    populations stored as real*2 and, on the ORIGINAL path, collided in real*2 as well
    (col.F90 declares its temporaries real(mykind); MIXEDPRECISION's real(4) temporaries exist
    only in col_MC*).

The reference's real(4)-compute path (col_MC, no ORIGINAL: CODE/SRC/collision.F90:99-126) is not
usable on this case: built with the same case macros it diverges at the open x faces within five
steps, in double precision as well (density 1e68 at x = 1 and x = l - 2 after 5 steps).

Writes tests/golden/half_baseline.json: the deck text, the two drag records as written, and the
builds' MLUPS on one core.  Usage (where /root/reference is present):
    make -f oracle/Makefile.ref serial halforig && python tests/golden/make_half_baseline.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from make_golden_decks import RUN, edit_deck  # noqa: E402
from oracle import refrun  # noqa: E402

LATTICE = (680, 340, 4)
STEPS, EVERY = 500, 25
RUNLEN = dict(itfin=STEPS, ivtim=10 * STEPS, isignal=10 * STEPS, itsave=10 * STEPS, icheck=EVERY)


def run(variant, deck):
    wd = tempfile.mkdtemp(prefix=f"euphalf_{variant}_")
    try:
        with open(os.path.join(wd, "bgk.input"), "w") as fh:
            fh.write(deck)
        res = subprocess.run([refrun.ref_binary(variant)], cwd=wd, capture_output=True, text=True,
                             timeout=7200, env=dict(os.environ, OMP_NUM_THREADS="1"))
        if res.returncode != 0:
            raise RuntimeError(f"{variant}: reference exited {res.returncode}\n{res.stdout[-2000:]}")
        with open(os.path.join(wd, "drag.0000.dat")) as fh:
            drag = fh.read()
        return {"drag_text": drag, "mlups": refrun._read_mlups(os.path.join(wd, "bgk.perf"))}
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def main():
    with open(os.path.join(RUN, "bgk.A.input")) as fh:
        deck = edit_deck(fh.read(), LATTICE, runlen=RUNLEN)
    out = {"deck_text": deck, "l": LATTICE[0], "m": LATTICE[1], "n": LATTICE[2], "steps": STEPS,
           "every": EVERY}
    for variant in ("serial", "halforig"):
        out[variant] = run(variant, deck)
        print(variant, "done", out[variant]["mlups"], "MLUPS", flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "half_baseline.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
