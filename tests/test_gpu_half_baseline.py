"""16-bit storage beside the reference's own reduced-precision build, through the Fortran host.

Case: the reference's deck A (CODE/RUN/bgk.A.input: the R = 100 cylinder of the in-code obstacle
generator, u0 = 0.1, svisc = 0.20, Re = 100) shrunk to 680 x 340 x 4, 500 steps from rest, the
control-volume drag of the four -DDRAG boxes every 25 steps (CODE/SRC/drag.F90:48-67).  The
fixture tests/golden/half_baseline.json holds the deck and the drag records written by two builds
of the reference: double precision (the build parity mode reproduces) and HALF_P (populations
stored and collided in binary16; tests/golden/make_half_baseline.py).

* Parity mode writes the double-precision build's drag record character for character.
* The error measure: the largest |F_x - F_x,ref| / |F_x,ref| over the records of steps 400-500 and
  the boxes whose F_x,ref is not zero, with F_x,ref from FP64 of the same code (physics mode in
  FP64 for this library's storages; the double-precision build for the reference's binary16 build).
* FP32, FP16 deviation and e3m12 storage stay below 1e-3 and below 1/100 of the reference's
  binary16 build's error (which is printed).
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from euplectella_hpc_amd import build as eup_build

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "half_baseline.json")) as _fh:
    FIX = json.load(_fh)


def drag_table(text):
    """{(box delta, step): F_x} of a drag.0000.dat record."""
    out = {}
    for line in text.splitlines():
        parts = line.replace("D", "E").split()
        if len(parts) >= 4:
            out[(int(parts[0]), int(parts[1]))] = float(parts[2])
    return out


def rel_error(got, ref, first_step):
    keys = [k for k in ref if k[1] >= first_step and ref[k] != 0.0]
    assert keys and all(k in got for k in keys)
    return max(abs(got[k] - ref[k]) / abs(ref[k]) for k in keys)


def host(env):
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "bgk.input"), "w") as fh:
            fh.write(FIX["deck_text"])
        res = subprocess.run([eup_build.FORTRAN_EXE], cwd=wd, capture_output=True, text=True, timeout=300,
                             env=dict(os.environ, EUP_DRAG="1", EUP_VTK="none", **env))
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        with open(os.path.join(wd, "drag.0000.dat")) as fh:
            return fh.read()


def test_parity_mode_writes_the_double_precision_builds_drag():
    print("[gpu] Fortran host, parity mode, deck A 680x340x4, 500 steps", flush=True)
    assert host({"EUP_MODE": "parity"}) == FIX["serial"]["drag_text"]


def test_16_bit_storage_beside_the_references_binary16_build():
    first = FIX["steps"] - 100
    ref_half = rel_error(drag_table(FIX["halforig"]["drag_text"]), drag_table(FIX["serial"]["drag_text"]), first)
    print(f"[gpu] reference binary16 build vs its double-precision build: {ref_half:.3e}", flush=True)
    f64 = drag_table(host({"EUP_MODE": "physics", "EUP_STORAGE": "f64"}))
    errs = {}
    for st in ("f32", "f16", "e3m12"):
        print(f"[gpu] Fortran host, physics mode {st}, deck A 680x340x4, 500 steps", flush=True)
        errs[st] = rel_error(drag_table(host({"EUP_MODE": "physics", "EUP_STORAGE": st})), f64, first)
        print(f"[gpu]   {st} vs FP64: {errs[st]:.3e}", flush=True)
    for st, e in errs.items():
        assert e < 1e-3 and e < ref_half / 100, (st, e, ref_half)
