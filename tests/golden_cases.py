"""Helpers shared by the parity tests: rebuild a golden case's inputs and hash states."""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
sys.path.insert(0, ROOT)

from make_golden import CASES, make_inputs  # noqa: E402,F401


def sha(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def omega_of(rec):
    """hencol.F90:30: omega = 2/(6*svisc + 1) in double, svisc as written in the deck."""
    return 2.0 / (6.0 * rec["svisc"] + 1.0)


def e14_6(x):
    """Value of ``x`` after Fortran E14.6 output (6 significant digits)."""
    return float(f"{x:.5e}")
