"""Golden fixtures for the Fortran host from the reference's own run decks (CODE/RUN/bgk.*.input).

The reference's decks target 4000 x 2000 x 400 lattices on up to 512 MPI tasks and assign lists
to scalar fields (svisc, itfin, irestart: one value per run of the Re chain of CODE/RUN/README).
The reference executable built here rejects those lists ("NAMELIST input group 'parameters' was
not terminated"), so each deck is taken with:
  * lx, ly, lz shrunk to a lattice that runs in seconds and still holds the in-code obstacle
    (and, with -DDRAG, the hard-wired drag boxes of R = 100);
  * every list line reduced to its first value (the first run of the chain: svisc = 0.20,
    irestart = 0, a fresh start from rest);
  * itfin, ivtim, isignal, itsave, icheck set to a short run;
every other line -- proc_x/y/z (reset to 1 by the serial build), u0, tstep, flag1/2/3, init_v,
the layout and spacing -- is kept as the deck has it.  Decks C and D read an obstacle list
(flag1 = 1), written here from this repository's cylinder generator.  Further cases pin the
other in-code obstacles (flag2 = 1 spherical shell, 3 half column) and the obstacle dumps
(flag3 = 1 ASCII VTK, 2 list).

The edited deck text is stored as input data; the reference's outputs as SHA-256 digests (save
dumps, obstacle dumps) or text (drag, probe, diagno records; the labels of bgk.perf and the step
column of fort.100).

Usage (in a container where /root/reference is present):
    make -f oracle/Makefile.ref && python tests/golden/make_golden_decks.py
"""
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from euplectella_hpc_amd import fortio, geometry  # noqa: E402
from oracle import refrun  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
RUN = "/root/reference/CODE/RUN"

RUNLEN = dict(itfin=6, ivtim=1000, isignal=3, itsave=6, icheck=2)

# case: source deck, lattice, extra fields (added or replacing a line), reference build, obstacle file
CASES = {
    "deck_A": dict(deck="bgk.A.input", lattice=(680, 340, 4), variant="serial"),
    "deck_B": dict(deck="bgk.B.input", lattice=(680, 340, 12), variant="serial"),
    "deck_C": dict(deck="bgk.C.input", lattice=(680, 340, 4), variant="serial",
                   obs=("cylinder", dict(radius=40, centre=(170, 170)))),
    "deck_D": dict(deck="bgk.D.input", lattice=(680, 340, 8), variant="serial",
                   obs=("cylinder", dict(radius=30, centre=(160, 172)))),
    "deck_S": dict(deck="bgk.S.input", lattice=(700, 340, 16), variant="serial"),
    # in-code obstacles and dumps on smaller lattices (no -DDRAG: its boxes need R = 100)
    "deck_A_sphere_vtk": dict(deck="bgk.A.input", lattice=(96, 64, 48), variant="nodrag",
                              extra=dict(flag2=1, flag3=1)),
    "deck_A_halfcol_list": dict(deck="bgk.A.input", lattice=(96, 64, 72), variant="nodrag",
                                extra=dict(flag2=3, flag3=2)),
    "deck_A_cyl_list": dict(deck="bgk.A.input", lattice=(680, 340, 4), variant="serial",
                            extra=dict(flag3=2)),
}


def edit_deck(text, lattice, extra=None, runlen=None):
    """The deck with lx/ly/lz and the run length replaced, lists reduced to their first value."""
    lx, ly, lz = lattice
    fields = dict(lx=lx, ly=ly, lz=lz, **(runlen or RUNLEN), **(extra or {}))
    out, seen = [], set()
    for line in text.splitlines():
        m = re.match(r"^(\s*)(\w+)(\s*=\s*)([^/]*?)(\s*/?\s*)$", line)
        if m and m.group(2).lower() != "parameters":
            key = m.group(2).lower()
            val = m.group(4)
            if key in fields:
                val = str(fields[key])
            elif "," in val:
                val = val.split(",")[0]
            seen.add(key)
            line = f"{m.group(1)}{m.group(2)}{m.group(3)}{val}{m.group(5)}"
        out.append(line)
    # fields the deck does not set yet go before its closing line
    missing = [f"{k}={v}" for k, v in fields.items() if k not in seen]
    if missing:
        out[-1:-1] = missing
    return "\n".join(out) + "\n"


def perf_labels(text):
    """bgk.perf without its numbers (and the width they take): the line labels and their order."""
    num = r"(?<![\w.])[-+]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?(?![\w.])"
    return [" ".join(re.sub(num, "#", ln).split()) for ln in text.splitlines()]


def run_case(name, case):
    with open(os.path.join(RUN, case["deck"])) as fh:
        deck = edit_deck(fh.read(), case["lattice"], case.get("extra"))
    l, m, n = case["lattice"]
    wd = tempfile.mkdtemp(prefix=f"eupdeck_{name}_")
    try:
        with open(os.path.join(wd, "bgk.input"), "w") as fh:
            fh.write(deck)
        rec = {"deck_text": deck, "l": l, "m": m, "n": n, "variant": case["variant"]}
        if "obs" in case:
            kind, kw = case["obs"]
            mask = getattr(geometry, kind)(l, m, n, **kw)
            fortio.write_obs(os.path.join(wd, fortio.obs_file_name()), mask)
            rec["obs"] = [kind, kw]
            rec["obs_sha256"] = hashlib.sha256(mask.astype(np.int32).tobytes()).hexdigest()
        exe = refrun.ref_binary(case["variant"])
        res = subprocess.run([exe], cwd=wd, capture_output=True, text=True, timeout=1800,
                             env=dict(os.environ, OMP_NUM_THREADS="1"))
        if res.returncode != 0:
            raise RuntimeError(f"{name}: reference exited {res.returncode}\n{res.stdout[-2000:]}")
        files = sorted(os.listdir(wd))
        save = fortio.save_file_name(RUNLEN["itfin"])
        _, state = fortio.read_raw(os.path.join(wd, save), (n, m, l))
        rec["state_sha256"] = hashlib.sha256(np.ascontiguousarray(state).tobytes()).hexdigest()
        for key, fn in (("drag_text", "drag.0000.dat"), ("probe_text", "probe_g.0000.dat"),
                        ("diagno_text", "diagno.dat")):
            with open(os.path.join(wd, fn)) as fh:
                rec[key] = fh.read()
        with open(os.path.join(wd, "bgk.perf")) as fh:
            rec["perf_labels"] = perf_labels(fh.read())
        with open(os.path.join(wd, "fort.100")) as fh:
            rec["profile_steps"] = [int(r.split()[0]) for r in fh.read().splitlines() if r.strip()]
        for fn in ("tec_ob.0000.00000000.vtk", "obs.0000.00000000.bin"):
            if fn in files and not (fn.startswith("obs.") and "obs" in case):
                with open(os.path.join(wd, fn), "rb") as fh:
                    data = fh.read()
                rec.setdefault("dumps", {})[fn] = {"sha256": hashlib.sha256(data).hexdigest(),
                                                   "bytes": len(data)}
        nobs = re.search(r"num\. obs\s+-->\s*(-?\d+)", res.stdout)
        rec["nobs_reported"] = int(nobs.group(1)) if nobs else None
        return rec
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def main(names=None):
    path = os.path.join(OUT, "golden_decks.json")
    index = {}
    if names and os.path.exists(path):
        with open(path) as fh:
            index = json.load(fh)
    for name, case in CASES.items():
        if names and name not in names:
            continue
        index[name] = run_case(name, case)
        print(name, "done", index[name]["state_sha256"][:16], "nobs", index[name]["nobs_reported"],
              flush=True)
    with open(path, "w") as fh:
        json.dump(index, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
