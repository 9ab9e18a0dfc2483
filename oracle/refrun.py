"""Driver for the reference solver built by oracle/Makefile.ref (test infrastructure only).

Runs the reference executable in a scratch directory on a deck, obstacle file and restore dump
written by this repository's own generators, and collects its outputs: the raw save dumps
(CODE/SRC/save_raw.F90:54-74), the drag records (CODE/SRC/drag.F90:67), the probe records
(CODE/SRC/probe_global.F90:77,93) and the MLUPS line of bgk.perf (CODE/SRC/finalize.F90:107-108).

Only tests, the smoke check and the CPU-baseline leg of bench.py use this module.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

from euplectella_hpc_amd import fortio

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def ref_binary(variant="serial"):
    path = os.path.join(REF_DIR, f"bgk3d_{variant}.x")
    return path if os.path.exists(path) else None


def run_reference(l, m, n, svisc, u0, steps, mask, pops0, *, variant="nodrag", itsave=None,
                  icheck=None, ivtim=None, threads=1, workdir=None, keep=False, timeout=3600,
                  irestart=1):
    """Run ``steps`` time steps of the reference from the ghosted state ``pops0`` (irestart = 1),
    or from its own rest state (irestart = 0, init_v = 0: CODE/SRC/init.F90:128-236; ``pops0`` is
    then ignored).

    Returns a dict with ``saves`` ({itime: pops}), ``drag`` (list of (delta, itime, fx, fy)),
    ``probe`` (list of (itime, u, v, w, rho)), ``vtk`` ({file name: bytes} of the binary VTK
    files written every ``ivtim`` steps, CODE/SRC/diagnostic.F90:40-90), ``diagno_text`` (the
    diagno.dat written at a fresh start, CODE/SRC/initialize.F90:145), ``mlups`` and ``stdout``.
    """
    exe = ref_binary(variant)
    if exe is None:
        raise FileNotFoundError(f"reference build {variant!r} missing; run make -f oracle/Makefile.ref")
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="eupref_") if own else fortio.ensure_dir(workdir)
    try:
        big = steps + 10
        fortio.write_deck(os.path.join(wd, "bgk.input"), svisc=float(svisc), u0=float(u0),
                          itfin=int(steps), ivtim=int(ivtim if ivtim else big), isignal=big,
                          itsave=int(itsave if itsave else big), icheck=int(icheck if icheck else big),
                          irestart=int(irestart), init_v=0, lx=l, ly=m, lz=n, proc_x=1, proc_y=1,
                          proc_z=1, tstep=1, flag1=1, flag2=2, flag3=0)
        fortio.write_obs(os.path.join(wd, fortio.obs_file_name()), mask)
        if irestart:
            fortio.write_raw(os.path.join(wd, fortio.restore_file_name()), 0, pops0)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(threads)
        env.setdefault("OMP_PROC_BIND", "close")
        proc = subprocess.run([exe], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"reference exited {proc.returncode}:\n{proc.stdout[-2000:]}\n"
                               f"{proc.stderr[-2000:]}")
        out = {"stdout": proc.stdout, "workdir": wd}
        saves = {}
        for name in sorted(os.listdir(wd)):
            mt = re.match(r"save\.00\.00\.00\.(\d{8})\.bin$", name)
            if mt:
                saves[int(mt.group(1))] = fortio.read_raw(os.path.join(wd, name), (n, m, l))[1]
        out["saves"] = saves
        out["vtk"] = {}
        for name in sorted(os.listdir(wd)):
            if re.match(r"tec_(3d|xy|xz|yz|om)\.\d{4}\.\d{8}\.vtk$", name):
                with open(os.path.join(wd, name), "rb") as fh:
                    out["vtk"][name] = fh.read()
        out["drag"] = _read_table(os.path.join(wd, "drag.0000.dat"), 5)
        out["probe"] = _read_table(os.path.join(wd, "probe_g.0000.dat"), 5)
        out["drag_text"] = _read_text(os.path.join(wd, "drag.0000.dat"))
        out["probe_text"] = _read_text(os.path.join(wd, "probe_g.0000.dat"))
        out["diagno_text"] = _read_text(os.path.join(wd, "diagno.dat"))
        out["mlups"] = _read_mlups(os.path.join(wd, "bgk.perf"))
        return out
    finally:
        if own and not keep:
            shutil.rmtree(wd, ignore_errors=True)


def _read_table(path, ncol):
    rows = []
    if not os.path.exists(path):
        return rows
    with open(path) as fh:
        for line in fh:
            parts = line.replace("D", "E").split()
            if len(parts) >= ncol - 1:
                rows.append(tuple(float(p) for p in parts[:ncol]))
    return rows


def _read_text(path):
    if not os.path.exists(path):
        return ""
    with open(path) as fh:
        return fh.read()


def _read_mlups(path):
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if "Mlups" in line:
                nums = re.findall(r"[-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?|\d+", line)
                if nums:
                    return float(nums[-1].replace("D", "E").replace("d", "e"))
    return None


def time_reference(l, m, n, svisc, u0, steps, mask, *, threads=None, variant="omp", timeout=900):
    """Time ``steps`` steps of the reference's OpenMP build on a lattice started from rest.

    The reference's own initial condition (init_v = 0, CODE/SRC/init.F90:128-236) avoids writing
    a multi-GB restore file; the step cost does not depend on the values.  I/O intervals are set
    beyond the last step so that the loop time holds no output.  Returns the reference's MLUPS
    (bgk.perf, CODE/SRC/finalize.F90:107-108) and the thread count used.
    """
    exe = ref_binary(variant)
    if exe is None:
        raise FileNotFoundError(f"reference build {variant!r} missing")
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    wd = tempfile.mkdtemp(prefix="eupref_t_")
    try:
        big = steps + 10
        fortio.write_deck(os.path.join(wd, "bgk.input"), svisc=float(svisc), u0=float(u0),
                          itfin=int(steps), ivtim=big, isignal=big, itsave=big, icheck=big,
                          irestart=0, init_v=0, lx=l, ly=m, lz=n, proc_x=1, proc_y=1, proc_z=1,
                          tstep=1, flag1=1, flag2=2, flag3=0)
        fortio.write_obs(os.path.join(wd, fortio.obs_file_name()), mask)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(threads)
        env.setdefault("OMP_PROC_BIND", "close")
        proc = subprocess.run([exe], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"reference exited {proc.returncode}: {proc.stdout[-1000:]}")
        return {"mlups": _read_mlups(os.path.join(wd, "bgk.perf")), "threads": threads,
                "steps": steps}
    finally:
        shutil.rmtree(wd, ignore_errors=True)
