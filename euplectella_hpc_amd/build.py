"""Build libeup.so (HIP, gfx950) in-tree with hipcc.

Compile flags:
  --offload-arch=gfx950       MI355X only
  -ffp-contract=off           no fused multiply-add: the FP64 paths are bit-identical to the
                              reference's serial build
  -Xclang -target-feature -Xclang -packed-fp32-ops
                              no packed FP32 instructions in the generated code
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["eup_api.hip", "eup_parity.hip", "eup_physics.hip", "eup_diag.hip", "eup_grey.hip",
           "eup_membw.hip"]
OUT = os.path.join(HERE, "libeup.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_EXE = os.path.join(FORTRAN_DIR, "bgk3d_eup.x")


def build_fortran(verbose=False):
    """Fortran host (ISO_C_BINDING) linked against libeup.so; needs amdflang."""
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        if verbose:
            print("amdflang not found: Fortran host not built")
        return None
    cmd = [fc, "-O2", "-module-dir", FORTRAN_DIR,
           os.path.join(FORTRAN_DIR, "eup_mod.F90"), os.path.join(FORTRAN_DIR, "bgk3d_eup.F90"),
           "-L" + HERE, "-leup", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_EXE]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("amdflang failed building the Fortran host")
    if verbose:
        print("built", FORTRAN_EXE)
    return FORTRAN_EXE


def build(verbose=False, out=None, defines=(), fortran=True):
    """Compile libeup.so; ``out``/``defines`` build a variant elsewhere (kernel experiments)."""
    target = out or OUT
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + \
        ["-I" + os.path.join(ROOT, "include")] + srcs + ["-o", target + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libeup.so")
    os.replace(target + ".tmp", target)
    if verbose:
        print("built", target)
    if fortran and out is None:
        build_fortran(verbose)
    return target


if __name__ == "__main__":
    build(verbose=True)
