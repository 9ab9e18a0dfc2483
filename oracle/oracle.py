"""ctypes binding of the C oracle (oracle/eup_oracle.c) -- test infrastructure only.

Arrays follow the reference layout: populations ``pops[19, n+2, m+2, l+2]`` float64 (ghosted,
i fastest) and the solid mask ``mask[n, m, l]`` (nonzero = solid), converted here to the int32
``obs`` array of the reference.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def load(path):
    """A build of eup_oracle.c at ``path`` with its argument types set (the tests also load copies
    with a planted mistake, to show that a pin catches it)."""
    L = ctypes.CDLL(path)
    i, d = ctypes.c_int, ctypes.c_double
    L.orc_parity_steps.argtypes = [_dp, _dp, _ip, i, i, i, d, d, d, i]
    L.orc_physics_steps.argtypes = [_dp, _dp, _ip, i, i, i, d, d, d, i, i]
    L.orc_phase_bcond_obs.argtypes = [_dp, _ip, i, i, i]
    L.orc_phase_bcond_bc.argtypes = [_dp, i, i, i, d]
    L.orc_phase_movef.argtypes = [_dp, _dp, i, i, i]
    L.orc_phase_col.argtypes = [_dp, _dp, i, i, i, d, d]
    L.orc_collide_site.argtypes = [_dp, _dp, d, d]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.orc_vtk_velocity.argtypes = [_dp, i, i, i, i, i, _fp]
    L.orc_vtk_vorticity.argtypes = [_dp, i, i, i, i, _fp]
    L.orc_flux_x.argtypes = [_dp, i, i, i, d, _ip, _dp]
    L.orc_flux_y.argtypes = [_dp, i, i, i, d, _ip, _dp]
    L.orc_probe.argtypes = [_dp, i, i, i, i, i, i, _dp]
    L.orc_probe.restype = ctypes.c_int
    L.orc_diagno.argtypes = [_dp, _ip, i, i, i, _dp]
    L.orc_mem_force.argtypes = [_dp, _ip, i, i, i, i, _dp]
    L.orc_grey_steps.argtypes = [_dp, _dp, i, i, i, d, d, i, i]
    L.orc_grey_force.argtypes = [_dp, _dp, i, i, i, d, i, _dp]
    L.orc_grey_objective.argtypes = [_dp, _dp, i, i, i, d, d, i, i, i, i, d, _dp, _dp]
    L.orc_collide_adj.argtypes = [_dp, _dp, d, _dp]
    L.orc_grey_step_adj.argtypes = [_dp, _dp, i, i, i, d, d, i, _dp, _dp, d, _dp]
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _obs(mask):
    return np.ascontiguousarray(mask, dtype=np.int32)


def _shape(pops):
    _, n2, m2, l2 = pops.shape
    return l2 - 2, m2 - 2, n2 - 2


def omega_from_svisc(svisc):
    """hencol.F90:30: omega = 2/(6*svisc + 1) in double."""
    return 2.0 / (6.0 * float(svisc) + 1.0)


def parity_steps(pops, mask, omega, u_inflow, nsteps, fgrad=0.0):
    """Advance ``pops`` (modified in place and returned) by ``nsteps`` serial reference steps."""
    assert pops.dtype == np.float64 and pops.flags.c_contiguous
    l, m, n = _shape(pops)
    obs = _obs(mask)
    scratch = np.empty_like(pops)
    lib().orc_parity_steps(_d(pops), _d(scratch), obs.ctypes.data_as(_ip), l, m, n, omega, fgrad,
                           u_inflow, int(nsteps))
    return pops


def physics_steps(pops, mask, omega, u_inflow, nsteps, fgrad=0.0, x_open=True, L=None):
    assert pops.dtype == np.float64 and pops.flags.c_contiguous
    l, m, n = _shape(pops)
    obs = _obs(mask)
    scratch = np.zeros_like(pops)
    (L or lib()).orc_physics_steps(_d(pops), _d(scratch), obs.ctypes.data_as(_ip), l, m, n, omega, fgrad,
                            u_inflow, int(bool(x_open)), int(nsteps))
    return pops


def phase_bcond_obs(pops, mask):
    l, m, n = _shape(pops)
    lib().orc_phase_bcond_obs(_d(pops), _obs(mask).ctypes.data_as(_ip), l, m, n)


def phase_bcond_bc(pops, u_inflow):
    l, m, n = _shape(pops)
    lib().orc_phase_bcond_bc(_d(pops), l, m, n, u_inflow)


def phase_movef(pops, out):
    l, m, n = _shape(pops)
    lib().orc_phase_movef(_d(pops), _d(out), l, m, n)


def phase_col(pops, b, omega, fgrad=0.0):
    l, m, n = _shape(pops)
    lib().orc_phase_col(_d(pops), _d(b), l, m, n, omega, fgrad)


def collide_site(x, omega, fgrad=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(19)
    lib().orc_collide_site(_d(x), _d(out), omega, fgrad)
    return out


def flux_x(pops, omega, box):
    l, m, n = _shape(pops)
    b = np.ascontiguousarray(box, dtype=np.int32)
    out = np.zeros(1)
    lib().orc_flux_x(_d(pops), l, m, n, omega, b.ctypes.data_as(_ip), _d(out))
    return float(out[0])


def flux_y(pops, omega, box):
    l, m, n = _shape(pops)
    b = np.ascontiguousarray(box, dtype=np.int32)
    out = np.zeros(1)
    lib().orc_flux_y(_d(pops), l, m, n, omega, b.ctypes.data_as(_ip), _d(out))
    return float(out[0])


def probe(pops, i, j, k):
    l, m, n = _shape(pops)
    out = np.zeros(4)
    rc = lib().orc_probe(_d(pops), l, m, n, i, j, k, _d(out))
    return None if rc else out


def diagno(pops, mask):
    l, m, n = _shape(pops)
    out = np.zeros(5)
    lib().orc_diagno(_d(pops), _obs(mask).ctypes.data_as(_ip), l, m, n, _d(out))
    return out


VTK_YZ, VTK_XZ, VTK_XY, VTK_3D = 0, 1, 2, 3


def vtk_velocity(pops, plane, index=0):
    """Velocity of a plane (VTK_YZ / VTK_XZ / VTK_XY at the 1-based ``index``) or of the whole
    lattice (VTK_3D) as the reference's binary VTK writers compute it: float32 ``[..., 3]``."""
    l, m, n = _shape(pops)
    count = {VTK_YZ: m * n, VTK_XZ: l * n, VTK_XY: l * m, VTK_3D: l * m * n}[plane]
    out = np.zeros((count, 3), np.float32)
    lib().orc_vtk_velocity(_d(pops), l, m, n, int(plane), int(index),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def vtk_vorticity(pops, k0):
    """Vorticity of the xy plane z = k0 as the reference's vtk_om_bin computes it:
    float32 ``[(m-2) * (l-2)]``, i fastest."""
    l, m, n = _shape(pops)
    out = np.zeros((m - 2) * (l - 2), np.float32)
    lib().orc_vtk_vorticity(_d(pops), l, m, n, int(k0),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def mem_force(pops, mask, x_open=True):
    l, m, n = _shape(pops)
    out = np.zeros(3)
    lib().orc_mem_force(_d(pops), _obs(mask).ctypes.data_as(_ip), l, m, n, int(bool(x_open)),
                        _d(out))
    return out


def drag_boxes(lx, ly, lz, radius=100, deltas=(52, 57, 62, 67)):
    """Control boxes of the reference drag diagnostic (CODE/SRC/drag.F90:48-58)."""
    boxes = []
    for delta in deltas:
        boxes.append((delta, (lx // 4 - radius - delta, lx // 4 + radius + delta,
                              ly // 2 - radius - delta, ly // 2 + radius + delta, 1, lz)))
    return boxes


# --------------------------------------------------------------------------------------------
# grey lattice: dense interior layout a[19, n, m, l], s[n, m, l]
# --------------------------------------------------------------------------------------------

def _dense_shape(a):
    _, n, m, l = a.shape
    return l, m, n


def grey_steps(a, s, omega, u_in, nsteps, x_open=True, L=None):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    l, m, n = _dense_shape(a)
    s = np.ascontiguousarray(s, dtype=np.float64)
    rc = (L or lib()).orc_grey_steps(_d(a), _d(s), l, m, n, omega, u_in, int(bool(x_open)), int(nsteps))
    assert rc == 0
    return a


def grey_force(a, s, u_in, x_open=True):
    l, m, n = _dense_shape(a)
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_grey_force(_d(a), _d(s), l, m, n, u_in, int(bool(x_open)), _d(out))
    return out


def grey_objective(a0, s, omega, u_in, t0, t1, K, norm, grad=True, x_open=True):
    """J = mean over t0..t1 of F_x(a^t)/norm from a0; returns (J, dJ/ds or None)."""
    a0 = np.ascontiguousarray(a0, dtype=np.float64)
    l, m, n = _dense_shape(a0)
    s = np.ascontiguousarray(s, dtype=np.float64)
    J = np.zeros(1)
    gs = np.zeros((n, m, l)) if grad else None
    rc = lib().orc_grey_objective(_d(a0), _d(s), l, m, n, omega, u_in, int(bool(x_open)), t0, t1, K,
                                  norm, _d(J), _d(gs) if grad else ctypes.POINTER(ctypes.c_double)())
    assert rc == 0
    return float(J[0]), gs


def collide_adj(f, lam, omega):
    f = np.ascontiguousarray(f, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.empty(19)
    lib().orc_collide_adj(_d(f), _d(lam), omega, _d(out))
    return out


def grey_step_adj(a, s, omega, u_in, lam_in, wobj=0.0, x_open=True):
    l, m, n = _dense_shape(a)
    a = np.ascontiguousarray(a, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    lam_in = np.ascontiguousarray(lam_in, dtype=np.float64)
    lam_out = np.empty_like(lam_in)
    gs = np.zeros((n, m, l))
    lib().orc_grey_step_adj(_d(a), _d(s), l, m, n, omega, u_in, int(bool(x_open)), _d(lam_in),
                            _d(lam_out), wobj, _d(gs))
    return lam_out, gs
