"""ctypes binding of libeup.so (the C-ABI in include/eup.h) and a thin Python host.

``Lattice`` mirrors the reference's per-run flow: build the boundary flags and collision
parameters (build_bcond + hencol), read the obstacle mask (build_obs), restore the populations
(restore_raw), then call the time step (boundaries / propagation / collision) and the diagnostics
(drag, probe_global, diagno).  Every compute call goes to the HIP library; there is no CPU
fallback: if ``libeup.so`` is missing the import of this module's functions raises.
"""
import ctypes
import os

import numpy as np

from . import fortio
from .fortio import VTK_3D, VTK_XY, VTK_XZ, VTK_YZ

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EUP_LIBRARY") or os.path.join(HERE, "libeup.so")

EUP_OK = 0
ERRORS = {-1: "EUP_EINVAL", -2: "EUP_EHIP", -3: "EUP_ESTATE", -4: "EUP_ERANGE"}

STORE_F64, STORE_F32, STORE_F16, STORE_BF16, STORE_E3M12 = 0, 1, 2, 3, 4
STORAGE = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16,
           "e3m12": STORE_E3M12}
MODE_PARITY, MODE_PHYSICS, MODE_GREY = 0, 1, 2
MODES = {"parity": MODE_PARITY, "physics": MODE_PHYSICS, "grey": MODE_GREY}
FACE_PERIODIC, FACE_INFLOW, FACE_OUTFLOW = 0, 4, 5
FLAG_NO_TILE_SKIP = 1
FLAG_SHIFTED_DEVIATION = 2

#: every function the header declares (checked by the CPU tests)
EXPORTS = ("eup_create", "eup_destroy", "eup_last_error", "eup_set_mask", "eup_set_pops",
           "eup_get_pops", "eup_set_state", "eup_get_state", "eup_step", "eup_step_timed",
           "eup_sync", "eup_drag_cv", "eup_drag_mem", "eup_probe", "eup_diagno", "eup_vtk_velocity",
           "eup_vtk_vorticity", "eup_get_info",
           "eup_stream", "eup_abi_version", "eup_set_grey", "eup_set_grey_device",
           "eup_grey_force", "eup_grey_objective", "eup_grey_objective_device", "eup_membw",
           "eup_init_equilibrium", "eup_set_omega", "eup_add_obstacles", "eup_commit_obstacles",
           "eup_drag_cv_noalloc")


class EupConfig(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int32), ("m", ctypes.c_int32), ("n", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("front", ctypes.c_int32), ("rear", ctypes.c_int32),
                ("up", ctypes.c_int32), ("down", ctypes.c_int32),
                ("left", ctypes.c_int32), ("right", ctypes.c_int32),
                ("device", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("omega", ctypes.c_double), ("fgrad", ctypes.c_double),
                ("u_inflow", ctypes.c_double)]


class EupInfo(ctypes.Structure):
    _fields_ = [("sites", ctypes.c_int64), ("nsolid", ctypes.c_int64),
                ("tiles", ctypes.c_int64), ("tiles_fluid", ctypes.c_int64),
                ("tiles_mixed", ctypes.c_int64), ("tiles_solid", ctypes.c_int64),
                ("bbox", ctypes.c_int32 * 6), ("bytes_per_site", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64)]


class EupError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libeup.so (raises FileNotFoundError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    # One HIP runtime per process: PyTorch-ROCm bundles its own libamdhip64.so.7.  Importing torch
    # first makes the dynamic linker bind libeup's libamdhip64.so.7 to that already-loaded copy,
    # so library kernels and torch tensors share one runtime (device pointers, streams).
    if not os.environ.get("EUP_NO_TORCH"):
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    i32 = ctypes.c_int32
    L.eup_create.argtypes = [ctypes.POINTER(EupConfig), ctypes.POINTER(vp)]
    L.eup_destroy.argtypes = [vp]
    L.eup_destroy.restype = None
    L.eup_last_error.argtypes = [vp]
    L.eup_last_error.restype = ctypes.c_char_p
    L.eup_set_mask.argtypes = [vp, ip]
    L.eup_add_obstacles.argtypes = [vp, ip, ctypes.c_int64]
    L.eup_commit_obstacles.argtypes = [vp]
    L.eup_set_pops.argtypes = [vp, i32, dp]
    L.eup_get_pops.argtypes = [vp, i32, dp]
    L.eup_set_state.argtypes = [vp, dp]
    L.eup_get_state.argtypes = [vp, dp]
    L.eup_step.argtypes = [vp, i32]
    L.eup_step_timed.argtypes = [vp, i32, dp]
    L.eup_sync.argtypes = [vp]
    L.eup_drag_cv.argtypes = [vp, ip, dp, dp]
    L.eup_drag_cv_noalloc.argtypes = [vp, ip, dp, dp]
    L.eup_drag_mem.argtypes = [vp, dp]
    L.eup_probe.argtypes = [vp, i32, i32, i32, dp]
    L.eup_diagno.argtypes = [vp, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.eup_vtk_velocity.argtypes = [vp, i32, i32, fp]
    L.eup_vtk_vorticity.argtypes = [vp, i32, fp]
    L.eup_get_info.argtypes = [vp, ctypes.POINTER(EupInfo)]
    L.eup_stream.argtypes = [vp]
    L.eup_stream.restype = vp
    L.eup_abi_version.argtypes = []
    L.eup_abi_version.restype = i32
    L.eup_membw.argtypes = [i32, ctypes.c_int64, i32, dp]
    L.eup_set_omega.argtypes = [vp, ctypes.c_double]
    L.eup_init_equilibrium.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_uint64]
    L.eup_set_grey.argtypes = [vp, dp]
    L.eup_set_grey_device.argtypes = [vp, vp, i32]
    L.eup_grey_force.argtypes = [vp, dp]
    L.eup_grey_objective.argtypes = [vp, i32, i32, i32, ctypes.c_double, dp, dp]
    L.eup_grey_objective_device.argtypes = [vp, i32, i32, i32, ctypes.c_double, dp, vp, i32]
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def omega_from_svisc(svisc):
    """hencol.F90:30: omega = 2/(6*svisc + 1), in double."""
    return 2.0 / (6.0 * float(svisc) + 1.0)


def svisc_from_tau(tau):
    """Kinematic viscosity (2 tau - 1)/6 as a deck would carry it."""
    return (2.0 * float(tau) - 1.0) / 6.0


class Lattice:
    """One D3Q19 lattice on one GPU (a libeup handle).

    ``mode`` is "parity" (the reference's serial FP64 step, bit for bit) or "physics" (fused
    pull kernel, ``storage`` f64 / f32 / f16 / bf16 / e3m12).  ``x_open`` selects inlet/outlet x faces
    (the CYLINDER_INF build) or a periodic x axis.
    """

    def __init__(self, l, m, n, *, svisc, u_inflow, mode="physics", storage="f32", x_open=True,
                 fgrad=0.0, device=0, tile_skip=True, shifted=False):
        self.L = load()
        self.shape = (l, m, n)
        self.svisc = float(svisc)
        self.omega = omega_from_svisc(svisc)
        self.u_inflow = float(u_inflow)
        cfg = EupConfig()
        cfg.l, cfg.m, cfg.n = l, m, n
        cfg.mode = MODES[mode]
        cfg.storage = STORAGE[storage]
        cfg.front = FACE_OUTFLOW if x_open else FACE_PERIODIC
        cfg.rear = FACE_INFLOW if x_open else FACE_PERIODIC
        cfg.up = cfg.down = cfg.left = cfg.right = FACE_PERIODIC
        cfg.device = device
        cfg.flags = (0 if tile_skip else FLAG_NO_TILE_SKIP) | (FLAG_SHIFTED_DEVIATION if shifted else 0)
        cfg.omega = self.omega
        cfg.fgrad = float(fgrad)
        cfg.u_inflow = self.u_inflow
        self.cfg = cfg
        h = ctypes.c_void_p()
        rc = self.L.eup_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != EUP_OK:
            raise EupError(rc, self.L.eup_last_error(None).decode())
        self.h = h

    # -- plumbing -------------------------------------------------------------------------------
    def _check(self, rc):
        if rc != EUP_OK:
            raise EupError(rc, self.L.eup_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.eup_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ----------------------------------------------------------------------------------
    def set_mask(self, mask):
        """mask[k, j, i] (n, m, l), nonzero = solid."""
        l, m, n = self.shape
        obs = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.int32)
        if obs.shape != (n, m, l):
            raise ValueError(f"mask shape {obs.shape} != {(n, m, l)}")
        self._check(self.L.eup_set_mask(self.h, obs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    def add_obstacles(self, ijk):
        """Mark the sites ``ijk`` (int array [count, 3] of 1-based (i, j, k), the records of the
        obstacle file) solid; call commit_obstacles() when the list is complete
        (eup_add_obstacles)."""
        a = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        self._check(self.L.eup_add_obstacles(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             int(a.shape[0])))

    def commit_obstacles(self):
        """Finish the obstacle list given to add_obstacles (eup_commit_obstacles)."""
        self._check(self.L.eup_commit_obstacles(self.h))

    def set_obstacle_file(self, path, chunk=1 << 20):
        """Stream an obstacle file (obs.RRRR.TTTTTTTT.bin) into the lattice, ``chunk`` records at
        a time, without a dense mask on the host."""
        for block in fortio.iter_obs(path, chunk):
            self.add_obstacles(block)
        self.commit_obstacles()

    def set_obstacle_dense_file(self, path, chunk=1 << 20):
        """Stream a dense obstacle field (restore.obs.bin, the HDF5-free form of the reference's
        restore.obs.h5) into the lattice one z plane at a time: its solid sites go through
        eup_add_obstacles, so no dense mask is held on the host.  Returns the file's time step."""
        it = fortio.iter_obs_dense(path, self.shape, chunk)
        while True:
            try:
                block = next(it)
            except StopIteration as stop:
                itime = stop.value
                break
            self.add_obstacles(block)
        self.commit_obstacles()
        return itime

    def set_state(self, pops):
        l, m, n = self.shape
        p = np.ascontiguousarray(pops, dtype=np.float64)
        if p.shape != (19, n + 2, m + 2, l + 2):
            raise ValueError("pops must have shape (19, n+2, m+2, l+2)")
        self._check(self.L.eup_set_state(self.h, _dptr(p)))

    def set_svisc(self, svisc):
        """Continue with a new viscosity (the reference's restart with a new deck svisc)."""
        self.svisc = float(svisc)
        self.omega = omega_from_svisc(svisc)
        self._check(self.L.eup_set_omega(self.h, self.omega))

    def init_equilibrium(self, ux=0.0, uy=0.0, uz=0.0, noise=0.0, seed=0):
        """Equilibrium start on the device: u = (ux, uy + noise*U[-1,1), uz) per site."""
        self._check(self.L.eup_init_equilibrium(self.h, float(ux), float(uy), float(uz),
                                                float(noise), int(seed)))

    def set_pops(self, q, arr):
        """Population q (1..19) as a ghosted ``[n+2, m+2, l+2]`` float64 array."""
        l, m, n = self.shape
        p = np.ascontiguousarray(arr, dtype=np.float64)
        if p.shape != (n + 2, m + 2, l + 2):
            raise ValueError("population must have shape (n+2, m+2, l+2)")
        self._check(self.L.eup_set_pops(self.h, int(q), _dptr(p)))

    def get_pops(self, q):
        l, m, n = self.shape
        out = np.empty((n + 2, m + 2, l + 2), dtype=np.float64)
        self._check(self.L.eup_get_pops(self.h, int(q), _dptr(out)))
        return out

    def get_state(self):
        l, m, n = self.shape
        out = np.empty((19, n + 2, m + 2, l + 2), dtype=np.float64)
        self._check(self.L.eup_get_state(self.h, _dptr(out)))
        return out

    # -- stepping -------------------------------------------------------------------------------
    def step(self, nsteps=1):
        self._check(self.L.eup_step(self.h, int(nsteps)))

    def step_timed(self, nsteps):
        ms = ctypes.c_double()
        self._check(self.L.eup_step_timed(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._check(self.L.eup_sync(self.h))

    # -- diagnostics ----------------------------------------------------------------------------
    def drag_cv(self, box, noalloc=False):
        """Control-volume drag (fluxX, fluxY) on ``box``; ``noalloc`` calls eup_drag_cv_noalloc."""
        b = np.ascontiguousarray(box, dtype=np.int32)
        if b.shape != (6,):
            raise ValueError(f"box must hold 6 indices (xstart, xstop, ystart, ystop, zstart, zstop), "
                             f"got shape {b.shape}")
        fx, fy = ctypes.c_double(), ctypes.c_double()
        fn = self.L.eup_drag_cv_noalloc if noalloc else self.L.eup_drag_cv
        self._check(fn(self.h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                       ctypes.byref(fx), ctypes.byref(fy)))
        return fx.value, fy.value

    def drag_mem(self):
        F = np.zeros(3)
        self._check(self.L.eup_drag_mem(self.h, _dptr(F)))
        return F

    def vtk_velocity(self, plane, index=0):
        """Velocity of a plane (VTK_YZ / VTK_XZ / VTK_XY at the 1-based ``index``) or of the
        lattice (VTK_3D) as the reference's binary VTK writers compute it: float32 [sites, 3]
        in the writer's loop order (eup_vtk_velocity)."""
        l, m, n = self.shape
        count = {VTK_YZ: m * n, VTK_XZ: l * n, VTK_XY: l * m, VTK_3D: l * m * n}[plane]
        out = np.empty((count, 3), np.float32)
        self._check(self.L.eup_vtk_velocity(self.h, int(plane), int(index),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def vtk_vorticity(self, k0):
        """Vorticity of the xy plane z = k0 as the reference's vtk_om_bin computes it:
        float32 [(m-2) * (l-2)], i fastest (eup_vtk_vorticity)."""
        l, m, n = self.shape
        out = np.empty((m - 2) * (l - 2), np.float32)
        self._check(self.L.eup_vtk_vorticity(self.h, int(k0),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def write_vtk(self, itime, directory=".", branch="default", rank=0):
        """Write the binary VTK files the reference writes every ``ivtim`` steps
        (CODE/SRC/diagnostic.F90:75-90): branch "default" = velocity on the xy plane n/2-1, xz plane
        m/2-1, yz plane l/2-1 and the vorticity of the xy plane n/2+1; "hpc" (the -DHPC build) =
        the 3-D velocity and the xy plane n/2+1.  Returns the paths written."""
        l, m, n = self.shape
        shape = (l, m, n)
        files = []
        if branch == "hpc":
            items = [("3d", fortio.vtk_velocity_bytes(VTK_3D, 0, shape, self.vtk_velocity(VTK_3D))),
                     ("xy", fortio.vtk_velocity_bytes(VTK_XY, n // 2 + 1, shape,
                                                      self.vtk_velocity(VTK_XY, n // 2 + 1)))]
        elif branch == "default":
            items = [("xy", fortio.vtk_velocity_bytes(VTK_XY, n // 2 - 1, shape,
                                                      self.vtk_velocity(VTK_XY, n // 2 - 1))),
                     ("xz", fortio.vtk_velocity_bytes(VTK_XZ, m // 2 - 1, shape,
                                                      self.vtk_velocity(VTK_XZ, m // 2 - 1))),
                     ("yz", fortio.vtk_velocity_bytes(VTK_YZ, l // 2 - 1, shape,
                                                      self.vtk_velocity(VTK_YZ, l // 2 - 1))),
                     ("om", fortio.vtk_vorticity_bytes(shape, self.vtk_vorticity(n // 2 + 1)))]
        else:
            raise ValueError("branch must be 'default' or 'hpc'")
        for kind, data in items:
            path = os.path.join(directory, fortio.vtk_file_name(kind, itime, rank))
            fortio.write_bytes(path, data)
            files.append(path)
        return files

    def probe(self, i, j, k):
        out = np.zeros(4)
        self._check(self.L.eup_probe(self.h, int(i), int(j), int(k), _dptr(out)))
        return out

    def diagno(self):
        out = np.zeros(5)
        self._check(self.L.eup_diagno(self.h, _dptr(out)))
        return out

    # -- grey lattice -----------------------------------------------------------------------------
    def set_grey(self, s):
        """Solid fraction s[k, j, i] in [0, 1] (grey mode)."""
        l, m, n = self.shape
        a = np.ascontiguousarray(s, dtype=np.float64)
        if a.shape != (n, m, l):
            raise ValueError(f"grey field shape {a.shape} != {(n, m, l)}")
        self._check(self.L.eup_set_grey(self.h, _dptr(a)))

    def set_grey_device(self, ptr, dtype):
        """Solid fraction from device memory (``dtype`` 'f64' or 'f32')."""
        self._check(self.L.eup_set_grey_device(self.h, ctypes.c_void_p(ptr), STORAGE[dtype]))

    def grey_force(self):
        F = np.zeros(3)
        self._check(self.L.eup_grey_force(self.h, _dptr(F)))
        return F

    def grey_objective(self, t0, t1, every, norm, grad=True):
        """(J, dJ/ds) of the time-averaged drag F_x / norm over steps t0..t1 (state unchanged)."""
        l, m, n = self.shape
        J = ctypes.c_double()
        g = np.zeros((n, m, l)) if grad else None
        self._check(self.L.eup_grey_objective(self.h, int(t0), int(t1), int(every), float(norm),
                                              ctypes.byref(J),
                                              _dptr(g) if grad else ctypes.POINTER(ctypes.c_double)()))
        return J.value, g

    def grey_objective_device(self, t0, t1, every, norm, grad_ptr=None, dtype="f64"):
        J = ctypes.c_double()
        self._check(self.L.eup_grey_objective_device(
            self.h, int(t0), int(t1), int(every), float(norm), ctypes.byref(J),
            ctypes.c_void_p(grad_ptr) if grad_ptr else None, STORAGE[dtype]))
        return J.value

    def info(self):
        inf = EupInfo()
        self._check(self.L.eup_get_info(self.h, ctypes.byref(inf)))
        return {k: (list(getattr(inf, k)) if k == "bbox" else getattr(inf, k))
                for k, _ in EupInfo._fields_}

    def stream(self):
        return self.L.eup_stream(self.h)


def membw(device=0, mib=4096, reps=10):
    """HBM GB/s of copy / read / write streams of ``mib`` MiB buffers (eup_membw)."""
    L = load()
    out = np.zeros(3)
    rc = L.eup_membw(int(device), int(mib) << 20, int(reps), _dptr(out))
    if rc != EUP_OK:
        raise EupError(rc, L.eup_last_error(None).decode())
    return {"copy": out[0], "read": out[1], "write": out[2]}


def drag_boxes(lx, ly, lz, radius=100, deltas=(52, 57, 62, 67)):
    """Control boxes of the reference drag diagnostic (CODE/SRC/drag.F90:48-58)."""
    return [(d, (lx // 4 - radius - d, lx // 4 + radius + d, ly // 2 - radius - d,
                 ly // 2 + radius + d, 1, lz)) for d in deltas]
