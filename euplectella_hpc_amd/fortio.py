"""Reference on-disk formats: namelist deck, obstacle list, raw restore / save dumps.

All binary files are Fortran *unformatted sequential* files: every record is framed by a 4-byte
little-endian length marker before and after the payload.

* deck ``bgk.input``: namelist ``&parameters`` (CODE/SRC/input.F90:46-62)
* ``obs.RRRR.TTTTTTTT.bin``: record ``nobs`` (int32), then ``nobs`` records ``i, j, k`` (int32,
  1-based) (reader CODE/SRC/build.obs.F90:147-186, writer :506-531)
* ``restore.obs.bin``: the HDF5-free equivalent of the reference's ``restore.obs.h5`` (datasets
  ``/time/timestep``, ``/size/dimension``, ``/obstacles/obs``; reader
  CODE/SRC/restore.obs.hdf5.serial.F90:85-160, writer save.obs.hdf5.serial.F90): record ``itime``
  (int32), record ``l, m, n`` (int32), then one record per z plane k = 1..n of ``l*m`` reals (FP64
  or FP32, i fastest).  As in the reference, the reals become the integer obstacle field by
  truncation and a site is solid where that integer is 1.
* ``restore.XX.YY.ZZ.bin`` / ``save.XX.YY.ZZ.TTTTTTTT.bin``: record ``itime`` (int32), then 19
  records, population q = 1..19, each ``(l+2)(m+2)(n+2)`` reals with i fastest
  (CODE/SRC/restore_raw.F90:43-63, CODE/SRC/save_raw.F90:54-74)
* ``tec_{xy,xz,yz,3d,om}.RRRR.TTTTTTTT.vtk``: binary VTK files, ASCII header then big-endian
  float32 data (CODE/SRC/vtk_*.binary.F90)
"""
import os
import re
import struct

import numpy as np

_MARK = struct.Struct("<i")


def write_records(path, records):
    """Write a sequence of byte strings / numpy arrays as Fortran sequential records."""
    with open(path, "wb") as fh:
        for rec in records:
            data = rec.tobytes() if isinstance(rec, np.ndarray) else bytes(rec)
            if len(data) > 0x7FFFFFFF:
                raise ValueError("record larger than 2 GiB is not supported by this writer")
            fh.write(_MARK.pack(len(data)))
            fh.write(data)
            fh.write(_MARK.pack(len(data)))


def iter_records(path):
    """Yield the payload of each Fortran sequential record as bytes."""
    with open(path, "rb") as fh:
        while True:
            head = fh.read(4)
            if not head:
                return
            if len(head) != 4:
                raise ValueError(f"{path}: truncated record marker")
            (n,) = _MARK.unpack(head)
            if n < 0:
                raise ValueError(f"{path}: sub-record markers are not supported")
            data = fh.read(n)
            tail = fh.read(4)
            if len(data) != n or len(tail) != 4 or _MARK.unpack(tail)[0] != n:
                raise ValueError(f"{path}: corrupt record framing")
            yield data


# ---------------------------------------------------------------------------------------------
# obstacle list
# ---------------------------------------------------------------------------------------------

def obs_file_name(rank=0, itime=0):
    return f"obs.{rank:04d}.{itime:08d}.bin"


def write_obs(path, mask):
    """Write the solid mask ``mask[k, j, i]`` (C order, shape (n, m, l), nonzero = solid).

    Records are emitted in the reference writer's loop order (k outer, j, i inner) with 1-based
    indices.
    """
    mask = np.asarray(mask)
    k, j, i = np.nonzero(mask)
    recs = [np.array([len(i)], dtype="<i4")]
    ijk = np.stack([i + 1, j + 1, k + 1], axis=1).astype("<i4")
    recs.extend(ijk)
    write_records(path, recs)


def iter_obs(path, chunk=1 << 20):
    """The (i, j, k) records of an obstacle file in blocks of at most ``chunk`` rows (int32
    [rows, 3], 1-based), read sequentially without holding the whole list."""
    with open(path, "rb") as fh:
        head = fh.read(12)
        if len(head) < 12:
            raise ValueError(f"{path}: truncated obstacle file")
        n1, nobs, n2 = struct.unpack("<iii", head)
        if n1 != 4 or n2 != 4:
            raise ValueError(f"{path}: bad record header")
        done = 0
        while done < nobs:
            rows = min(chunk, nobs - done)
            raw = np.frombuffer(fh.read(20 * rows), dtype="<i4")
            if raw.size != 5 * rows:
                raise ValueError(f"{path}: truncated obstacle file")
            rec = raw.reshape(rows, 5)
            if np.any(rec[:, 0] != 12) or np.any(rec[:, 4] != 12):
                raise ValueError(f"{path}: bad record marker")
            yield np.ascontiguousarray(rec[:, 1:4])
            done += rows


def read_obs(path, shape):
    """Read an obstacle list into a mask ``[n, m, l]`` (uint8)."""
    n, m, l = shape
    mask = np.zeros((n, m, l), dtype=np.uint8)
    it = iter_records(path)
    nobs = struct.unpack("<i", next(it))[0]
    cnt = 0
    for rec in it:
        i, j, k = struct.unpack("<3i", rec)
        if not (1 <= i <= l and 1 <= j <= m and 1 <= k <= n):
            raise ValueError(f"{path}: obstacle index out of range {(i, j, k)}")
        mask[k - 1, j - 1, i - 1] = 1
        cnt += 1
    if cnt != nobs:
        raise ValueError(f"{path}: header says {nobs} obstacles, found {cnt}")
    return mask


# ---------------------------------------------------------------------------------------------
# dense obstacle field (HDF5-free restore.obs.h5)
# ---------------------------------------------------------------------------------------------

def obs_dense_file_name():
    return "restore.obs.bin"


def write_obs_dense(path, values, itime=0, dtype="<f8"):
    """Write the obstacle field ``values[k, j, i]`` (shape (n, m, l); 1 = solid, as reals) with
    its time step and dimensions, one record per z plane."""
    values = np.asarray(values)
    if values.ndim != 3:
        raise ValueError("obstacle field must be 3-D [n, m, l]")
    n, m, l = values.shape

    def records():
        yield np.array([itime], dtype="<i4")
        yield np.array([l, m, n], dtype="<i4")
        for k in range(n):
            yield np.ascontiguousarray(values[k], dtype=dtype)

    write_records(path, records())


def _read_record(fh, path):
    head = fh.read(4)
    if len(head) != 4:
        raise ValueError(f"{path}: truncated dense obstacle file")
    (nb,) = _MARK.unpack(head)
    if nb < 0:
        raise ValueError(f"{path}: sub-record markers are not supported")
    data = fh.read(nb)
    tail = fh.read(4)
    if len(data) != nb or len(tail) != 4 or _MARK.unpack(tail)[0] != nb:
        raise ValueError(f"{path}: corrupt record framing")
    return data


def iter_obs_dense(path, shape, chunk=1 << 20):
    """The solid sites of a dense obstacle file as int32 [rows, 3] blocks of 1-based (i, j, k),
    at most ``chunk`` rows each, reading one z plane at a time.  ``shape`` is (l, m, n); a file
    of other dimensions is refused, as the reference stops on a size mismatch
    (restore.obs.hdf5.serial.F90:123-136).  Returns the file's time step through the generator's
    return value."""
    l, m, n = shape
    with open(path, "rb") as fh:
        itime = struct.unpack("<i", _read_record(fh, path))[0]
        dims = np.frombuffer(_read_record(fh, path), dtype="<i4")
        if dims.size != 3 or tuple(int(d) for d in dims) != (l, m, n):
            raise ValueError(f"{path}: size doesn't match: file {tuple(dims)}, lattice {(l, m, n)}")
        pending = []
        rows = 0
        for k in range(n):
            data = _read_record(fh, path)
            if len(data) == 8 * l * m:
                vals = np.frombuffer(data, dtype="<f8")
            elif len(data) == 4 * l * m:
                vals = np.frombuffer(data, dtype="<f4")
            else:
                raise ValueError(f"{path}: plane {k + 1} holds {len(data)} bytes, not l*m reals")
            # Fortran real -> integer assignment truncates toward zero; solid where it gives 1
            solid = np.trunc(vals) == 1
            jj, ii = np.nonzero(solid.reshape(m, l))
            if ii.size:
                trip = np.empty((ii.size, 3), dtype=np.int32)
                trip[:, 0] = ii + 1
                trip[:, 1] = jj + 1
                trip[:, 2] = k + 1
                pending.append(trip)
                rows += ii.size
            while rows >= chunk:
                allr = np.concatenate(pending)
                yield allr[:chunk]
                rest = allr[chunk:]
                pending = [rest] if rest.size else []
                rows = rest.shape[0]
        if rows:
            yield np.concatenate(pending)
        if fh.read(1):
            raise ValueError(f"{path}: data after the last plane")
    return itime


def read_obs_dense(path, shape):
    """Read a dense obstacle file into a mask ``[n, m, l]`` (uint8)."""
    l, m, n = shape
    mask = np.zeros((n, m, l), dtype=np.uint8)
    for trip in iter_obs_dense(path, shape):
        mask[trip[:, 2] - 1, trip[:, 1] - 1, trip[:, 0] - 1] = 1
    return mask


# ---------------------------------------------------------------------------------------------
# raw restore / save dumps
# ---------------------------------------------------------------------------------------------

def restore_file_name(cx=0, cy=0, cz=0):
    return f"restore.{cx:02d}.{cy:02d}.{cz:02d}.bin"


def save_file_name(itime, cx=0, cy=0, cz=0):
    return f"save.{cx:02d}.{cy:02d}.{cz:02d}.{itime:08d}.bin"


def write_raw(path, itime, pops):
    """Write populations ``pops[19, n+2, m+2, l+2]`` (float64) as a restore/save dump."""
    pops = np.asarray(pops)
    if pops.ndim != 4 or pops.shape[0] != 19:
        raise ValueError("pops must have shape (19, n+2, m+2, l+2)")
    recs = [np.array([itime], dtype="<i4")]
    recs.extend(np.ascontiguousarray(pops[q], dtype="<f8") for q in range(19))
    write_records(path, recs)


def read_raw(path, shape, dtype="<f8"):
    """Read a restore/save dump; ``shape`` = (n, m, l) interior sizes.  Returns (itime, pops)."""
    n, m, l = shape
    it = iter_records(path)
    itime = struct.unpack("<i", next(it))[0]
    pops = np.empty((19, n + 2, m + 2, l + 2), dtype=np.float64)
    for q in range(19):
        rec = next(it)
        arr = np.frombuffer(rec, dtype=dtype)
        if arr.size != (n + 2) * (m + 2) * (l + 2):
            raise ValueError(f"{path}: record {q + 1} has {arr.size} reals, expected "
                             f"{(n + 2) * (m + 2) * (l + 2)}")
        pops[q] = arr.reshape(n + 2, m + 2, l + 2)
    return itime, pops


# ---------------------------------------------------------------------------------------------
# namelist deck
# ---------------------------------------------------------------------------------------------

DECK_FIELDS = ("svisc", "u0", "itfin", "ivtim", "isignal", "itsave", "icheck", "irestart",
               "init_v", "lx", "ly", "lz", "proc_x", "proc_y", "proc_z", "tstep", "flag1",
               "flag2", "flag3", "ipad", "jpad", "kpad")

#: defaults applied before the namelist is read (CODE/SRC/input.F90:51-58)
DECK_DEFAULTS = {"flag1": 2, "flag2": 2, "flag3": 1, "ipad": 0, "jpad": 0, "kpad": 0}


def _fmt(v):
    if isinstance(v, float):
        return repr(float(v))  # shortest round-trip decimal: the reader recovers the same double
    return str(int(v))


def write_deck(path, **fields):
    """Write a ``&parameters`` namelist deck."""
    bad = set(fields) - set(DECK_FIELDS)
    if bad:
        raise ValueError(f"unknown deck fields: {sorted(bad)}")
    lines = ["&parameters"]
    for key in DECK_FIELDS:
        if key in fields:
            lines.append(f"{key} = {_fmt(fields[key])}")
    lines.append("/")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_deck(path):
    """Parse a ``&parameters`` namelist deck with scalar values into a dict."""
    with open(path) as fh:
        text = fh.read()
    m = re.search(r"&parameters(.*?)/", text, re.S | re.I)
    if not m:
        raise ValueError(f"{path}: no &parameters group")
    body = m.group(1)
    out = dict(DECK_DEFAULTS)
    for key, val in re.findall(r"(\w+)\s*=\s*([^\s,/]+)", body):
        key = key.lower()
        if key not in DECK_FIELDS:
            raise ValueError(f"{path}: unknown namelist field {key}")
        val = val.lower().replace("d", "e")
        out[key] = float(val) if key in ("svisc", "u0") else int(float(val))
    return out


def ensure_dir(path):
    os.makedirs(path, exist_ok=True)
    return path


# ----------------------------------------------------------------------------------------------
# binary VTK files (CODE/SRC/vtk_*.binary.F90): an ASCII header written with the writers' fixed
# edit descriptors, then big-endian float32 data (access='STREAM', CONVERT="BIG_ENDIAN")
# ----------------------------------------------------------------------------------------------

VTK_YZ, VTK_XZ, VTK_XY, VTK_3D = 0, 1, 2, 3
_VTK_TAG = {VTK_YZ: "yz", VTK_XZ: "xz", VTK_XY: "xy", VTK_3D: "3d"}


def vtk_file_name(kind, itime, rank=0):
    """'tec_<kind>.RRRR.TTTTTTTT.vtk' (vtk_xy.binary.F90:43-47); kind in xy, xz, yz, 3d, om."""
    return f"tec_{kind}.{rank:04d}.{itime:08d}.vtk"


def _vtk_header(dims, origin, count, data_lines):
    # A26, A5, A6, A25; (A11,I10,A1,I10,A1,I10); (A7,...); (A8,...); (A10,I10) -- the literal
    # 'POINT_DATA ' is cut to its first 10 characters by A10 -- and the data declaration in an Aw
    # one wider than the text, so right-justified behind one blank (vtk_xy.binary.F90:51-64)
    three = lambda v: "%10d %10d %10d" % tuple(v)
    lines = ["# vtk DataFile Version 2.0", "Campo", "BINARY", "DATASET STRUCTURED_POINTS",
             "DIMENSIONS " + three(dims), "ORIGIN " + three(origin), "SPACING " + three((1, 1, 1)),
             "POINT_DATA" + "%10d" % count] + data_lines
    return ("\n".join(lines) + "\n").encode("ascii")


def vtk_velocity_bytes(plane, index, shape, data, offset=(0, 0, 0)):
    """Bytes of the reference's velocity VTK file for ``plane`` (VTK_YZ / VTK_XZ / VTK_XY at the
    1-based ``index``, or VTK_3D), ``shape`` = (l, m, n), ``data`` float32 [sites, 3] in the
    writer's loop order (as eup_vtk_velocity returns it)."""
    l, m, n = shape
    ox, oy, oz = offset
    dims, origin = {
        VTK_YZ: ((1, m, n), (ox + 1 + index, oy + 1, oz + 1)),   # vtk_yz.binary.F90:55-59
        VTK_XZ: ((l, 1, n), (ox + 1, oy + index + 1, oz + 1)),   # vtk_xz.binary.F90:55-59
        VTK_XY: ((l, m, 1), (ox + 1, oy + 1, oz + 1 + index)),   # vtk_xy.binary.F90:55-59
        VTK_3D: ((l, m, n), (ox + 1, oy + 1, oz + 1)),           # vtk_3d.binary.F90:58-62
    }[plane]
    count = dims[0] * dims[1] * dims[2]
    arr = np.ascontiguousarray(data, dtype=np.float32).reshape(count, 3)
    return _vtk_header(dims, origin, count, [" VECTORS velocity float"]) + arr.astype(">f4").tobytes()


def vtk_vorticity_bytes(shape, data, offset=(0, 0, 0)):
    """Bytes of the reference's vorticity VTK file (vtk_om.binary.F90:73-86): (l-2) x (m-2)
    points with origin (2, 2, 1) whatever the plane, float32 data as eup_vtk_vorticity
    returns it."""
    l, m, _ = shape
    ox, oy, oz = offset
    count = (l - 2) * (m - 2)
    arr = np.ascontiguousarray(data, dtype=np.float32).reshape(count)
    return (_vtk_header((l - 2, m - 2, 1), (ox + 2, oy + 2, oz + 1), count,
                        [" SCALARS vorticity float", "LOOKUP_TABLE default"])
            + arr.astype(">f4").tobytes())


def write_bytes(path, data):
    with open(path, "wb") as fh:
        fh.write(data)
