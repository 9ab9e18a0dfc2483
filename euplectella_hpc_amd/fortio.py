"""Reference on-disk formats: namelist deck, obstacle list, raw restore / save dumps.

All binary files are Fortran *unformatted sequential* files: every record is framed by a 4-byte
little-endian length marker before and after the payload.

* deck ``bgk.input``: namelist ``&parameters`` (CODE/SRC/input.F90:46-62)
* ``obs.RRRR.TTTTTTTT.bin``: record ``nobs`` (int32), then ``nobs`` records ``i, j, k`` (int32,
  1-based) (reader CODE/SRC/build.obs.F90:147-186, writer :506-531)
* ``restore.XX.YY.ZZ.bin`` / ``save.XX.YY.ZZ.TTTTTTTT.bin``: record ``itime`` (int32), then 19
  records, population q = 1..19, each ``(l+2)(m+2)(n+2)`` reals with i fastest
  (CODE/SRC/restore_raw.F90:43-63, CODE/SRC/save_raw.F90:54-74)
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
