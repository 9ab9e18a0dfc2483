"""The reference's on-disk formats written and read by euplectella_hpc_amd.fortio (CPU).

Round trips of the obstacle list (whole and streamed in chunks), the raw restore / save dumps and
the namelist deck; the binary VTK files are pinned against the reference's own files in
test_oracle_golden.py.
"""
import os

import numpy as np

from euplectella_hpc_amd import fortio, geometry


def test_obstacle_file_streams_in_chunks(tmp_path):
    l, m, n = 20, 12, 9
    mask = geometry.sponge(l, m, n, 3, (6, 6), cell=4, seed=3)
    path = str(tmp_path / fortio.obs_file_name())
    fortio.write_obs(path, mask)
    k, j, i = np.nonzero(mask)
    want = np.stack([i + 1, j + 1, k + 1], axis=1)
    for chunk in (1, 7, 1 << 20):
        blocks = list(fortio.iter_obs(path, chunk))
        assert all(b.shape[0] <= chunk for b in blocks)
        got = np.concatenate(blocks) if blocks else np.zeros((0, 3), np.int32)
        assert np.array_equal(got, want)
    assert np.array_equal(fortio.read_obs(path, (n, m, l)) != 0, mask != 0)


def test_empty_obstacle_file(tmp_path):
    path = str(tmp_path / "obs.bin")
    fortio.write_obs(path, np.zeros((3, 4, 5), np.uint8))
    assert list(fortio.iter_obs(path)) == []


def test_raw_dump_round_trip(tmp_path):
    pops = geometry.initial_state(6, 5, 4, 0.05, seed=2)
    path = str(tmp_path / fortio.save_file_name(17))
    fortio.write_raw(path, 17, pops)
    itime, back = fortio.read_raw(path, (4, 5, 6))
    assert itime == 17 and np.array_equal(back, pops)
    assert os.path.getsize(path) == (4 + 4 + 4) + 19 * (8 + 8 * 8 * 7 * 6)


def test_deck_round_trip(tmp_path):
    path = str(tmp_path / "bgk.input")
    fields = dict(svisc=0.0333333333333333, u0=0.05, itfin=100, ivtim=50, isignal=100,
                  itsave=100, icheck=10, irestart=1, init_v=0, lx=32, ly=16, lz=16)
    fortio.write_deck(path, **fields)
    back = fortio.read_deck(path)
    for key, value in fields.items():
        assert back[key] == value, key


def test_dense_obstacle_field_streams_solid_sites_by_truncation(tmp_path):
    """restore.obs.bin (the HDF5-free restore.obs.h5): the reals become the integer obstacle field
    by truncation toward zero and a site is solid where that integer is 1, as the reference's
    reader assigns the real buffer to its integer obs array and tests obs == 1
    (restore.obs.hdf5.serial.F90:160, build.obs.F90:114-118)."""
    l, m, n = 13, 7, 5
    rng = np.random.default_rng(4)
    vals = rng.choice(np.array([0.0, 1.0, 0.999, 1.5, 1.999, 2.0, -0.5, -1.0]), size=(n, m, l))
    want = np.trunc(vals) == 1
    assert want.any() and (~want).any()
    for dtype in ("<f8", "<f4"):
        path = str(tmp_path / f"obs_{dtype[-1]}.bin")
        fortio.write_obs_dense(path, vals, itime=42, dtype=dtype)
        for chunk in (1, 5, 1 << 20):
            it = fortio.iter_obs_dense(path, (l, m, n), chunk)
            blocks = []
            while True:
                try:
                    blocks.append(next(it))
                except StopIteration as stop:
                    assert stop.value == 42
                    break
            assert all(0 < b.shape[0] <= chunk for b in blocks)
            got = np.concatenate(blocks)
            k, j, i = np.nonzero(want)
            assert np.array_equal(got, np.stack([i + 1, j + 1, k + 1], axis=1))
        assert np.array_equal(fortio.read_obs_dense(path, (l, m, n)) != 0, want)


def test_dense_obstacle_field_refuses_other_dimensions(tmp_path):
    import pytest
    path = str(tmp_path / fortio.obs_dense_file_name())
    fortio.write_obs_dense(path, np.zeros((4, 3, 5)))
    assert fortio.read_obs_dense(path, (5, 3, 4)).sum() == 0
    with pytest.raises(ValueError, match="size doesn't match"):
        fortio.read_obs_dense(path, (5, 3, 5))
    with open(path, "ab") as fh:
        fh.write(b"\0")
    with pytest.raises(ValueError):
        fortio.read_obs_dense(path, (5, 3, 4))
