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
