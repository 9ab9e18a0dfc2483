"""GPU parity: libeup (through its C-ABI) against the golden fixtures and the oracle.

Parity mode must reproduce the reference's serial FP64 step bit for bit (every site, ghosts
included).  Physics mode in FP64 must equal the oracle's physics step bit for bit; reduced
storage is checked against it with tolerances stated in each test.
"""
import os
import sys

import numpy as np
import pytest

from golden_cases import CASES, make_inputs, omega_of, sha, e14_6
from momentum_balance import box_momentum
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lattice(rec_or_shape, **kw):
    from euplectella_hpc_amd.lib import Lattice
    return Lattice(*rec_or_shape, **kw)


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def parity_lattice(name, golden):
    rec = golden[name]
    mask, pops = make_inputs(CASES[name])
    lat = lattice((rec["l"], rec["m"], rec["n"]), svisc=rec["svisc"], u_inflow=rec["u_in"],
                  mode="parity", storage="f64")
    lat.set_mask(mask)
    lat.set_state(pops)
    return rec, mask, pops, lat


@pytest.mark.parametrize("remap", ["0", "1", "2"])
def test_parity_tiny_full_state(golden, remap, monkeypatch):
    """Every site of the state (ghosts included) after 20 steps, bit for bit, with the
    stream-collide in each block order (dispatch order, XCD-contiguous ranges, XCD-owned rows)."""
    monkeypatch.setenv("EUP_PARITY_XCD_REMAP", remap)
    announce(f"parity c0_tiny 16x8x12, 20 steps, xcd_remap={remap}")
    rec, mask, pops, lat = parity_lattice("c0_tiny", golden)
    with lat:
        lat.step(rec["steps"])
        got = lat.get_state()
    ref = np.load(os.path.join(GOLDEN, "c0_tiny_state.npz"))[f"t{rec['steps']}"]
    bad = got.view(np.uint64) != ref.view(np.uint64)
    assert not bad.any(), f"{bad.sum()} values differ (first at {np.argwhere(bad)[:3].tolist()})"


@pytest.mark.parametrize("name,remap", [("c1_small", "2"), ("c2_ridged", "2"), ("c3_sponge", "2"),
                                         ("c1_cylinder", "2"), ("c3_sponge", "1"), ("c3_sponge", "0"),
                                         ("c4_scattered", "2"), ("c4_scattered", "1")])
def test_parity_state_hashes(golden, name, remap, monkeypatch):
    monkeypatch.setenv("EUP_PARITY_XCD_REMAP", remap)
    rec, mask, pops, lat = parity_lattice(name, golden)
    announce(f"parity {name} {rec['l']}x{rec['m']}x{rec['n']} up to {rec['steps']} steps")
    done = 0
    with lat:
        for t in sorted(int(s) for s in rec["state_sha256"]):
            lat.step(t - done)
            done = t
            got = lat.get_state()
            assert sha(got) == rec["state_sha256"][str(t)], f"{name}: state differs at step {t}"


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", ["c2_full", "c3_full"])
def test_full_size_parity_equals_the_reference_and_physics_fp64_the_oracle(golden, name):
    """The benchmark configurations at full size (ridged cylinder 256 x 128 x 128, 200 steps; the
    512 x 256 x 256 sponge, 100 steps), from the seeded impulsive start: parity mode's ghosted state
    equals the reference executable's save dumps at both save points, and physics mode in FP64
    equals the oracle's physics step on every interior site at the end, bit for bit (SHA-256 of the
    states; tests/golden/make_golden.py c2_full c3_full)."""
    from make_golden import FULL

    rec = golden[name]
    l, m, n = rec["l"], rec["m"], rec["n"]
    announce(f"{name} {l}x{m}x{n}: generating the mask and the start")
    mask, pops = make_inputs(FULL[name])
    assert sha(mask.astype(np.int32)) == rec["mask_sha256"] and sha(pops) == rec["init_sha256"]
    announce(f"{name}: parity mode, saves at {sorted(int(x) for x in rec['state_sha256'])}")
    with lattice((l, m, n), svisc=rec["svisc"], u_inflow=rec["u_in"], mode="parity", storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        done = 0
        for t in sorted(int(x) for x in rec["state_sha256"]):
            lat.step(t - done)
            done = t
            assert sha(lat.get_state()) == rec["state_sha256"][str(t)], f"parity state differs at step {t}"
    announce(f"{name}: physics mode FP64, {rec['steps']} steps")
    with lattice((l, m, n), svisc=rec["svisc"], u_inflow=rec["u_in"], mode="physics", storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(rec["steps"])
        got = lat.get_state()
    assert sha(got[:, 1:-1, 1:-1, 1:-1]) == rec["physics_interior_sha256"]


@pytest.mark.parametrize("remap", ["0", "1", "2"])
@pytest.mark.parametrize("shape", [(320, 12, 10), (512, 8, 6)])
def test_parity_block_orders_equal_oracle(shape, remap, monkeypatch):
    """Parity mode in each stream-collide block order against the oracle's serial step, bit for bit
    (ghosts included), on lattices whose rows span two blocks (320 x 12 x 10 also ends in a
    partial group of blocks)."""
    monkeypatch.setenv("EUP_PARITY_XCD_REMAP", remap)
    l, m, n = shape
    announce(f"parity cylinder {l}x{m}x{n} xcd_remap={remap} vs oracle, 12 steps")
    mask, pops = physics_case("cylinder", l, m, n, seed=11)
    svisc = (2 * 0.6 - 1) / 6
    ref = pops.copy()
    orc.parity_steps(ref, mask, orc.omega_from_svisc(svisc), 0.05, 12)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="parity", storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(12)
        got = lat.get_state()
    bad = got.view(np.uint64) != ref.view(np.uint64)
    assert not bad.any(), f"{bad.sum()} values differ (first at {np.argwhere(bad)[:3].tolist()})"


@pytest.mark.parametrize("name,nsteps", [("c1_small", 10), ("c2_ridged", 60)])
def test_parity_probe_records(golden, name, nsteps):
    rec, mask, pops, lat = parity_lattice(name, golden)
    announce(f"parity probe {name}")
    i, j, k = rec["probe_site"]
    with lat:
        for row in rec["probe"][:nsteps]:
            lat.step(1)
            got = lat.probe(i, j, k)
            assert [e14_6(v) for v in got] == [e14_6(v) for v in row[1:5]], f"step {row[0]}"


def test_parity_drag_records(golden):
    rec, mask, pops, lat = parity_lattice("drag_r100", golden)
    announce("parity drag_r100 680x340x4, 3 steps, control-volume drag every step")
    rows = iter(rec["drag"])
    from euplectella_hpc_amd.lib import drag_boxes
    with lat:
        for step in range(1, rec["steps"] + 1):
            lat.step(1)
            for delta, box in drag_boxes(rec["l"], rec["m"], rec["n"]):
                d, t, fx, fy = next(rows)
                assert (int(d), int(t)) == (delta, step)
                gx, gy = lat.drag_cv(box)
                assert gx == fx and gy == fy, f"delta {delta} step {step}: {gx} {gy} vs {fx} {fy}"
        assert sha(lat.get_state()) == rec["state_sha256"][str(rec["steps"])]


@pytest.mark.parametrize("name", ["c0_tiny", "c1_small"])
def test_parity_vtk_files_match_reference(golden, name, tmp_path):
    """eup_vtk_velocity / eup_vtk_vorticity on the parity state: the binary VTK files of both
    output branches equal the reference's byte for byte."""
    import hashlib
    rec, mask, pops, lat = parity_lattice(name, golden)
    announce(f"parity {name}: VTK files after {rec['steps']} steps")
    with lat:
        lat.step(rec["steps"])
        for branch in ("default", "hpc"):
            d = tmp_path / branch
            d.mkdir()
            paths = lat.write_vtk(rec["steps"], str(d), branch)
            want = rec["vtk"][branch]
            assert sorted(os.path.basename(p) for p in paths) == sorted(want)
            for p in paths:
                data = open(p, "rb").read()
                assert hashlib.sha256(data).hexdigest() == want[os.path.basename(p)]["sha256"], p


# ----------------------------------------------------------------------------------------------
# physics mode
# ----------------------------------------------------------------------------------------------

def physics_case(kind, l, m, n, seed, u=0.05):
    from euplectella_hpc_amd import geometry
    if kind == "cylinder":
        mask = geometry.cylinder(l, m, n, max(2, m // 5), (l // 4, m // 2))
    elif kind == "sponge":
        mask = geometry.sponge(l, m, n, m // 4, (l // 4, m // 2), cell=8, seed=seed)
    elif kind == "random":
        # scattered solid sites: most tiles mixed, some sites with every neighbour solid
        mask = (np.random.default_rng(seed).random((n, m, l)) < 0.2).astype(np.uint8)
    elif kind == "block":
        # a solid block covering whole 64-site tiles (all-solid tiles) next to a cylinder
        mask = geometry.cylinder(l, m, n, max(2, m // 6), (l // 5, m // 2))
        mask[:, 2:m - 3, 64:128] = 1
    else:
        mask = np.zeros((n, m, l), np.uint8)
    pops = geometry.initial_state(l, m, n, u, seed=seed)
    return mask, pops


def interior(p):
    return p[:, 1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("remap", ["0", "1", "2"])
@pytest.mark.parametrize("kind,x_open,shape", [("cylinder", True, (70, 24, 20)),
                                                ("cylinder", False, (64, 24, 20)),
                                                ("sponge", False, (100, 24, 20)),
                                                ("sponge", True, (128, 40, 48)),
                                                ("block", True, (200, 24, 20)),
                                                ("block", False, (192, 24, 20)),
                                                ("random", True, (5, 3, 2)),
                                                ("random", False, (7, 3, 2)),
                                                ("random", False, (67, 2, 3)),
                                                ("random", True, (130, 5, 4)),
                                                ("cylinder", True, (320, 12, 10)),
                                                ("random", False, (512, 8, 6))])
def test_physics_f64_equals_oracle(kind, x_open, shape, remap, monkeypatch):
    """The fused kernel in its three block orders (dispatch order, XCD-contiguous ranges,
    XCD-owned tile rows) against the oracle, bit for bit.  320 and 512 sites in x give rows of
    5 and 8 tiles (two blocks per row, so the row order deals out groups of 16 blocks; 320 x 12 x
    10 also ends in a partial group)."""
    monkeypatch.setenv("EUP_FUSED_XCD_REMAP", remap)
    l, m, n = shape
    announce(f"physics f64 {kind} {shape} x_open={x_open} xcd_remap={remap} vs oracle, 25 steps")
    mask, pops = physics_case(kind, l, m, n, seed=7)
    svisc = (2 * 0.55 - 1) / 6
    ref = pops.copy()
    orc.physics_steps(ref, mask, orc.omega_from_svisc(svisc), 0.05, 25, x_open=x_open)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="physics", storage="f64",
                 x_open=x_open) as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(25)
        got = lat.get_state()
        if kind == "block":
            assert lat.info()["tiles_solid"] > 0
    a, b = interior(got), interior(ref)
    bad = a.view(np.uint64) != b.view(np.uint64)
    assert not bad.any(), f"{bad.sum()} values differ, max abs {np.abs(a - b).max()}"


@pytest.mark.parametrize("shape,x_open", [((128, 48, 48), True), ((200, 24, 24), False),
                                          ((320, 24, 20), True)])
@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16", "e3m12"])
def test_physics_tile_skip_is_exact(storage, shape, x_open, monkeypatch):
    """The pure-fluid tile path against the per-site mixed path (every tile marked mixed), in
    every block order."""
    l, m, n = shape
    mask, pops = physics_case("sponge", l, m, n, seed=3)
    svisc = (2 * 0.52 - 1) / 6
    out = []
    for remap in ("0", "1", "2"):
        monkeypatch.setenv("EUP_FUSED_XCD_REMAP", remap)
        for skip in (True, False):
            announce(f"physics {storage} sponge {l}x{m}x{n} x_open={x_open} xcd_remap={remap} "
                     f"tile_skip={skip}")
            with lattice((l, m, n), svisc=svisc, u_inflow=0.06, mode="physics", storage=storage,
                         tile_skip=skip, x_open=x_open) as lat:
                lat.set_mask(mask)
                lat.set_state(pops)
                lat.step(20)
                out.append(lat.get_state())
                if skip:
                    info = lat.info()
                    assert info["tiles_mixed"] > 0 and info["tiles_fluid"] > 0
    for o in out[1:]:
        assert np.array_equal(out[0], o)


@pytest.mark.parametrize("storage,tol,shifted,x_open", [("f32", 2e-6, False, True),
                                                        ("f16", 2e-4, False, True),
                                                        ("bf16", 2e-3, False, True),
                                                        ("f16", 5e-5, True, True),
                                                        ("bf16", 5e-4, True, True),
                                                        ("f32", 2e-6, False, False),
                                                        ("f16", 2e-4, False, False),
                                                        ("e3m12", 1e-4, False, True),
                                                        ("e3m12", 2e-5, True, True),
                                                        ("e3m12", 1e-4, False, False)])
def test_physics_reduced_storage_close_to_f64(storage, tol, shifted, x_open):
    """Populations after 40 steps within ``tol`` (absolute; populations are O(1/36)) of FP64;
    the free-stream-shifted deviations have the tighter bound.  Open and periodic x."""
    l, m, n = 96, 32, 32
    announce(f"physics {storage} shifted={shifted} x_open={x_open} vs oracle f64, cylinder "
             f"{l}x{m}x{n}, 40 steps")
    mask, pops = physics_case("cylinder", l, m, n, seed=5)
    svisc = (2 * 0.6 - 1) / 6
    ref = pops.copy()
    orc.physics_steps(ref, mask, orc.omega_from_svisc(svisc), 0.05, 40, x_open=x_open)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="physics", storage=storage,
                 shifted=shifted, x_open=x_open) as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(40)
        got = lat.get_state()
    fluid = mask == 0
    err = np.abs(interior(got) - interior(ref))[:, fluid].max()
    assert err < tol, f"{storage}: max abs population error {err}"


@pytest.mark.parametrize("storage,tol", [("f32", 2e-6), ("f16", 2e-4), ("bf16", 2e-3), ("e3m12", 1e-4)])
@pytest.mark.parametrize("x_open,shape", [(True, (130, 6, 5)), (False, (130, 6, 5)),
                                          (False, (7, 3, 2))])
def test_physics_reduced_storage_scattered_solids(storage, tol, x_open, shape):
    """Scattered solid sites (every tile mixed, lanes with many solid neighbours): the reduced
    storages' mixed-tile path (all loads at once, bounce-back chosen per lane) against FP64;
    also a lattice narrower than one tile with periodic x (first and last tile the same)."""
    l, m, n = shape
    announce(f"physics {storage} x_open={x_open} scattered solids {l}x{m}x{n} vs oracle f64")
    mask, pops = physics_case("random", l, m, n, seed=8)
    svisc = (2 * 0.6 - 1) / 6
    ref = pops.copy()
    orc.physics_steps(ref, mask, orc.omega_from_svisc(svisc), 0.05, 40, x_open=x_open)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="physics", storage=storage,
                 x_open=x_open) as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(40)
        got = lat.get_state()
        assert lat.info()["tiles_mixed"] > 0
    fluid = mask == 0
    err = np.abs(interior(got) - interior(ref))[:, fluid].max()
    assert err < tol, f"{storage}: max abs population error {err}"


def test_physics_drag_mem_and_diagno_match_oracle():
    l, m, n = 96, 32, 24
    announce(f"physics f64 drag_mem / diagno / probe vs oracle {l}x{m}x{n}")
    mask, pops = physics_case("cylinder", l, m, n, seed=9)
    svisc = (2 * 0.6 - 1) / 6
    ref = pops.copy()
    orc.physics_steps(ref, mask, orc.omega_from_svisc(svisc), 0.05, 30)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="physics", storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(30)
        F = lat.drag_mem()
        d = lat.diagno()
        p = lat.probe(40, 20, 5)
    Fr = orc.mem_force(ref, mask)
    # summation order differs (wavefront tree vs sequential): absolute bound scaled by |F|
    assert np.abs(F - Fr).max() <= 1e-12 * np.abs(Fr).max(), (F, Fr)
    dr = orc.diagno(ref, mask)
    assert np.allclose(d, dr, rtol=1e-13, atol=1e-16), (d, dr)
    pr = orc.probe(ref, 40, 20, 5)
    assert np.array_equal(p, pr)


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_physics_vtk_fields_equal_oracle(storage):
    """VTK velocity planes, volume and vorticity of a physics-mode state (tile-major layout, any
    storage) equal the oracle's on the same state read back as FP64."""
    l, m, n = 70, 24, 20
    announce(f"physics {storage} sponge {l}x{m}x{n}: VTK fields vs oracle")
    mask, pops = physics_case("sponge", l, m, n, seed=12)
    with lattice((l, m, n), svisc=(2 * 0.6 - 1) / 6, u_inflow=0.05, mode="physics",
                 storage=storage) as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(15)
        state = lat.get_state()
        for plane, index in ((orc.VTK_YZ, 1), (orc.VTK_YZ, l), (orc.VTK_XZ, 7), (orc.VTK_XY, n),
                             (orc.VTK_3D, 0)):
            got = lat.vtk_velocity(plane, index)
            assert np.array_equal(got.view(np.uint32),
                                  orc.vtk_velocity(state, plane, index).view(np.uint32)), (plane, index)
        for k0 in (1, 11):
            assert np.array_equal(lat.vtk_vorticity(k0).view(np.uint32),
                                  orc.vtk_vorticity(state, k0).view(np.uint32)), k0


@pytest.mark.parametrize("mode,storage", [("parity", "f64"), ("physics", "f64"), ("physics", "f16")])
def test_streamed_obstacle_list_equals_dense_mask(mode, storage, tmp_path):
    """eup_add_obstacles (chunks of 97 records, one chunk given twice) + eup_commit_obstacles give
    the same solid count, bounding box and states as eup_set_mask with the dense array."""
    from euplectella_hpc_amd import fortio
    l, m, n = 70, 24, 20
    announce(f"{mode} {storage} sponge {l}x{m}x{n}: streamed obstacle list vs dense mask")
    mask, pops = physics_case("sponge", l, m, n, seed=21)
    path = str(tmp_path / fortio.obs_file_name())
    fortio.write_obs(path, mask)
    out, infos = [], []
    for streamed in (False, True):
        with lattice((l, m, n), svisc=(2 * 0.6 - 1) / 6, u_inflow=0.05, mode=mode,
                     storage=storage) as lat:
            if streamed:
                blocks = list(fortio.iter_obs(path, 97))
                for b in blocks + blocks[:1]:
                    lat.add_obstacles(b)
                lat.commit_obstacles()
            else:
                lat.set_mask(mask)
            lat.set_state(pops)
            lat.step(10)
            out.append(lat.get_state())
            info = lat.info()
            infos.append((info["nsolid"], tuple(info["bbox"])))
    assert infos[0] == infos[1] and infos[0][0] == int(mask.sum())
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


@pytest.mark.parametrize("mode", ["parity", "physics"])
def test_dense_obstacle_file_equals_dense_mask(mode, tmp_path):
    """Lattice.set_obstacle_dense_file (restore.obs.bin streamed plane by plane through
    eup_add_obstacles) gives the same solid count, bounding box and states as eup_set_mask."""
    from euplectella_hpc_amd import fortio
    l, m, n = 70, 24, 20
    announce(f"{mode} sponge {l}x{m}x{n}: dense obstacle file vs dense mask")
    mask, pops = physics_case("sponge", l, m, n, seed=22)
    path = str(tmp_path / fortio.obs_dense_file_name())
    fortio.write_obs_dense(path, mask.astype(np.float64), itime=3)
    out, infos = [], []
    for from_file in (False, True):
        with lattice((l, m, n), svisc=(2 * 0.6 - 1) / 6, u_inflow=0.05, mode=mode,
                     storage="f64") as lat:
            if from_file:
                assert lat.set_obstacle_dense_file(path, chunk=131) == 3
            else:
                lat.set_mask(mask)
            lat.set_state(pops)
            lat.step(10)
            out.append(lat.get_state())
            info = lat.info()
            infos.append((info["nsolid"], tuple(info["bbox"])))
    assert infos[0] == infos[1] and infos[0][0] == int(mask.sum())
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_obstacle_list_commit_replace_and_range_check():
    from euplectella_hpc_amd.lib import EupError
    l, m, n = 40, 16, 12
    announce("physics obstacle list: empty commit, replacement, range check")
    with lattice((l, m, n), svisc=0.1, u_inflow=0.05, mode="physics", storage="f32") as lat:
        lat.commit_obstacles()
        assert lat.info()["nsolid"] == 0
        lat.add_obstacles([[3, 4, 5], [3, 4, 5], [10, 1, 12]])
        lat.commit_obstacles()
        info = lat.info()
        assert info["nsolid"] == 2 and tuple(info["bbox"]) == (3, 10, 1, 4, 5, 12)
        lat.add_obstacles([[40, 16, 1]])  # a new list replaces the committed one
        lat.commit_obstacles()
        info = lat.info()
        assert info["nsolid"] == 1 and tuple(info["bbox"]) == (40, 40, 16, 16, 1, 1)
        with pytest.raises(EupError):
            lat.add_obstacles([[41, 1, 1]])
        with pytest.raises(EupError):
            lat.add_obstacles([[1, 0, 1]])


def test_momentum_exchange_closes_the_momentum_balance():
    """Discrete momentum balance of one step around the body: the momentum stored in a box after
    the step minus before = inflow through its surface - the momentum-exchange force (collision
    conserves momentum, bounce-back hands 2 c f to the body).  Checked on two boxes after a
    developed start; the mirror-symmetric cylinder (no noise) has no lift.  R = 5, u_in = 0.02,
    tau = 0.8 (Re = 2)."""
    from euplectella_hpc_amd import geometry
    from euplectella_hpc_amd.lib import svisc_from_tau
    l, m, n = 96, 48, 4
    announce(f"physics f64 cylinder {l}x{m}x{n}: momentum exchange closes the box balance")
    mask = geometry.cylinder(l, m, n, 5, (24, (m + 1) / 2))
    pops = geometry.initial_state(l, m, n, 0.02, noise=0.0)
    with lattice((l, m, n), svisc=svisc_from_tau(0.8), u_inflow=0.02, mode="physics",
                 storage="f64") as lat:
        lat.set_mask(mask)
        lat.set_state(pops)
        lat.step(3000)
        F = lat.drag_mem()
        P0 = lat.get_state()
        lat.step(1)
        P1 = lat.get_state()
    for box in ((8, 60, 8, 41), (14, 34, 14, 35)):
        M0, G = box_momentum(P0, box)
        M1, _ = box_momentum(P1, box)
        err = (M1 - M0) - (G - F)
        announce(f"  box {box}: F_mem_x {F[0]:.12e}, inflow_x {G[0]:.12e}, dM_x {M1[0] - M0[0]:.3e}, "
                 f"balance error {np.abs(err).max():.2e}")
        assert np.abs(err).max() < 1e-10 * abs(F[0])
    assert abs(F[1]) < 1e-12 * abs(F[0]) and abs(F[2]) < 1e-12 * abs(F[0])


def test_parity_diagno_and_mem_match_oracle(golden):
    rec, mask, pops, lat = parity_lattice("c1_small", golden)
    announce("parity c1_small diagno / drag_mem vs oracle")
    with lat:
        lat.step(10)
        d = lat.diagno()
        F = lat.drag_mem()
        got = lat.get_state()
    assert np.allclose(d, orc.diagno(got, mask), rtol=1e-13, atol=1e-16)
    Fr = orc.mem_force(got, mask)
    assert np.abs(F - Fr).max() <= 1e-12 * np.abs(Fr).max(), (F, Fr)


def hash_uniform(seed, n):
    """numpy restatement of the library's counter-based uniform in [-1, 1) (splitmix64)."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (i + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


@pytest.mark.parametrize("mode", ["parity", "physics"])
def test_init_equilibrium_matches_the_closed_form(mode):
    from euplectella_hpc_amd import d3q19
    l, m, n = 20, 12, 10
    announce(f"init_equilibrium {mode} {l}x{m}x{n}")
    ux, noise, seed = 0.06, 1e-4, 77
    with lattice((l, m, n), svisc=0.05, u_inflow=ux, mode=mode, storage="f64") as lat:
        lat.set_mask(np.zeros((n, m, l), np.uint8))
        lat.init_equilibrium(ux, 0.0, 0.0, noise, seed)
        got = lat.get_state()
    if mode == "parity":
        uy = (noise * hash_uniform(seed, (n + 2) * (m + 2) * (l + 2))).reshape(n + 2, m + 2, l + 2)
        region = (slice(None),) * 4
    else:
        uy = np.zeros((n + 2, m + 2, l + 2))
        uy[1:-1, 1:-1, 1:-1] = (noise * hash_uniform(seed, n * m * l)).reshape(n, m, l)
        region = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    usq = ux * ux + uy * uy
    for q in range(19):
        cu = d3q19.C[q, 0] * ux + d3q19.C[q, 1] * uy
        e = d3q19.W[q] * (1.0 + (3.0 * cu + 1.5 * (3.0 * cu * cu - usq)))
        assert np.abs(got[q][region[1:]] - e[region[1:]]).max() < 1e-17, q


@pytest.mark.parametrize("mode", ["parity", "physics"])
def test_rejected_mask_leaves_the_committed_mask_in_place(mode):
    """eup_set_mask with a value other than 0 / 1 returns EUP_EINVAL before it touches the device:
    the mask set before it stays in force (same solid count, bounding box and steps, bit for bit)."""
    import ctypes
    from euplectella_hpc_amd import geometry
    from euplectella_hpc_amd.lib import EUP_OK
    l, m, n = 48, 16, 8
    announce(f"{mode} f64 {l}x{m}x{n}: rejected mask keeps the committed one")
    mask = geometry.cylinder(l, m, n, 3, (12, (m + 1) / 2))
    pops = geometry.initial_state(l, m, n, 0.04, noise=1e-4, seed=5)
    bad = np.ascontiguousarray(mask != 0, dtype=np.int32)
    bad[n // 2, m // 2, l // 2] = 2
    states = []
    for reject in (False, True):
        with lattice((l, m, n), svisc=0.05, u_inflow=0.04, mode=mode, storage="f64") as lat:
            lat.set_mask(mask)
            info0 = lat.info()
            if reject:
                rc = lat.L.eup_set_mask(lat.h, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
                assert rc != EUP_OK
                info1 = lat.info()
                assert info1["nsolid"] == info0["nsolid"] and info1["bbox"] == info0["bbox"]
            lat.set_state(pops)
            lat.step(50)
            states.append(lat.get_state())
    assert np.array_equal(states[0].view(np.uint64), states[1].view(np.uint64))


def test_drag_box_ending_at_plane_zero(golden):
    """A control box whose x range ends at 0 puts its second face on plane 0.  Parity mode holds
    the reference's ghost plane there and equals the oracle's flux_X / flux_Y bit for bit; the
    physics and grey layouts have no plane 0 and refuse the box (EUP_ERANGE)."""
    from euplectella_hpc_amd import geometry
    from euplectella_hpc_amd.lib import EupError, ERRORS
    box = [-3, 0, 2, 6, 2, 10]
    rec, mask, pops, lat = parity_lattice("c0_tiny", golden)
    announce("drag box ending at plane 0: parity vs oracle, physics and grey refuse")
    with lat:
        lat.step(5)
        state = lat.get_state()
        fx, fy = lat.drag_cv(box)
        omega = lat.omega
    assert fx == orc.flux_x(state, omega, box) and fy == orc.flux_y(state, omega, box)
    l, m, n = 16, 8, 12
    p0 = geometry.initial_state(l, m, n, 0.04, noise=1e-4, seed=3)
    for mode in ("physics", "grey"):
        with lattice((l, m, n), svisc=0.05, u_inflow=0.04, mode=mode, storage="f64") as lat:
            if mode == "grey":
                lat.set_grey(np.zeros((n, m, l)))
            else:
                lat.set_mask(np.zeros((n, m, l), np.uint8))
            lat.set_state(p0)
            with pytest.raises(EupError) as err:
                lat.drag_cv(box)
            assert ERRORS[err.value.code] == "EUP_ERANGE"
            fx, fy = lat.drag_cv([2, 9, 2, 6, 2, 10])   # a box inside the lattice still works
            assert np.isfinite(fx) and np.isfinite(fy)
