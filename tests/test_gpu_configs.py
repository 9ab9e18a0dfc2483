"""The benchmark configurations at their own sizes on the GPU, through the C-ABI.

* C2 ridged cylinder 256x128x128 and C3 reticular sponge 512x256x256: after the start-up ramp the
  benchmark runs (tau 0.6 / 0.55 / 0.52 stages) and 100 steps at the configuration's own tau, the
  FP32, FP16, BF16 and e3m12 storages are compared with FP64 physics mode on every fluid site and in the
  momentum-exchange drag coefficient C_D = F_x / (0.5 u_in^2 D l_z).  These are implementation
  checks at full size (tile rows of 4 and 8 tiles, every mixed tile of the geometry); the accuracy
  of the time-mean drag is measured by bench.py.
* C5 capacity run 1536x768x768 (906 M sites; FP32 state 138 GB, FP16 69 GB): FP32 and FP16 one
  after the other, 30 steps from the equilibrium start at tau 0.502, checked through properties:
  the tile classes add up and match the mask, the state is finite at 65 sampled fluid sites, the
  global mean density stays within 1e-3 of 1 (diagno), and FP16 agrees with FP32 at the sampled
  sites within the tolerance below.
"""
import sys

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def drag_coefficient(lat, wl):
    return lat.drag_mem()[0] / (0.5 * wl["u_in"] ** 2 * wl["diameter"] * wl["n"])


def compare_with_f64(name, storages, steps, pop_tol, cd_tol):
    wl = bench.WORKLOADS[name]
    mask = bench.workload_mask(wl)
    fluid = mask == 0
    lats = {}
    try:
        for st in ["f64"] + storages:
            announce(f"{name}: {st} start-up ramp + {steps} steps at tau {wl['tau']}")
            lat = bench.build_lattice(wl, st, mask)
            lat.step(steps)
            lats[st] = lat
        cd = {st: drag_coefficient(lat, wl) for st, lat in lats.items()}
        err = dict.fromkeys(storages, 0.0)
        for q in range(1, 20):
            ref = lats["f64"].get_pops(q)[1:-1, 1:-1, 1:-1][fluid]
            assert np.all(np.isfinite(ref))
            for st in storages:
                got = lats[st].get_pops(q)[1:-1, 1:-1, 1:-1][fluid]
                err[st] = max(err[st], float(np.abs(got - ref).max()))
    finally:
        for lat in lats.values():
            lat.close()
    dcd = {st: abs(cd[st] - cd["f64"]) / abs(cd["f64"]) for st in storages}
    for st in storages:
        announce(f"  {st}: max |df| = {err[st]:.3e}, |dC_D|/C_D = {dcd[st]:.3e} (C_D f64 = {cd['f64']:.5f})")
    for st in storages:
        assert err[st] < pop_tol[st], (st, err[st])
        assert dcd[st] < cd_tol[st], (st, dcd[st])


@pytest.mark.timeout(600)
def test_c2_ridged_256x128x128_reduced_storage_against_fp64():
    # measured on MI355X: max |df| 1.4e-6 / 8.4e-5 / 9.8e-4, |dC_D|/C_D 2.3e-5 / 2.5e-4 / 5.7e-3
    compare_with_f64("ridged_256x128x128", ["f32", "f16", "bf16"], 100,
                     pop_tol={"f32": 1e-5, "f16": 5e-4, "bf16": 5e-3},
                     cd_tol={"f32": 2e-4, "f16": 2e-3, "bf16": 3e-2})


@pytest.mark.timeout(600)
def test_c3_sponge_512x256x256_reduced_storage_against_fp64():
    # measured on MI355X: max |df| 2.5e-6 / 2.4e-4 / 5.8e-5, |dC_D|/C_D 3.5e-5 / 6.2e-4 / 5.9e-5
    # (FP32 / FP16 / e3m12)
    compare_with_f64("sponge_512x256x256", ["f32", "f16", "e3m12"], 100,
                     pop_tol={"f32": 2e-5, "f16": 1e-3, "e3m12": 3e-4},
                     cd_tol={"f32": 2e-4, "f16": 3e-3, "e3m12": 3e-4})


@pytest.mark.timeout(900)
def test_c5_sponge_1536x768x768_fp32_and_fp16_at_full_size():
    from euplectella_hpc_amd.lib import Lattice, svisc_from_tau

    wl = bench.WORKLOADS["sponge_1536x768x768"]
    l, m, n = wl["l"], wl["m"], wl["n"]
    announce("C5 1536x768x768: generating the sponge")
    mask = bench.workload_mask(wl)
    nsolid = int(mask.sum())
    rng = np.random.default_rng(5)
    sites = [(l * 8 // 20, m * 9 // 16, n // 2 - 5)]  # the reference's probe site
    while len(sites) < 65:
        i, j, k = (int(rng.integers(1, l + 1)), int(rng.integers(1, m + 1)), int(rng.integers(1, n + 1)))
        if mask[k - 1, j - 1, i - 1] == 0:
            sites.append((i, j, k))
    steps = 30
    probes = {}
    for st in ("f32", "f16"):
        announce(f"C5 {st}: create, mask, equilibrium start, {steps} steps at tau {wl['tau']}")
        with Lattice(l, m, n, svisc=svisc_from_tau(wl["tau"]), u_inflow=wl["u_in"], mode="physics",
                     storage=st) as lat:
            lat.set_mask(mask)
            info = lat.info()
            assert info["nsolid"] == nsolid
            assert info["tiles"] == ((l + 63) // 64) * m * n
            assert info["tiles_fluid"] + info["tiles_mixed"] + info["tiles_solid"] == info["tiles"]
            assert info["tiles_mixed"] > 0
            lat.init_equilibrium(wl["u_in"], 0.0, 0.0, 1e-4, wl["seed"])
            lat.step(steps)
            lat.sync()
            probes[st] = np.array([lat.probe(*s) for s in sites])
            mean = lat.diagno()
        announce(f"  {st}: tiles {info['tiles_fluid']} fluid / {info['tiles_mixed']} mixed / "
                 f"{info['tiles_solid']} solid, <rho> = {mean[0]:.9f}, <u> = {mean[1]:.6f}")
        assert np.all(np.isfinite(probes[st]))
        assert abs(mean[0] - 1.0) < 1e-3
        assert 0.0 < mean[1] < 2 * wl["u_in"]
    du = np.abs(probes["f16"][:, :3] - probes["f32"][:, :3]).max()
    drho = np.abs(probes["f16"][:, 3] - probes["f32"][:, 3]).max()
    announce(f"C5 FP16 vs FP32 at {len(sites)} sites: max |du| = {du:.3e}, max |drho| = {drho:.3e}")
    # measured on MI355X: 4.5e-5 / 8.4e-5
    assert du < 5e-4 and drho < 5e-4


@pytest.mark.timeout(300)
def test_accuracy_window_on_the_cylinder():
    """bench.py's accuracy window on the C1 cylinder (128x64x64, tau 0.6), 400 steps from the
    seeded start, in two noise realisations: FP64 stays finite, the floor (FP64 from its
    FP32-rounded start) is far below FP32's error, FP32 and FP16 meet the bound, and the
    distribution summary counts them."""
    wl = bench.WORKLOADS["cylinder_128x64x64"]
    mask = bench.workload_mask(wl)
    windows = []
    for k in range(2):
        announce(f"C1 accuracy window, 400 steps, noise seed {wl['seed'] + k}")
        windows.append(bench.accuracy(wl, ["f32", "f16"], 400, 20, mask=mask, quiet=True, floor=True,
                                      noise_seed=wl["seed"] + k))
    for w in windows:
        announce(f"  seed {w['noise_seed']}: " + ", ".join(f"{st} {w[st]['rel_dcd']:.2e}"
                                                          for st in ("f32", "f16", bench.FLOOR)))
        assert w["finite_f64"] and w["cd_f64"] > 0
        assert w["f32"]["met"] and w["f16"]["met"]
        assert w[bench.FLOOR]["rel_dcd"] < w["f32"]["rel_dcd"]
    # measured on MI355X: FP32 1.9e-6 / 2.0e-6, FP16 5.9e-4 / 5.2e-4, floor 8e-10 / 4e-9 (the runs
    # are deterministic: the same seeds give the same values)
    d = bench.seed_distribution(windows, ["f32", "f16", bench.FLOOR])
    assert d["f32"]["runs"] == 2 and d["f32"]["met"] == 2 and d["f16"]["met"] == 2
