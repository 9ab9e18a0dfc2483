"""bench.py command-line contract that does not need a GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU only" in res.stderr
    assert res.stdout.strip() == ""


def test_workloads_define_the_configured_lattices():
    sys.path.insert(0, ROOT)
    import bench

    wl = bench.WORKLOADS["sponge_512x256x256"]
    assert (wl["l"], wl["m"], wl["n"], wl["tau"], wl["u_in"]) == (512, 256, 256, 0.505, 0.06)
    c5 = bench.WORKLOADS["sponge_1536x768x768"]
    assert (c5["l"], c5["m"], c5["n"], c5["radius"], c5["cell"]) == (1536, 768, 768, 144, 48)
    c1 = bench.WORKLOADS["cylinder_128x64x64"]
    assert (c1["radius"], c1["centre"], c1["tau"], c1["u_in"]) == (8, (32, 32), 0.6, 0.05)


def test_cpu_baseline_leg_runs_the_reference():
    """The CPU-baseline leg (the reference's OpenMP build, bounded sample) on a tiny lattice."""
    from oracle import refrun
    from euplectella_hpc_amd import geometry
    if refrun.ref_binary("omp") is None:
        pytest.skip("the reference build is not present")
    mask = geometry.cylinder(24, 12, 8, 2, (6, 6.5))
    r = refrun.time_reference(24, 12, 8, (2 * 0.6 - 1) / 6, 0.05, 5, mask, threads=2)
    assert r["mlups"] > 0 and r["threads"] == 2


def test_seed_distribution_counts_and_orders():
    sys.path.insert(0, ROOT)
    import bench

    windows = [{"f16": {"rel_dcd": v}, "f32": {"rel_dcd": 1e-5}} for v in (2e-3, 5e-4, 1e-3, 8e-3)]
    d = bench.seed_distribution(windows, ["f16", "f32", bench.FLOOR])
    assert d["f16"] == {"runs": 4, "met": 2, "min": 5e-4, "median": 1.5e-3, "max": 8e-3}
    assert d["f32"]["met"] == 4 and d["f32"]["runs"] == 4
    assert bench.FLOOR not in d  # no window held it
