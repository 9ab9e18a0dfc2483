"""GPU grey lattice and discrete adjoint (through the C-ABI and the torch binding).

* grey FP64 step == oracle grey step bit for bit; a binary field == physics mode on fluid sites;
* objective and gradient (FP64) match the oracle's adjoint;
* the FP32 adjoint matches the FP64 adjoint, and the FP64 adjoint matches central finite
  differences of the FP64 objective on 64 cells (relative 1e-4 is the stated tolerance);
* the fused adjoint pass equals the split kernels bit for bit;
* the torch.autograd.Function returns the same gradient.
"""
import numpy as np
import pytest

from euplectella_hpc_amd import geometry
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)


def lattice(shape, **kw):
    from euplectella_hpc_amd.lib import Lattice
    return Lattice(*shape, **kw)


def dense(p):
    return np.ascontiguousarray(p[:, 1:-1, 1:-1, 1:-1])


def grey_inputs(l, m, n, seed, n_spheres=4, rmin=2.0, rmax=4.0, sigma=1.0):
    s = geometry.grey_spheres(l, m, n, n_spheres=n_spheres, rmin=rmin, rmax=rmax, sigma=sigma,
                              seed=seed)
    pops = geometry.initial_state(l, m, n, 0.05, seed=seed)
    return s, pops


@pytest.mark.parametrize("x_open", [True, False])
def test_grey_f64_step_equals_oracle(x_open):
    l, m, n = 40, 20, 16
    announce(f"grey f64 step {l}x{m}x{n} x_open={x_open} vs oracle, 20 steps")
    s, pops = grey_inputs(l, m, n, seed=1)
    svisc = (2 * 0.6 - 1) / 6
    ref = dense(pops)
    orc.grey_steps(ref, s, orc.omega_from_svisc(svisc), 0.05, 20, x_open=x_open)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="grey", storage="f64",
                 x_open=x_open) as lat:
        lat.set_grey(s)
        lat.set_state(pops)
        lat.step(20)
        got = dense(lat.get_state())
        F = lat.grey_force()
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    Fr = orc.grey_force(ref, s, 0.05, x_open=x_open)
    assert np.abs(F - Fr).max() <= 1e-12 * np.abs(Fr).max()


def test_binary_grey_field_equals_physics_mode_on_fluid_sites():
    l, m, n = 48, 24, 16
    announce("grey f64 with a binary field vs physics f64")
    mask = geometry.cylinder(l, m, n, 4, (12, 12))
    pops = geometry.initial_state(l, m, n, 0.05, seed=2)
    svisc = (2 * 0.55 - 1) / 6
    out = {}
    for mode in ("grey", "physics"):
        with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode=mode, storage="f64") as lat:
            if mode == "grey":
                lat.set_grey(mask.astype(np.float64))
            else:
                lat.set_mask(mask)
            lat.set_state(pops)
            lat.step(15)
            out[mode] = dense(lat.get_state())
    fluid = mask == 0
    assert np.array_equal(out["grey"][:, fluid], out["physics"][:, fluid])


@pytest.mark.parametrize("x_open", [True, False])
def test_objective_and_gradient_f64_match_oracle(x_open):
    l, m, n = 24, 12, 10
    announce(f"grey f64 objective + adjoint {l}x{m}x{n} x_open={x_open} vs oracle adjoint")
    s, pops = grey_inputs(l, m, n, seed=3, n_spheres=3)
    svisc = (2 * 0.6 - 1) / 6
    om = orc.omega_from_svisc(svisc)
    t0, t1, K = 10, 30, 7
    Jr, gr = orc.grey_objective(dense(pops), s, om, 0.05, t0, t1, K, 1.0, x_open=x_open)
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="grey", storage="f64",
                 x_open=x_open) as lat:
        lat.set_grey(s)
        lat.set_state(pops)
        J, g = lat.grey_objective(t0, t1, K, 1.0)
        # the call leaves the state at a^0
        assert np.array_equal(dense(lat.get_state()), dense(pops))
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    assert np.abs(g - gr).max() <= 1e-10 * np.abs(gr).max()


def test_objective_windows_at_the_edges_match_oracle():
    """Windows of one step, checkpoint intervals longer than the run, a window starting at step 1."""
    l, m, n = 16, 8, 6
    s, pops = grey_inputs(l, m, n, seed=9, n_spheres=2, rmin=1.5, rmax=2.5)
    svisc = (2 * 0.6 - 1) / 6
    om = orc.omega_from_svisc(svisc)
    # K = 10^7: a segment holds min(K, t1) states, not K (K states of this lattice would be 1.2 TB)
    for t0, t1, K in [(1, 1, 1), (1, 1, 4), (1, 3, 2), (2, 9, 9), (2, 9, 20), (2, 9, 10_000_000)]:
        announce(f"grey f64 objective + adjoint t0={t0} t1={t1} K={K} vs oracle")
        Jr, gr = orc.grey_objective(dense(pops), s, om, 0.05, t0, t1, K, 1.0)
        with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="grey", storage="f64") as lat:
            lat.set_grey(s)
            lat.set_state(pops)
            J, g = lat.grey_objective(t0, t1, K, 1.0)
        assert abs(J - Jr) <= 1e-12 * abs(Jr)
        assert np.abs(g - gr).max() <= 1e-10 * np.abs(gr).max()


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("x_open", [True, False])
def test_fused_adjoint_sweep_equals_split_kernels(storage, x_open, monkeypatch):
    """The fused stream + collide adjoint pass gives the split kernels' objective and gradient bit
    for bit, for segments of one step, several steps with a partial last one, and one segment."""
    l, m, n = 24, 12, 10
    s, pops = grey_inputs(l, m, n, seed=5, n_spheres=3)
    svisc = (2 * 0.6 - 1) / 6
    t0, t1 = 10, 30
    for K in (1, 7, 31):
        announce(f"grey {storage} x_open={x_open} adjoint fused vs split, K={K}")
        out = {}
        for fused in ("0", "1"):
            monkeypatch.setenv("EUP_ADJ_FUSED", fused)
            with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="grey", storage=storage,
                         x_open=x_open) as lat:
                lat.set_grey(s)
                lat.set_state(pops)
                out[fused] = lat.grey_objective(t0, t1, K, 1.0)
        assert out["0"][0] == out["1"][0]
        assert np.array_equal(out["0"][1].view(np.uint64), out["1"][1].view(np.uint64))
        assert np.abs(out["1"][1]).max() > 0


C4 = dict(l=128, m=64, n=64, tau=0.6, u_in=0.05, t0=1000, t1=2000, every=100, seed=44)


def c4_inputs():
    c = C4
    s = geometry.grey_spheres(c["l"], c["m"], c["n"], n_spheres=12, rmin=4.0, rmax=8.0, sigma=1.5,
                              seed=c["seed"])
    pops = geometry.initial_state(c["l"], c["m"], c["n"], c["u_in"], seed=c["seed"])
    # frontal area of the cluster: projection of s onto the y-z plane
    area = float(s.max(axis=2).sum())
    norm = 0.5 * 1.0 * c["u_in"] ** 2 * area
    return s, pops, norm


def c4_lattice(storage):
    from euplectella_hpc_amd.lib import svisc_from_tau
    c = C4
    return lattice((c["l"], c["m"], c["n"]), svisc=svisc_from_tau(c["tau"]), u_inflow=c["u_in"],
                   mode="grey", storage=storage)


def test_c4_fp32_adjoint_matches_fp64_adjoint_and_finite_differences():
    """C4: 128x64x64, 12 blurred spheres, mean C_D over steps 1000-2000, checkpoint every 100;
    central differences on 64 random cells of the grey transition zone."""
    c = C4
    announce("C4 grey 128x64x64: FP64 and FP32 objective + adjoint, 2000 steps")
    s, pops, norm = c4_inputs()
    grads = {}
    with c4_lattice("f64") as lat:
        lat.set_grey(s)
        lat.set_state(pops)
        J64, grads["f64"] = lat.grey_objective(c["t0"], c["t1"], c["every"], norm)
        # central differences of the FP64 objective on cells inside the grey transition zone
        g = grads["f64"]
        rng = np.random.default_rng(0)
        # cells whose central-difference points stay inside [0, 1], with a gradient of at least
        # 1 % of the largest one (the largest gradients sit at the fluid edge of the blur, s << 1)
        sf = s.ravel()
        eps = 1e-5
        cand = np.flatnonzero((np.abs(g.ravel()) > 0.01 * np.abs(g).max()) & (sf > 4 * eps) &
                              (sf < 1 - 4 * eps))
        assert cand.size >= 64
        cells = rng.choice(cand, 64, replace=False)
        worst = 0.0
        fds = []
        for cidx in cells:
            Js = []
            for sign in (1, -1):
                sp = sf.copy()
                sp[cidx] += sign * eps
                lat.set_grey(sp.reshape(s.shape))
                Js.append(lat.grey_objective(c["t0"], c["t1"], c["every"], norm, grad=False)[0])
            fd = (Js[0] - Js[1]) / (2 * eps)
            fds.append(fd)
            rel = abs(fd - g.ravel()[cidx]) / abs(g.ravel()[cidx])
            announce(f"  cell {cidx}: s={sf[cidx]:.4f} adjoint {g.ravel()[cidx]:.8e} fd {fd:.8e} rel {rel:.2e}")
            worst = max(worst, rel)
        announce(f"C4 FD check: J={J64:.6e}, worst relative error {worst:.2e} on {len(cells)} cells")
        assert worst < 1e-4
    with c4_lattice("f32") as lat:
        lat.set_grey(s)
        lat.set_state(pops)
        J32, grads["f32"] = lat.grey_objective(c["t0"], c["t1"], c["every"], norm)
    announce(f"C4 FP32 J={J32:.6e} vs FP64 {J64:.6e}")
    assert abs(J32 - J64) < 1e-3 * abs(J64)
    g64, g32 = grads["f64"], grads["f32"]
    big = np.abs(g64) > 0.05 * np.abs(g64).max()
    rel = np.abs(g32[big] - g64[big]) / np.abs(g64[big])
    assert np.median(rel) < 1e-3 and np.abs(g32 - g64).max() < 1e-2 * np.abs(g64).max()
    # the FP32 gradient on the 64 cells of the FD check, against the FP64 adjoint and against the
    # central differences of the FP64 objective (bound 5e-4: FP32 rounding of a 2000-step forward
    # and adjoint sweep; measured on MI355X worst 6.6e-5 / 6.6e-5, median 1.9e-5)
    g32c, g64c, fds = g32.ravel()[cells], g64.ravel()[cells], np.array(fds)
    rel64 = np.abs(g32c - g64c) / np.abs(g64c)
    relfd = np.abs(g32c - fds) / np.abs(fds)
    announce(f"C4 FP32 gradient on the 64 cells: worst relative error {rel64.max():.2e} vs the FP64 "
             f"adjoint, {relfd.max():.2e} vs central differences (median {np.median(rel64):.2e})")
    assert rel64.max() < 5e-4 and relfd.max() < 5e-4


def test_torch_autograd_function_returns_the_adjoint_gradient():
    import torch
    from euplectella_hpc_amd.autograd import grey_drag

    l, m, n = 32, 16, 12
    announce("torch GreyDrag autograd function vs C-ABI gradient")
    s_np, pops = grey_inputs(l, m, n, seed=5, n_spheres=3)
    svisc = (2 * 0.6 - 1) / 6
    with lattice((l, m, n), svisc=svisc, u_inflow=0.05, mode="grey", storage="f64") as lat:
        lat.set_grey(s_np)
        lat.set_state(pops)
        Jc, gc = lat.grey_objective(8, 20, 5, 1.0)
        s = torch.tensor(s_np, dtype=torch.float64, device="cuda", requires_grad=True)
        J = grey_drag(s, lat, 8, 20, 5, 1.0)
        J.backward()
    assert abs(J.item() - Jc) <= 1e-14 * abs(Jc)
    assert np.array_equal(s.grad.cpu().numpy(), gc)
