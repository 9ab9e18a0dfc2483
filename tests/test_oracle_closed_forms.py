"""Closed-form and invariant pins of the C oracle.

The reference's own executable pins the parity step, the drag and the probe bit for bit
(test_oracle_golden.py).  The functions it does not exercise -- the physics-mode step, the
momentum-exchange force, the global means -- and the individual phases are pinned here by what
the mathematics fixes: conserved moments, fixed points, exact permutations, the BGK relaxation of
the second moment, the shear-wave decay rate, the Poiseuille profile and hand-countable forces.
"""
import numpy as np
import pytest

from euplectella_hpc_amd import d3q19, geometry
from oracle import oracle as orc

C = d3q19.C.astype(np.float64)
W = d3q19.W


def eq(rho, u):
    u = np.asarray(u, dtype=np.float64)
    return d3q19.equilibrium(rho, u[0], u[1], u[2])


def uniform_state(l, m, n, rho, u):
    e = eq(rho, u)
    return np.ascontiguousarray(np.broadcast_to(e[:, None, None, None], (19, n + 2, m + 2, l + 2)))


# --------------------------------------------------------------------------------------------
# collision (col.F90:127-229)
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(5))
def test_collision_conserves_mass_and_momentum(seed):
    rng = np.random.default_rng(seed)
    x = eq(1.0 + 0.05 * rng.standard_normal(), 0.08 * rng.standard_normal(3)) \
        * (1 + 0.05 * rng.standard_normal(19))
    out = orc.collide_site(x, omega=1.7)
    assert abs(out.sum() - x.sum()) < 1e-15
    assert np.abs(C.T @ out - C.T @ x).max() < 1e-15


def test_collision_force_adds_ten_fgrad_rho_to_x_momentum():
    x = eq(1.02, (0.03, -0.01, 0.02))
    g = 1e-5
    out = orc.collide_site(x, omega=1.2, fgrad=g)
    # F_q = +/- fgrad*rho on the ten populations with c_x = +/-1 (FORCING_X)
    assert abs((C[:, 0] @ out) - (C[:, 0] @ x) - 10 * g * x.sum()) < 1e-15
    assert abs(out.sum() - x.sum()) < 1e-15


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.05, 0.0, 0.0), (0.03, -0.02, 0.04)])
def test_equilibrium_is_a_fixed_point(u):
    e = eq(1.0, u)
    out = orc.collide_site(e, omega=1.9)
    assert np.abs(out - e).max() < 1e-15


@pytest.mark.parametrize("seed", range(3))
def test_second_moment_relaxes_by_one_minus_omega(seed):
    """Pi' - rho(uu + I/3) = (1 - omega)(Pi - rho(uu + I/3)) for the D3Q19 2nd-order equilibrium."""
    rng = np.random.default_rng(seed)
    x = eq(1.0, 0.05 * rng.standard_normal(3)) * (1 + 0.02 * rng.standard_normal(19))
    om = 1.6
    out = orc.collide_site(x, omega=om)
    rho = x.sum()
    u = (C.T @ x) / rho
    pieq = rho * (np.outer(u, u) + np.eye(3) / 3)
    pi = np.einsum("q,qa,qb->ab", x, C, C)
    pi2 = np.einsum("q,qa,qb->ab", out, C, C)
    assert np.abs((pi2 - pieq) - (1 - om) * (pi - pieq)).max() < 1e-15


# --------------------------------------------------------------------------------------------
# phases of the parity step
# --------------------------------------------------------------------------------------------

def test_streaming_is_the_pull_permutation():
    l, m, n = 5, 4, 3
    a = np.arange(19 * (n + 2) * (m + 2) * (l + 2), dtype=np.float64).reshape(19, n + 2, m + 2, l + 2)
    b = np.full_like(a, -1.0)
    orc.phase_movef(a, b)
    for q in range(18):
        cx, cy, cz = d3q19.C[q]
        for k in range(1, n + 1):
            for j in range(1, m + 1):
                for i in range(1, l + 1):
                    assert b[q, k, j, i] == a[q, k - cz, j - cy, i - cx]
    assert (b[18] == -1).all()


def test_bounce_back_single_solid_site():
    l = m = n = 4
    a = np.random.default_rng(0).random((19, n + 2, m + 2, l + 2))
    before = a.copy()
    mask = np.zeros((n, m, l), np.uint8)
    mask[1, 2, 1] = 1                      # site (i, j, k) = (2, 3, 2)
    orc.phase_bcond_obs(a, mask)
    i, j, k = 2, 3, 2
    for q in range(18):
        p = d3q19.OPP[q]
        cx, cy, cz = d3q19.C[q]
        assert a[q, k, j, i] == before[p, k + cz, j + cy, i + cx]
    changed = (a != before)
    changed[:, k, j, i] = False
    assert not changed.any()


def test_ghost_fill_inlet_outlet_and_periodic_copies():
    l, m, n = 6, 5, 4
    a = np.random.default_rng(1).random((19, n + 2, m + 2, l + 2)) * 0.05
    before = a.copy()
    u_in = 0.07
    orc.phase_bcond_bc(a, u_in)
    ein = eq(1.0, (u_in, 0, 0))
    for q in range(5):                                   # inlet: q = 1..5 on the whole x = 0 plane
        mism = np.abs(a[q, :, :, 0] - ein[q])
        if q == 2:                                       # a03(0,0,k) <- a03(l,m,k) (bcond.bc.F90:547)
            assert np.array_equal(a[2, 1:n + 1, 0, 0], before[2, 1:n + 1, m, l])
            mism[1:n + 1, 0] = 0
        assert mism.max() < 1e-17
    for k in range(1, n + 1):                            # outlet: eq(rho=1, u = momentum at x = l)
        for j in range(1, m + 1):
            mom = C.T @ before[:, k, j, l]
            e = eq(1.0, mom)
            for t, q in enumerate(range(9, 14)):
                if q == 9 and j == m + 1:
                    continue
                assert abs(a[q, k, j, l + 1] - e[q]) < 1e-16
    # z and y periodic copies over the interior range
    for q in (1, 8, 10, 14, 15):
        assert np.array_equal(a[q, n + 1, 1:m + 1, 1:l + 1], before[q, 1, 1:m + 1, 1:l + 1])
    for q in (3, 5, 6, 12, 17):
        assert np.array_equal(a[q, 0, 1:m + 1, 1:l + 1], before[q, n, 1:m + 1, 1:l + 1])
    for q in (2, 6, 7, 8, 11):
        assert np.array_equal(a[q, 1:n + 1, 0, 1:l + 1], before[q, 1:n + 1, m, 1:l + 1])
    for q in (0, 9, 15, 16, 17):
        assert np.array_equal(a[q, 1:n + 1, m + 1, 1:l + 1], before[q, 1:n + 1, 1, 1:l + 1])
    # the four y-z edge ghost lines the serial path never writes (frozen values)
    for q, jj, kk in ((15, m + 1, n + 1), (8, 0, n + 1), (6, 0, 0), (17, m + 1, 0)):
        assert np.array_equal(a[q, kk, jj, 1:l + 1], before[q, kk, jj, 1:l + 1])


# --------------------------------------------------------------------------------------------
# physics-mode step
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.04, 0.01, -0.02)])
def test_physics_uniform_state_is_stationary_in_a_periodic_box(u):
    l, m, n = 6, 5, 4
    p = uniform_state(l, m, n, 1.0, u)
    ref = p.copy()
    orc.physics_steps(p, np.zeros((n, m, l), np.uint8), 1.7, 0.0, 10, x_open=False)
    assert np.abs(p - ref)[:, 1:-1, 1:-1, 1:-1].max() < 2e-15


def test_physics_mass_is_conserved_with_obstacles_in_a_periodic_box():
    l, m, n = 16, 12, 10
    mask = geometry.cylinder(l, m, n, 3, (6, 6))
    p = geometry.initial_state(l, m, n, 0.05, noise=1e-3, seed=2)
    fluid = mask == 0
    m0 = p[:, 1:-1, 1:-1, 1:-1][:, fluid].sum()
    orc.physics_steps(p, mask, 1.5, 0.0, 50, x_open=False)
    m1 = p[:, 1:-1, 1:-1, 1:-1][:, fluid].sum()
    assert abs(m1 - m0) < 1e-11 * m0


def test_physics_shear_wave_decays_at_the_lattice_viscosity():
    l, m, n = 4, 32, 4
    tau = 0.8
    nu = (tau - 0.5) / 3
    k = 2 * np.pi / m
    amp = 1e-3
    jj = np.arange(m + 2) - 0.5   # ghost-padded j index; interior j = 1..m -> y = j - 0.5
    ux = amp * np.sin(k * jj)
    p = np.empty((19, n + 2, m + 2, l + 2))
    for j in range(m + 2):
        p[:, :, j, :] = eq(1.0, (ux[j], 0, 0))[:, None, None]
    steps = 200
    orc.physics_steps(p, np.zeros((n, m, l), np.uint8), 1 / tau, 0.0, steps, x_open=False)
    rho = p.sum(axis=0)
    u = (np.tensordot(C[:, 0], p, axes=(0, 0)) / rho)[1:-1, 1:-1, 1:-1]
    prof = u[:, :, 0].mean(axis=0)
    a_num = 2 * np.mean(prof * np.sin(k * (np.arange(1, m + 1) - 0.5)))
    assert abs(a_num / amp - np.exp(-nu * k * k * steps)) < 2e-3


def test_physics_poiseuille_profile_between_bounce_back_walls():
    """Body force 10*fgrad per unit mass between solid rows j = 1 and j = m (walls halfway)."""
    l, m, n = 4, 20, 4
    tau = 1.0
    nu = (tau - 0.5) / 3
    g = 1e-6
    mask = np.zeros((n, m, l), np.uint8)
    mask[:, 0, :] = 1
    mask[:, m - 1, :] = 1
    p = uniform_state(l, m, n, 1.0, (0, 0, 0))
    orc.physics_steps(p, mask, 1 / tau, 0.0, 6000, fgrad=g, x_open=False)
    rho = p.sum(axis=0)
    ux = (np.tensordot(C[:, 0], p, axes=(0, 0)) / rho)[1:-1, 1:-1, 1:-1][0, :, 0]
    y = np.arange(1, m + 1, dtype=float)
    y0, y1 = 1.5, m - 0.5
    exact = 10 * g / (2 * nu) * (y - y0) * (y1 - y)
    inner = slice(1, m - 1)
    assert np.abs(ux[inner] - exact[inner]).max() < 0.02 * exact.max()


# --------------------------------------------------------------------------------------------
# momentum exchange and global means
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.05, 0.0, 0.0), (0.02, -0.03, 0.01)])
def test_mem_force_single_solid_site_in_uniform_equilibrium_is_2_rho_u(u):
    """F = sum_p 2 c_p w_p rho (1 + 3 c_p.u + ...) = 2 rho u (odd lattice moments vanish)."""
    l = m = n = 5
    rho = 1.03
    p = uniform_state(l, m, n, rho, u)
    mask = np.zeros((n, m, l), np.uint8)
    mask[2, 2, 2] = 1
    F = orc.mem_force(p, mask)
    assert np.abs(F - 2 * rho * np.asarray(u)).max() < 1e-15


def test_mem_force_symmetric_body_at_rest_is_zero():
    l, m, n = 20, 16, 6
    mask = geometry.cylinder(l, m, n, 4, (10.5, 8.5))
    p = uniform_state(l, m, n, 1.0, (0, 0, 0))
    assert np.abs(orc.mem_force(p, mask)).max() < 1e-15


def test_diagno_closed_forms():
    l, m, n = 8, 6, 5
    u = (0.03, -0.01, 0.02)
    rho = 1.01
    p = uniform_state(l, m, n, rho, u)
    out = orc.diagno(p, np.zeros((n, m, l), np.uint8))
    # sequential sums of l*m*n terms: a few ulp per term
    assert np.abs(out - [rho, *u, np.dot(u, u)]).max() < 1e-14
    mask = np.zeros((n, m, l), np.uint8)
    mask[:, :, :2] = 1                     # solid fraction 1/4: rho = 1, u = 0 there
    phi = 0.25
    out = orc.diagno(p, mask)
    expect = [phi + (1 - phi) * rho, *(np.asarray(u) * (1 - phi)), (1 - phi) * np.dot(u, u)]
    assert np.abs(out - expect).max() < 1e-14


def test_probe_closed_form():
    p = uniform_state(6, 5, 4, 0.98, (0.04, 0.01, -0.02))
    out = orc.probe(p, 3, 2, 2)
    assert np.abs(out - [0.04, 0.01, -0.02, 0.98]).max() < 1e-16
    assert orc.probe(p, 7, 2, 2) is None


def test_flux_x_uniform_equilibrium_gives_pressure_only():
    """With f = f_eq, f_neq = 0 and sigma_xx = -rho/3 on both faces: fluxX = 0 per face pair,
    and the single-face value is +/- (#sites) rho/3."""
    l, m, n = 10, 8, 4
    rho = 1.02
    p = uniform_state(l, m, n, rho, (0.05, 0.0, 0.0))
    box = (3, 7, 2, 6, 1, n)
    assert abs(orc.flux_x(p, 1.5, box)) < 1e-14
    assert abs(orc.flux_y(p, 1.5, box)) < 1e-15
    # only face 2 inside (case B) -> norm(2) = 0 : fluxX = -F(i1) = +(#sites)*rho/3
    boxb = (3, l + 5, 2, 6, 1, n)
    assert abs(orc.flux_x(p, 1.5, boxb) - 5 * n * rho / 3) < 1e-13


def test_mem_force_closes_the_momentum_balance():
    """orc_mem_force against the one-step momentum balance of orc_physics_steps around a
    mirror-symmetric cylinder (Re = 2): the balance closes to rounding and there is no lift."""
    from momentum_balance import box_momentum
    from euplectella_hpc_amd import geometry
    l, m, n = 96, 48, 4
    mask = geometry.cylinder(l, m, n, 5, (24, (m + 1) / 2))
    P = geometry.initial_state(l, m, n, 0.02, noise=0.0)
    om = orc.omega_from_svisc((2 * 0.8 - 1) / 6)
    orc.physics_steps(P, mask, om, 0.02, 300)
    P0 = P.copy()
    F = orc.mem_force(P0, mask)
    orc.physics_steps(P, mask, om, 0.02, 1)
    for box in ((8, 60, 8, 41), (14, 34, 14, 35)):
        M0, G = box_momentum(P0, box)
        M1, _ = box_momentum(P, box)
        assert np.abs((M1 - M0) - (G - F)).max() < 1e-10 * abs(F[0])
    assert F[0] > 0 and abs(F[1]) < 1e-12 * F[0] and abs(F[2]) < 1e-12 * F[0]
