"""Invariants that the mathematics fixes, checked on the CPU oracle for the
parts of the reference that cannot be compiled here (their modules need GSL):
the VR collision step, the inverse-CDF sampling, the 2-D and 1-D convection
and the wall boundary conditions. Each check is chosen so that a plausible
slip (a dropped term, a wrong sign, a swapped partner, a wrong upwind side)
fails it."""
import ctypes as C

import numpy as np
import pytest

from tests._oracle import load_oracle, oracle_solver


def moments_cons(f, v):
    return np.array([f.sum(), f @ v[:, 0], f @ v[:, 1], f @ v[:, 2], f @ (v ** 2).sum(1)])


@pytest.mark.parametrize("over", [{}, dict(num_repl=3), dict(omega=0.81, alpha=1.4),
                                  dict(sample_mode=0, coln_rms=0.05)])
def test_collision_step_conserves_mass_momentum_energy(pkg, over):
    cfg, f = pkg.homogeneous_relaxation(n_v=13, samples=20000)
    cfg = cfg.with_(**over)
    v = pkg.node_velocities(cfg)
    s = oracle_solver(cfg)
    s.initialize(f)
    for n in range(1, 4):
        a = moments_cons(s.download_f()[0], v)
        cnt = s.compute_collision_integral(n)
        assert cnt[0] > 0
        b = moments_cons(s.download_f()[0], v)
        s.calculate_physical_properties()
        # changes are pure round-off of the per-sample weights
        assert np.all(np.abs(b - a) <= 1e-13 * (np.abs(a) + 1.0)), (n, b - a)


def test_collision_relaxes_towards_equilibrium(pkg):
    """Two counter-streaming Maxwellians relax: the temperature anisotropy
    Txx - (Tyy+Tzz)/2 decays monotonically on average, density and energy fixed."""
    cfg, f = pkg.homogeneous_relaxation(n_v=15, samples=40000, dt=0.1)
    v = pkg.node_velocities(cfg)
    s = oracle_solver(cfg)
    s.initialize(f)

    def aniso(g):
        n = g.sum()
        u = (g @ v) / n
        c = v - u
        t = (g[:, None] * c ** 2).sum(0) / n
        return t[0] - 0.5 * (t[1] + t[2])

    a0 = aniso(s.download_f()[0])
    for n in range(1, 9):
        s.step(n)
    a1 = aniso(s.download_f()[0])
    assert a0 > 0.5 and a1 < 0.6 * a0, (a0, a1)


def test_sampling_histogram_follows_cdf(pkg, orc):
    """Picks of integral 1 (equilibrium partner) and 3 (deviation partners)
    follow their CDFs: chi-square against the node weights."""
    cfg, f = pkg.homogeneous_relaxation(n_v=7, samples=60000)
    s = oracle_solver(cfg)
    s.initialize(f)
    N = cfg.num_points
    eqc = np.zeros(N + 1); neqc = np.zeros(N + 1)
    tabs = [np.zeros(1000, np.int32) for _ in range(4)]
    cnt = np.zeros(3, np.int32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    I = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    assert orc.orc_cell_cdf(s.h, 0, P(eqc), P(neqc), *(I(t) for t in tabs), I(cnt)) == 0
    irec, _, n = s.collide_trace(1, 0, 200000)
    for t, col, cum in ((1, 2, eqc), (1, 3, neqc), (3, 2, neqc), (3, 3, neqc)):
        picks = irec[irec[:, 0] == t, col]
        p = np.diff(cum) / cum[-1]
        obs = np.bincount(picks, minlength=N)
        exp = p * len(picks)
        assert np.all(obs[p == 0] == 0)
        m = exp > 5
        chi2 = ((obs[m] - exp[m]) ** 2 / exp[m]).sum()
        dof = m.sum() - 1
        assert chi2 < dof + 6 * np.sqrt(2 * dof), (t, col, chi2, dof)


def test_interior_sweep_keeps_uniform_state(pkg):
    """A spatially uniform f is a fixed point of the interior donor-cell sweep."""
    cfg, f = pkg.lid_driven_cavity(nx=6, ny=6, n_v=7, vmax=3.0, samples=0)
    s = oracle_solver(cfg)
    rng = np.random.default_rng(0)
    g = np.tile(rng.random(cfg.num_points), (36, 1))
    s.initialize(g)
    s.compute_convection(1)
    out = s.download_f().reshape(6, 6, -1)
    assert np.array_equal(out[1:-1, 1:-1], g.reshape(6, 6, -1)[1:-1, 1:-1])


def test_interior_sweep_is_upwind_shift(pkg):
    """With |c_x| = 1 a node moving in +x takes its left neighbour's value and
    one moving in -x its right neighbour's (pins the upwind side)."""
    cfg, _ = pkg.lid_driven_cavity(nx=8, ny=8, n_v=5, vmax=2.0, samples=0)
    cfg = cfg.with_(dt=cfg.dx / 2.0, dy=1e30)  # x CFL = |v|/2: exactly 1 at |v| = 2; no y transport
    s = oracle_solver(cfg)
    rng = np.random.default_rng(3)
    g = rng.random((64, cfg.num_points))
    s.initialize(g)
    s.compute_convection(1)
    out = s.download_f().reshape(8, 8, 5, 5, 5)
    g = g.reshape(8, 8, 5, 5, 5)  # [iy][ix][k][j][i]
    assert np.allclose(out[2:-2, 2:-2, :, :, 4], g[2:-2, 1:-3, :, :, 4], rtol=0, atol=1e-15)
    assert np.allclose(out[2:-2, 2:-2, :, :, 0], g[2:-2, 3:-1, :, :, 0], rtol=0, atol=1e-15)
    assert np.array_equal(out[2:-2, 2:-2, :, :, 2], g[2:-2, 2:-2, :, :, 2])


@pytest.mark.parametrize("acc", [1.0, 0.0, 0.5])
def test_closed_box_conserves_mass(pkg, acc):
    """Walls at rest with full accommodation balance the incoming flux (diffuse
    ghost scaled by flux_in/flux_out); specular reflection on a symmetric grid
    maps each incoming node onto its mirror. Either way no mass crosses a wall."""
    cfg, f = pkg.lid_driven_cavity(nx=5, ny=4, n_v=9, vmax=3.0, samples=0, lid=0.0)
    cfg = cfg.with_(acc_coeff=acc)
    s = oracle_solver(cfg)
    rng = np.random.default_rng(11)
    g = f * (1.0 + 0.3 * rng.random(f.shape))
    s.initialize(g)
    m0 = s.download_f().sum()
    for n in range(1, 4):
        s.compute_convection(n)
    m1 = s.download_f().sum()
    assert abs(m1 - m0) <= 1e-13 * m0, (m1 - m0) / m0


def test_moving_lid_drives_flow(pkg):
    """The lid ghost carries x-momentum: after a few steps the top row moves in +x."""
    cfg, f = pkg.lid_driven_cavity(nx=6, ny=6, n_v=9, vmax=3.0, samples=200, lid=0.3)
    s = oracle_solver(cfg)
    s.initialize(f)
    for n in range(1, 6):
        s.step(n)
    p = s.calculate_physical_properties().reshape(6, 6, -1)
    assert np.all(p[-1, 1:-1, 1] > 1e-3) and abs(p[0, 2, 1]) < p[-1, 2, 1]


def test_inflow_cells_hold_the_wall_state(pkg):
    """Fixed in/out flow sets the end cells to the wall Maxwellian (n, u, T; v = w = 0)."""
    cfg, f = pkg.normal_shock_1d(nx=10, n_v=11, vmax=6.0, samples=200)
    s = oracle_solver(cfg)
    s.initialize(f)
    s.step(1)
    p = s.calculate_physical_properties()
    w = np.asarray(cfg.wall)
    for cell, side in ((0, 0), (9, 1)):
        assert abs(p[cell, 0] / w[side, 0] - 1) < 5e-3
        assert abs(p[cell, 1] - w[side, 1]) < 5e-3 * (1 + abs(w[side, 1]))
        assert abs(p[cell, 4] / w[side, 4] - 1) < 2e-2
