"""The oracle's 1-D sweep options pinned by closed forms and invariants:
the limited second-order upwind sweep (not in the reference: its accuracy
order, TVD property and exactness on linear data) and the 1-D diffuse /
specular wall (the outgoing half-Maxwellian scaled to the incoming density;
mirrored mass and momentum of the specular part)."""
import numpy as np
import pytest

from tests._oracle import oracle_solver


def channel(pkg, nx, n_v=3, vmax=1.0, dx=1.0, dt=0.5, order=1, bc=3, acc=1.0, wall=None):
    """1-D run with no collisions (coll_solver 0): every velocity node is
    advected on its own at dt/dx x."""
    ax = pkg.uniform_axis(n_v, -vmax, vmax)
    w = np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0]), (4, 1)) if wall is None else wall
    return pkg.SolverConfig(dim=1, nx=nx, ny=1, dx=dx, vx=ax, vy=ax.copy(), vz=ax.copy(), dt=dt, coll_solver=0,
                            samples_per_cell=0, enable_convection=1, bc_left=bc, bc_right=bc, acc_coeff=acc,
                            wall=w, conv_order=order)


def advect(pkg, cfg, f, steps):
    s = oracle_solver(cfg)
    s.upload_f(f)
    for n in range(1, steps + 1):
        s.compute_convection(n)
    out = s.download_f()
    s.close()
    return out


def bump(x):
    return 1.0 + np.exp(-((x - 0.35) / 0.08) ** 2)


def l1_error(pkg, nx, order, t_end=0.25):
    """Advect a smooth bump on [0, 1] with speeds -1, 0, 1 (CFL 0.5) for t_end;
    L1 error over the cells away from the ends against the exact shift."""
    dx = 1.0 / nx
    dt = 0.5 * dx
    steps = int(round(t_end / dt))
    cfg = channel(pkg, nx, dx=dx, dt=dt, order=order)
    xc = (np.arange(nx) + 0.5) * dx
    v = pkg.node_velocities(cfg)[:, 0]
    f0 = np.stack([bump(xc)] * cfg.num_points, axis=1)
    out = advect(pkg, cfg, f0, steps)
    exact = bump(xc[:, None] - v[None, :] * steps * dt)
    inner = (xc > 0.05) & (xc < 0.95)
    return np.abs(out - exact)[inner].sum() * dx / cfg.num_points


def test_second_order_converges_at_second_order(pkg):
    e1 = [l1_error(pkg, n, 1) for n in (100, 200, 400)]
    e2 = [l1_error(pkg, n, 2) for n in (100, 200, 400)]
    r1 = np.log2(e1[1] / e1[2])
    r2 = np.log2(e2[1] / e2[2])
    assert 0.8 < r1 < 1.2, e1
    assert r2 > 1.7, e2
    assert e2[2] < 0.1 * e1[2]


def test_second_order_is_tvd_and_keeps_extrema(pkg):
    nx = 40
    cfg = channel(pkg, nx, order=2)
    rng = np.random.default_rng(3)
    prof = np.where(np.arange(nx) < nx // 2, 2.0, 1.0) + 0.3 * rng.random(nx)
    f0 = np.stack([prof] * cfg.num_points, axis=1)
    out = advect(pkg, cfg, f0, 6)
    for l in range(cfg.num_points):
        tv0 = np.abs(np.diff(f0[:, l])).sum()
        tv1 = np.abs(np.diff(out[:, l])).sum()
        assert tv1 <= tv0 * (1 + 1e-12)
        assert out[:, l].min() >= f0[:, l].min() - 1e-12 and out[:, l].max() <= f0[:, l].max() + 1e-12


def test_linear_data_translates_exactly(pkg):
    nx = 16
    cfg = channel(pkg, nx, order=2)
    v = pkg.node_velocities(cfg)[:, 0]
    a, b = 3.0, 0.125
    i = np.arange(nx)[:, None]
    f0 = a + b * i + 0.0 * v[None, :]
    out = advect(pkg, cfg, f0, 1)
    exact = a + b * (i - 0.5 * v[None, :])
    # cells 2..nx-3: both slopes of their stencils are interior
    assert np.allclose(out[2:nx - 2], exact[2:nx - 2], rtol=0, atol=1e-14)


def wall_channel(pkg, acc, n_v=8):
    rng = np.random.default_rng(8)
    wall = np.tile(np.array([1.0, 0.2, -0.1, 0.05, 1.3]), (4, 1))
    cfg = channel(pkg, 6, n_v=n_v, vmax=4.0, dx=0.5, dt=0.05, bc=1, acc=acc, wall=wall)
    f0 = 0.01 + rng.random((6, cfg.num_points))
    return cfg, f0, advect(pkg, cfg, f0, 1)


def halves(cfg, side):
    x = cfg.vx
    neg = np.nonzero(x < 0)[0].max()
    pos = np.nonzero(x > 0)[0].min()
    i = np.arange(cfg.num_points) % cfg.nvx
    inc = (i <= neg) if side == 0 else (i >= pos)
    out = (i >= pos) if side == 0 else (i <= neg)
    return inc, out


@pytest.mark.parametrize("side", [0, 1])
def test_diffuse_wall_emits_the_wall_maxwellian(pkg, side):
    cfg, f0, out = wall_channel(pkg, 1.0)
    ix, nb = (0, 1) if side == 0 else (5, 4)
    inc, outg = halves(cfg, side)
    m = pkg.maxwellian(cfg, 1.0, 0.2, -0.1, 0.05, 1.3, renormalise=False, pi=pkg.config.PI_REF)
    got = out[ix][outg]
    # shape of the wall Maxwellian, density = sum of the incoming half (not flux-weighted)
    assert np.allclose(got / got.sum(), m[outg] / m[outg].sum(), rtol=1e-12, atol=0)
    assert abs(got.sum() - f0[ix][inc].sum()) <= 1e-12 * f0[ix][inc].sum()
    # incoming half: upwind from the neighbour cell
    xv = pkg.node_velocities(cfg)[:, 0]
    want = f0[ix] - np.abs(cfg.dt / cfg.dx * xv) * (f0[ix] - f0[nb])
    assert np.allclose(out[ix][inc], want[inc], rtol=1e-14, atol=0)


@pytest.mark.parametrize("side", [0, 1])
def test_specular_wall_mirrors_mass_and_momentum(pkg, side):
    cfg, f0, out = wall_channel(pkg, 0.0)
    ix = 0 if side == 0 else 5
    inc, outg = halves(cfg, side)
    xv = pkg.node_velocities(cfg)[:, 0]
    mass_in = f0[ix][inc].sum()
    mom_in = (xv * f0[ix])[inc].sum()
    # even n_v: no node at x = 0, the mirror of every incoming node is an outgoing node
    assert abs(out[ix][outg].sum() - mass_in) <= 1e-12 * mass_in
    assert abs((xv * out[ix])[outg].sum() + mom_in) <= 1e-12 * abs(mom_in)
