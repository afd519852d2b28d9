"""The oracle's wall heat fluxes (boundary_conditions.f90:872-923 and the
right / bottom / top analogues), pinned by what the physics fixes. The
reference's BoundaryConditions module needs GSL and cannot be built here, so
these invariants are the pins:

* a gas at rest in equilibrium with diffuse walls at its own temperature
  exchanges no heat: every wall cell's flux vanishes to round-off (fails on a
  wrong half, a missing flux_in / flux_out scaling or a ghost from the wrong
  side);
* a gas hotter than the walls heats every wall: the flux points out of the gas
  (negative on the left and bottom walls, positive on the right and top), and
  on a symmetric grid opposite walls see equal and opposite fluxes (fails on a
  swapped half or a wrong wall-normal component); colder gas reverses the
  signs;
* the flux is recorded only on the steps the interval selects."""
import numpy as np
import pytest

from tests._oracle import oracle_solver


def box(pkg, t_gas, t_wall=1.0, n_v=11, acc=1.0):
    cfg, _ = pkg.lid_driven_cavity(nx=4, ny=3, n_v=n_v, vmax=4.0, samples=0, lid=0.0, t_wall=t_wall)
    cfg = cfg.with_(coll_solver=0, acc_coeff=acc)
    f = np.tile(pkg.maxwellian(cfg, 1.0, 0.0, 0.0, 0.0, t_gas), (cfg.nx * cfg.ny, 1))
    return cfg, f


def heat_after_one_step(pkg, cfg, f, every=1):
    s = oracle_solver(cfg)
    s.initialize(f)
    s.set_wall_heat_flux_freq(every)
    s.step(1)
    q, nt = s.wall_heat_flux()
    scale = 0.5 * cfg.mass * np.abs(f[0] @ (pkg.node_velocities(cfg)[:, 0] * (pkg.node_velocities(cfg) ** 2).sum(1)))
    s.close()
    return q, nt, scale


def test_no_heat_flux_at_equilibrium(pkg):
    cfg, f = box(pkg, 1.0)
    q, nt, _ = heat_after_one_step(pkg, cfg, f)
    v = pkg.node_velocities(cfg)
    half = 0.5 * cfg.mass * np.abs(f[0][v[:, 0] < 0] @ (v[v[:, 0] < 0, 0] * (v[v[:, 0] < 0] ** 2).sum(1)))
    assert nt == 1 and half > 0.1
    assert np.all(np.abs(q[:2, :cfg.ny]) <= 1e-12 * half), q
    assert np.all(np.abs(q[2:, :cfg.nx]) <= 1e-12 * half), q


@pytest.mark.parametrize("t_gas, sign", [(1.6, 1.0), (0.6, -1.0)])
def test_hot_gas_heats_the_walls(pkg, t_gas, sign):
    cfg, f = box(pkg, t_gas)
    q, nt, _ = heat_after_one_step(pkg, cfg, f)
    left, right, bottom, top = q[0, :cfg.ny], q[1, :cfg.ny], q[2, :cfg.nx], q[3, :cfg.nx]
    assert np.all(sign * left < 0) and np.all(sign * right > 0)
    assert np.all(sign * bottom < 0) and np.all(sign * top > 0)
    assert np.allclose(left, -right, rtol=1e-12, atol=0)
    assert np.allclose(bottom, -top, rtol=1e-12, atol=0)
    # uniform gas: every cell along a wall sees the same flux
    assert np.allclose(left, left[0], rtol=1e-12, atol=0) and np.allclose(bottom, bottom[0], rtol=1e-12, atol=0)


def test_recorded_on_the_interval_only(pkg):
    cfg, f = box(pkg, 1.4)
    s = oracle_solver(cfg)
    s.initialize(f)
    assert s.wall_heat_flux()[1] == -1
    s.set_wall_heat_flux_freq(2)
    s.step(1)
    assert s.wall_heat_flux()[1] == -1
    s.step(2)
    q, nt = s.wall_heat_flux()
    assert nt == 2 and np.all(q[0, :cfg.ny] < 0)
    s.step(3)
    assert s.wall_heat_flux()[1] == 2
    s.close()
