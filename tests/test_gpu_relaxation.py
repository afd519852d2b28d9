"""Space-homogeneous relaxation, the reference's CPU-runnable configuration:
21^3 nodes on +-5, two Maxwellians of equal density at T and 2T moving at
+-1 along x, hard spheres, 100 000 samples per step, dt = 0.05, 200 steps,
FP64; density, velocity, temperature and the L1 distance from equilibrium
(dvm_l1) every step. dvm_l1 equals the oracle's orc_l1 on the same f."""
import numpy as np
import pytest

from tests._oracle import oracle_solver
from tests.test_oracle_l1 import _maxwellian

pytestmark = pytest.mark.gpu


def test_l1_matches_oracle(pkg):
    cfg, _ = pkg.lid_driven_cavity(nx=4, ny=1, n_v=17, vmax=4.0, samples=10)
    rng = np.random.default_rng(3)
    mix = 0.5 * (_maxwellian(pkg, cfg, 1.0, (0.5, 0, 0), 0.7) + _maxwellian(pkg, cfg, 1.0, (-0.5, 0.2, 0), 1.6))
    f = np.stack([mix, _maxwellian(pkg, cfg, 2.0, (0.1, 0.2, 0.3), 1.2), rng.random(cfg.num_points) * 1e-3,
                  np.zeros(cfg.num_points)])
    print("[l1] 4 cells, 17^3", flush=True)
    with pkg.Solver(cfg) as g:
        g.initialize(f)
        lg = g.l1_to_equilibrium()
    o = oracle_solver(cfg)
    o.initialize(f)
    lo = o.l1_to_equilibrium()
    o.close()
    assert lg[3] == -1.0 and lo[3] == -1.0
    # L1 is a sum of |f - M| with f close to M: the moments' round-off (the two
    # sides reduce in different orders) moves it by ~1e-14 in units of n, so
    # the bar is absolute in those units (L1 lies in [0, 2])
    assert np.all(np.abs(lg[:3] - lo[:3]) <= 1e-12), (lg, lo)


def test_relaxation_200_steps(pkg):
    cfg, f = pkg.homogeneous_relaxation()  # 21^3 on +-5, 100 000 samples, dt 0.05
    print(f"[relaxation] {cfg.name}: 200 steps", flush=True)
    props, l1 = [], []
    with pkg.Solver(cfg) as s:
        props.append(s.initialize(f)[0])
        l1.append(s.l1_to_equilibrium()[0])
        for n in range(1, 201):
            s.step(n)
            props.append(s.calculate_physical_properties()[0])
            l1.append(s.l1_to_equilibrium()[0])
    props, l1 = np.array(props), np.array(l1)
    print(f"  L1 {l1[0]:.3e} -> {l1[10]:.3e} (10) -> {l1[50]:.3e} (50) -> {l1[-1]:.3e} (200)", flush=True)
    # collisions keep density, bulk velocity and temperature (round-off)
    assert np.max(np.abs(props[:, 0] - props[0, 0])) <= 1e-11 * props[0, 0]
    assert np.max(np.abs(props[:, 1:4] - props[0, 1:4])) <= 1e-11
    assert np.max(np.abs(props[:, 4] - props[0, 4])) <= 1e-11 * props[0, 4]
    # and drive f to its equilibrium
    assert l1[0] > 0.1
    assert l1[10] < l1[0] and l1[-1] < 0.1 * l1[0]
    assert np.max(l1[150:]) < 0.15 * l1[0]
