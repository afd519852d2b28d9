"""The oracle's output moments (heat flux, viscous stress, entropy; used by the
Tecplot writer) pinned by closed forms: the stress is traceless (T is defined
by the same second moment), an isotropic Maxwellian at rest on a symmetric
grid has no heat flux and no stress, the Maxwellian with the same moments has
the largest entropy (Gibbs inequality, uniform grid), and collisions raise
the entropy (H-theorem)."""
import numpy as np

from tests._oracle import oracle_solver


def ext_of(pkg, cfg, f):
    s = oracle_solver(cfg)
    s.initialize(f)
    e = s.output_moments()
    p = s.calculate_physical_properties()
    s.close()
    return e, p


def test_stress_is_traceless(pkg):
    cfg, f = pkg.homogeneous_relaxation(n_v=15, samples=100)
    e, p = ext_of(pkg, cfg, f)
    assert abs(e[0, 3] + e[0, 4] + e[0, 5]) <= 1e-12 * p[0, pkg.capi.P_PRESSURE]
    assert e[0, 3] < 0 < e[0, 4]  # counter-streaming along x: excess x-momentum flux


def test_isotropic_maxwellian_has_no_flux_and_no_stress(pkg):
    cfg, _ = pkg.homogeneous_relaxation(n_v=15, samples=100)
    f = pkg.maxwellian(cfg, 1.3, 0.0, 0.0, 0.0, 1.7).reshape(1, -1)
    e, p = ext_of(pkg, cfg, f)
    assert np.all(np.abs(e[0, :9]) <= 1e-13 * p[0, pkg.capi.P_PRESSURE])


def test_maxwellian_maximises_entropy(pkg):
    cfg, f = pkg.homogeneous_relaxation(n_v=15, samples=100)
    e, p = ext_of(pkg, cfg, f)
    m = pkg.maxwellian(cfg, *p[0, :4], p[0, pkg.capi.P_TRTEMP]).reshape(1, -1)
    em, _ = ext_of(pkg, cfg, m)
    assert em[0, 9] > e[0, 9]


def test_collisions_raise_entropy(pkg):
    cfg, f = pkg.homogeneous_relaxation(n_v=15, samples=20000)
    s = oracle_solver(cfg)
    s.initialize(f)
    ent = [s.output_moments()[0, 9]]
    for n in range(1, 9):
        s.step(n)
        ent.append(s.output_moments()[0, 9])
    s.close()
    assert ent[-1] > ent[0] and np.all(np.diff(ent[::2]) > 0)
