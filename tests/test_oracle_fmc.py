"""The oracle's full Monte Carlo kernel (full_mc.f90:110-641, iColl = 2),
pinned by what the method fixes rather than by a retyped formula:

* both partners are inverse-CDF picks of |f| (compute_cdf_fmc takes abs of
  phi, not f - f_eq): every pick of the trace equals an independent numpy
  inverse CDF of the same Philox uniforms;
* the depletion weight is sgn f(A) sgn f(B) (dt/Kn) sigma (sum |f|)^2 g /
  (2 num): checked per sample from the picked nodes (hard spheres, sigma = 1),
  which fails on a dropped like-species half, a sign taken from f - f_eq or a
  density in place of sum |f|;
* the sample count follows compute_num_coll_pairs with both density ratios
  sum|f| / n, against its closed form;
* each step keeps mass, momentum and energy to round-off (depletions equal
  replenishments; the remap keeps moments 0-2), also with negative f."""
import numpy as np
import pytest

from tests._oracle import load_oracle, oracle_solver


def fmc_case(pkg, samples=3000, n_v=9, negative=True, **over):
    cfg, f = pkg.homogeneous_relaxation(n_v=n_v, samples=samples)
    cfg = cfg.with_(coll_solver=pkg.capi.DVM_COLL_FMC, **over)
    if negative:
        rng = np.random.default_rng(5)
        idx = rng.choice(cfg.num_points, 30, replace=False)
        f = f.copy()
        f[0, idx] *= -0.3
    return cfg, f


def uniforms(orc, seed, sample, tag, cell, ntime):
    import ctypes as C
    ctr = (C.c_uint32 * 4)(sample, tag, cell, ntime)
    key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
    out = (C.c_uint32 * 4)()
    orc.orc_philox4x32_10(ctr, key, out)
    return np.array(list(out), dtype=np.float64) * 2.0 ** -32


def test_partners_are_inverse_cdf_picks_of_abs_f(pkg):
    orc = load_oracle()
    cfg, f = fmc_case(pkg)
    s = oracle_solver(cfg)
    s.initialize(f)
    irec, drec, n = s.collide_trace(3, 0, 100000)
    assert n == cfg.samples_per_cell and len(irec) == n
    cum = np.concatenate([[0.0], np.cumsum(np.abs(f[0]))])   # sequential, as compute_cdf_fmc
    for r in irec[:400]:
        assert r[0] == 3                                        # every sample is an integral-3 sample
        u = uniforms(orc, cfg.seed, int(r[1]), 3, 0, 3)
        for col, ui in ((2, u[0]), (3, u[1])):
            want = int(np.searchsorted(cum, ui * cum[-1], side="left")) - 1
            assert r[col] == max(want, 0), (r[1], col)


def test_depletion_weight_and_count(pkg):
    cfg, f = fmc_case(pkg)
    s = oracle_solver(cfg)
    s.initialize(f)
    irec, drec, n = s.collide_trace(1, 0, 100000)
    v = pkg.node_velocities(cfg)
    absd = np.abs(f[0]).sum()
    factor = 0.5 * (cfg.dt / cfg.kn) * 1.0 * absd * absd / cfg.samples_per_cell
    live = irec[:, 4] == 0
    lA, lB = irec[live, 2], irec[live, 3]
    g = np.linalg.norm(v[lB] - v[lA], axis=1)
    want = np.sign(f[0][lA]) * np.sign(f[0][lB]) * factor * g
    assert np.allclose(drec[live, 0], want, rtol=1e-12, atol=0)
    assert np.array_equal(drec[live, 0], drec[live, 1])
    assert (np.sign(drec[live, 0]) < 0).any()                   # negative partners present


def test_reference_sample_count_closed_form(pkg):
    cfg, f = fmc_case(pkg, sample_mode=0, coln_rms=0.005)
    s = oracle_solver(cfg)
    s.initialize(f)
    dens = s.calculate_physical_properties()[0, 0]
    absd = dens - 2.0 * (0.5 * (dens - np.abs(f[0]).sum()))
    ratio = absd / dens
    b3 = np.prod([np.diff(a).min() for a in (cfg.vx, cfg.vy, cfg.vz)]) * cfg.mass ** 1.5
    temp = s.calculate_physical_properties()[0, 4]
    want = np.round(0.5 * cfg.dt * temp ** (2.0 / 3.0) * ratio * ratio / (cfg.kn * cfg.coln_rms ** 2 * b3))
    cnt = s.compute_collision_integral(1)
    assert cnt[0] == want and want > 100


@pytest.mark.parametrize("over", [{}, dict(num_repl=2), dict(omega=0.81, alpha=1.4)])
def test_step_conserves_mass_momentum_energy(pkg, over):
    cfg, f = fmc_case(pkg, samples=20000, n_v=11, **over)
    v = pkg.node_velocities(cfg)
    s = oracle_solver(cfg)
    s.initialize(f)
    for n in range(1, 4):
        g = s.download_f()[0]
        a = np.array([g.sum(), g @ v[:, 0], g @ v[:, 1], g @ v[:, 2], g @ (v ** 2).sum(1)])
        assert s.compute_collision_integral(n)[0] == 20000
        g = s.download_f()[0]
        b = np.array([g.sum(), g @ v[:, 0], g @ v[:, 1], g @ v[:, 2], g @ (v ** 2).sum(1)])
        s.calculate_physical_properties()
        assert np.all(np.abs(b - a) <= 1e-13 * (np.abs(a) + 1.0)), (n, b - a)
