"""Closed forms that pin the parts of the oracle no compiled reference module
covers (the collision, moment and shock modules need GSL and are not built
here). Each expected value comes from the mathematics or from a
reference-held fixture, not from the oracle's own expressions:

* the mean depletion of a sampled VR step at every node equals the loss term
  of the discrete collision operator, dt/Kn sigma [f G f - f_eq G f_eq] with
  G = |v_l - v_m|^(2 - 2 omega) (variance_reduction.f90:596-604, 679, 831, 976
  and vr_depletion_routine :1127-1237); a dropped 1/2, a missing 1/Kn, a
  wrong sigma, exponent or deviation sign moves it by tens of percent;
* the same for the full Monte Carlo kernel, dt/Kn sigma f G f
  (full_mc.f90:240-321);
* the equilibrium CDF's increments are compute_maxwellian (golden-pinned) x
  beta^3 with the coefficient of variance_reduction.f90:311 and, with the
  reference's pi (constants.f90:30), carry (pi/pi_ref)^(3/2) of the density;
* the cache rule of variance_reduction.f90:316-322 on crafted state changes;
* Bird's VSS deflection law cos chi = 2 u^(1/alpha) - 1 about the folded
  relative velocity (collision_utilities.f90:259-269, 301-328);
* the Rankine-Hugoniot ratios of software/shock_calculator/
  normal_shock_parameters.txt;
* the moments and output moments (physical_props.f90:285-485, 1871-2065)
  against vectorised numpy sums, and the grid integers against the reference's
  grid fields (velocity_grid.f90:542-597, 653-689).
"""
import ctypes as C
import math

import numpy as np
import pytest

from tests._oracle import golden_cases, load_oracle, oracle_solver

PI_REF = 3.1459265358979323  # constants.f90:30

P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
I = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))


def _lib():
    lib = load_oracle()
    lib.orc_maxwellian.restype = C.c_double
    lib.orc_maxwellian.argtypes = [C.c_double] * 9
    lib.orc_cross_section.restype = C.c_double
    lib.orc_cross_section.argtypes = [C.c_int32] + [C.c_double] * 5
    lib.orc_set_f_keep_cache.restype = C.c_int
    lib.orc_set_f_keep_cache.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.orc_grid_info.restype = C.c_int
    lib.orc_grid_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    return lib


def _beta3(pkg, cfg):
    bx, by, bz = (pkg.config.beta_axis(a) for a in (cfg.vx, cfg.vy, cfg.vz))
    return (bz[:, None, None] * by[None, :, None] * bx[None, None, :]).reshape(-1)


def _np_moments(f, v, mass):
    """n, u (3), T = 2 m sum f |c|^2 / (3 n) (physical_props.f90:879-1087)."""
    n = f.sum()
    u = (f @ v) / n
    c = v - u
    t = 2.0 * mass * (f * (c * c).sum(1)).sum() / (3.0 * n)
    return n, u, t


def _ref_equilibrium(lib, pkg, cfg, f, props=None):
    """The cell's discrete equilibrium: compute_maxwellian (dist_func.f90:317-344,
    golden-pinned) at the cell's moments times beta^3, with the coefficient
    n sqrt(m^3 / (pi_ref^3 T^3)) of variance_reduction.f90:311. `props`: the
    cell's moment record (pinned against numpy sums below), else numpy sums."""
    v = pkg.node_velocities(cfg)
    if props is None:
        n, u, t = _np_moments(f, v, cfg.mass)
    else:
        n, u, t = props[0], props[1:4], props[5]
    coeff = n * math.sqrt(cfg.mass ** 3 / (PI_REF ** 3 * t ** 3))
    mw = np.array([lib.orc_maxwellian(x, y, z, coeff, cfg.mass, u[0], u[1], u[2], t) for x, y, z in v])
    return mw * _beta3(pkg, cfg)


def _kernel_matrix(v, omega):
    g = np.sqrt(((v[:, None, :] - v[None, :, :]) ** 2).sum(-1))
    return g ** (2.0 - 2.0 * omega)  # zero on the diagonal: a self-collision is skipped


def _mean_depletion(s, N, reps, max_records):
    """Per-node sum of both partners' depletions over one collision phase,
    averaged over `reps` independent traces (time steps 1..reps key Philox
    differently; a trace leaves the state untouched)."""
    per = np.zeros((reps, N))
    for r in range(reps):
        irec, drec, n = s.collide_trace(r + 1, 0, max_records)
        assert n <= max_records
        np.add.at(per[r], irec[:, 2], drec[:, 0])
        np.add.at(per[r], irec[:, 3], drec[:, 1])
    return per.mean(0), per.std(0, ddof=1) / math.sqrt(reps)


def _check_against_loss(est, se, exact, v):
    # node-wise: relative L2 distance, and every node within 6 standard errors
    rel = np.linalg.norm(est - exact) / np.linalg.norm(exact)
    assert rel < 0.03, rel
    sig = se > 0
    z = np.abs(est - exact)[sig] / se[sig]
    assert np.max(z) < 6.0, np.max(z)
    # moments of the depletion: mass, momentum and energy removed by the loss term
    for psi in (np.ones(len(v)), v[:, 0], v[:, 0] ** 2, (v * v).sum(1)):
        a, b = psi @ est, psi @ exact
        assert abs(a - b) <= 0.02 * np.abs(psi * exact).sum(), (a, b)


@pytest.mark.parametrize("omega,kn,diam,cs", [(0.5, 0.37, 1.3, 1), (0.81, 0.8, 1.0, 2)])
def test_vr_mean_depletion_is_the_loss_term(pkg, omega, kn, diam, cs):
    lib = _lib()
    cfg, f = pkg.homogeneous_relaxation(n_v=7, vmax=3.5, samples=30000, dt=0.05)
    cfg = cfg.with_(omega=omega, kn=kn, diam=diam, csection_coeff=cs, viscosity=1.7, t_viscosity=1.2)
    s = oracle_solver(cfg)
    s.initialize(f)
    f0 = f[0]
    v = pkg.node_velocities(cfg)
    est, se = _mean_depletion(s, cfg.num_points, 40, 40000)
    m = cfg.mass
    sigma = lib.orc_cross_section(cs, m * m / (m + m), omega, diam, cfg.viscosity, cfg.t_viscosity)
    G = _kernel_matrix(v, omega)
    eq = _ref_equilibrium(lib, pkg, cfg, f0)
    exact = cfg.dt / cfg.kn * sigma * (f0 * (G @ f0) - eq * (G @ eq))
    _check_against_loss(est, se, exact, v)
    # a planted slip is visible: half the weight, or |f - f_eq| in place of the signed deviation
    assert np.linalg.norm(est - 0.5 * exact) / np.linalg.norm(exact) > 0.3
    dev = np.abs(f0 - eq)
    wrong = cfg.dt / cfg.kn * sigma * (dev * (G @ eq) + eq * (G @ dev) + dev * (G @ dev)) * 1.0
    assert np.linalg.norm(est - wrong) / np.linalg.norm(exact) > 0.3


def test_fmc_mean_depletion_is_the_loss_term(pkg):
    lib = _lib()
    cfg, f = pkg.homogeneous_relaxation(n_v=7, vmax=3.5, samples=30000, dt=0.05)
    cfg = cfg.with_(coll_solver=pkg.capi.DVM_COLL_FMC, kn=0.45, diam=1.2, omega=0.7)
    f0 = f[0].copy()
    # a few negative values (FMC signs are dsgn(f), full_mc.f90:337-338)
    rng = np.random.default_rng(5)
    f0[rng.choice(len(f0), 20, replace=False)] *= -0.5
    s = oracle_solver(cfg)
    s.initialize(f0[None, :])
    v = pkg.node_velocities(cfg)
    est, se = _mean_depletion(s, cfg.num_points, 40, 40000)
    m = cfg.mass
    sigma = lib.orc_cross_section(1, m * m / (m + m), cfg.omega, cfg.diam, 1.0, 1.0)
    G = _kernel_matrix(v, cfg.omega)
    exact = cfg.dt / cfg.kn * sigma * f0 * (G @ f0)
    _check_against_loss(est, se, exact, v)
    assert np.linalg.norm(est - 0.5 * exact) / np.linalg.norm(exact) > 0.3


def _cell_cdf(s, N):
    lib = _lib()
    eqc = np.zeros(N + 1)
    neqc = np.zeros(N + 1)
    tabs = [np.zeros(1000, np.int32) for _ in range(4)]
    cnt = np.zeros(3, np.int32)
    assert lib.orc_cell_cdf(s.h, 0, P(eqc), P(neqc), *(I(t) for t in tabs), I(cnt)) == 0
    return eqc, neqc


def test_equilibrium_cdf_is_the_reference_maxwellian(pkg):
    lib = _lib()
    ax = pkg.config.uniform_axis(11, -4.0, 4.0)
    cfg = pkg.SolverConfig(vx=ax, vy=pkg.config.uniform_axis(9, -3.5, 4.5), vz=ax * 0.9, samples_per_cell=100)
    f = pkg.config.maxwellian(cfg, 1.3, 0.4, -0.2, 0.1, 1.4)
    s = oracle_solver(cfg)
    props = s.initialize(f[None, :])[0]
    eqc, neqc = _cell_cdf(s, cfg.num_points)
    exp = _ref_equilibrium(lib, pkg, cfg, f, props)
    inc = np.diff(eqc)
    # the increments are a difference of prefix sums (Q5): equal up to the
    # rounding of the running sum
    assert np.all(np.abs(inc - exp) <= 4 * np.finfo(float).eps * eqc[1:]), np.max(np.abs(inc - exp))
    # the deviation CDF accumulates |f - f_eq| with f_eq the same increments
    assert np.allclose(np.diff(neqc), np.abs(f - inc), rtol=0, atol=8 * np.finfo(float).eps * eqc[-1])


def test_equilibrium_carries_the_reference_pi_deficit(pkg):
    """Q1 and Q5: with pi_ref = 3.14592... in the coefficient and no
    renormalisation, a resolved discrete equilibrium holds (pi / pi_ref)^(3/2)
    of the cell's density (the Gaussian integral of exp(-m c^2 / T) is
    (pi T / m)^(3/2))."""
    ax = pkg.config.uniform_axis(41, -6.0, 6.0)
    cfg = pkg.SolverConfig(vx=ax, vy=ax.copy(), vz=ax.copy(), samples_per_cell=100)
    f = pkg.config.maxwellian(cfg, 0.8, 0.3, 0.0, -0.2, 1.0)
    s = oracle_solver(cfg)
    s.initialize(f[None, :])
    eqc, _ = _cell_cdf(s, cfg.num_points)
    ratio = eqc[-1] / f.sum()
    assert abs(ratio - (math.pi / PI_REF) ** 1.5) < 1e-9, ratio
    assert abs(ratio - 1.0) > 2e-3  # not renormalised to the density


def test_equilibrium_cache_rule(pkg):
    """variance_reduction.f90:316-322: the equilibrium is rebuilt only if n, u,
    v or T moved by more than tol times their value; w is never compared; a
    negative u or v makes the test true whatever the change."""
    lib = _lib()
    ax = pkg.config.uniform_axis(9, -4.0, 4.0)
    tol = 0.01
    cfg = pkg.SolverConfig(vx=ax, vy=ax.copy(), vz=ax.copy(), samples_per_cell=0, cdf_tolerance=tol)
    N = cfg.num_points
    mw = lambda n, u, v, w, t: pkg.config.maxwellian(cfg, n, u, v, w, t)

    def run(base, moved):
        s = oracle_solver(cfg)
        s.initialize(mw(*base)[None, :])
        s.compute_collision_integral(1)  # commits the cache at `base`
        eq0, _ = _cell_cdf(s, N)
        f1 = mw(*moved)
        assert lib.orc_set_f_keep_cache(s.h, P(f1)) == 0
        props = s.calculate_physical_properties()[0]
        eq1, _ = _cell_cdf(s, N)
        rebuilt = _ref_equilibrium(lib, pkg, cfg, f1, props)
        s.close()
        if np.array_equal(eq1, eq0):
            return "kept"
        assert np.allclose(np.diff(eq1), rebuilt, rtol=1e-12, atol=1e-15)
        return "rebuilt"

    b = (1.0, 0.3, 0.2, 0.1, 1.2)
    assert run(b, (1.0 * (1 + 0.5 * tol), 0.3, 0.2, 0.1, 1.2)) == "kept"
    assert run(b, (1.0 * (1 + 2.0 * tol), 0.3, 0.2, 0.1, 1.2)) == "rebuilt"
    assert run(b, (1.0, 0.3 * (1 + 2 * tol), 0.2, 0.1, 1.2)) == "rebuilt"
    assert run(b, (1.0, 0.3, 0.2 * (1 + 2 * tol), 0.1, 1.2)) == "rebuilt"
    assert run(b, (1.0, 0.3, 0.2, 0.1, 1.2 * (1 + 2 * tol))) == "rebuilt"
    assert run(b, (1.0, 0.3, 0.2, 0.1, 1.2 * (1 + 0.5 * tol))) == "kept"
    assert run(b, (1.0, 0.3, 0.2, 0.1 + 0.5, 1.2)) == "kept"  # w is not compared
    # negative u (or v): |du| > tol * u holds for any change
    assert run((1.0, -0.3, 0.2, 0.1, 1.2), (1.0 * (1 + 0.1 * tol), -0.3, 0.2, 0.1, 1.2)) == "rebuilt"
    assert run((1.0, 0.3, -0.2, 0.1, 1.2), (1.0 * (1 + 0.1 * tol), 0.3, -0.2, 0.1, 1.2)) == "rebuilt"


def _stencil_velocity(cfg, st, m):
    """Velocity carried by a remap stencil with unit weight: sum_q m_q v_q
    (the remap keeps momentum, remapping.f90:92-394)."""
    ax = (cfg.vx, cfg.vy, cfg.vz)
    oi, oj, ok, a, b, c, flag = st.T
    pts = [(oi, oj, ok), (oi + a, oj, ok), (oi, oj + b, ok), (oi, oj, ok + c),
           (oi - a, oj, ok), (oi, oj - b, ok), (oi, oj, ok - c)]
    out = np.zeros((len(st), 3))
    for q, (i, j, k) in enumerate(pts):
        use = np.ones(len(st), bool) if q < 4 else (flag & (1 << (q - 4))) != 0
        i, j, k = (np.clip(x, 0, len(a_) - 1) for x, a_ in zip((i, j, k), ax))
        vel = np.stack([cfg.vx[i], cfg.vy[j], cfg.vz[k]], 1)
        out += np.where(use[:, None], m[:, q:q + 1] * vel, 0.0)
    return out


@pytest.mark.parametrize("alpha", [1.4, 1.0])
def test_vss_deflection_law(pkg, alpha):
    """cos chi between the post-collision relative velocity and the folded
    pre-collision one (|g_x|, |g_y|, |g_z|) follows Bird's VSS law,
    P(cos chi <= x) = ((1 + x) / 2)^alpha (alpha = 1: isotropic)."""
    cfg, f = pkg.homogeneous_relaxation(n_v=9, vmax=4.0, samples=60000)
    cfg = cfg.with_(omega=0.81, alpha=alpha)
    s = oracle_solver(cfg)
    s.initialize(f)
    irec, drec, n = s.collide_trace(1, 0, 60000)
    keep = irec[:, 4] == 0
    irec, drec = irec[keep], drec[keep]
    v = pkg.node_velocities(cfg)
    v1, v2 = v[irec[:, 2]], v[irec[:, 3]]
    p1 = _stencil_velocity(cfg, irec[:, 5:12], drec[:, 2:9])
    p2 = _stencil_velocity(cfg, irec[:, 12:19], drec[:, 9:16])
    g = v2 - v1
    gsq = (g * g).sum(1)
    # both particles keep the centre of mass and the relative speed
    assert np.allclose(p1 + p2, v1 + v2, atol=1e-12)
    assert np.allclose(((p1 - p2) ** 2).sum(1), gsq, rtol=1e-12)
    cos_chi = ((p1 - p2) * np.abs(g)).sum(1) / gsq
    x = np.sort(np.clip(cos_chi, -1, 1))
    cdf = ((1 + x) / 2) ** alpha
    emp = np.arange(1, len(x) + 1) / len(x)
    ks = np.max(np.abs(emp - cdf))
    assert ks < 1.95 / math.sqrt(len(x)), ks
    # the exponent 1/alpha (not alpha) is what the reference draws
    wrong = ((1 + x) / 2) ** (1.0 / alpha)
    if alpha != 1.0:
        assert np.max(np.abs(emp - wrong)) > 10 / math.sqrt(len(x))


def test_rankine_hugoniot_matches_reference_fixture(pkg):
    """software/shock_calculator/normal_shock_parameters.txt: Mach 3.9 (shock
    frame) gives P2/P1 = 18.9009090909, n2/n1 = 3.2424418605,
    T2/T1 = 5.8292206628, and with a molecular mass of 0.12990501E-25 kg over
    mass_ref 0.21800001E-24 kg at T_up = 1, u_up = 14.7371451425 and
    u_down = 4.5450761422. The file prints the specific heat ratio rounded
    (1.70175): the calculator assigns cp_mix from the last species alone
    (normal_shock.f90:89, He at mole fraction 0.97 with 3 d.o.f.), so
    gamma = cp / (cp - n k) with cp = 0.97 (3/2 + 1) n k (:97)."""
    cp = 0.97 * (3.0 * 0.5 + 1.0)
    gamma = cp / (cp - 1.0)
    assert abs(gamma - 1.70175) < 5e-6
    dens, vel, temp, press = pkg.config.rankine_hugoniot(3.9, gamma)
    assert abs(press - 18.9009090909) < 1e-9
    assert abs(dens - 3.2424418605) < 1e-9
    assert abs(temp - 5.8292206628) < 1e-9
    m = 0.12990501e-25 / 0.21800001e-24
    u_up = 3.9 * pkg.config.sound_speed(gamma, 1.0, m)
    assert abs(u_up - 14.7371451425) < 2e-7 * 14.7
    assert abs(u_up * vel - 4.5450761422) < 2e-7 * 4.5


def test_moments_match_numpy_sums(pkg):
    lib = _lib()
    case = golden_cases("remap_grid.npz")["nonuni"]
    cfg = pkg.SolverConfig(vx=case["ax"], vy=case["ay"], vz=case["az"], nx=3, ny=1, samples_per_cell=0,
                           mass=1.7)
    rng = np.random.default_rng(3)
    f = rng.uniform(0.0, 1.0, size=(3, cfg.num_points))
    f[2, :5] = -0.3  # negative values enter the sums as they are, |f| only in the entropy
    s = oracle_solver(cfg)
    s.upload_f(f)
    p = s.calculate_physical_properties()
    ext = s.output_moments()
    v = pkg.node_velocities(cfg)
    m = cfg.mass
    for c in range(3):
        n, u, t = _np_moments(f[c], v, m)
        spd = (f[c] * np.sqrt((v * v).sum(1))).sum() / n
        ref = np.array([n, u[0], u[1], u[2], t, t, 0.5 * n * t, spd])
        assert np.allclose(p[c], ref, rtol=1e-12, atol=1e-14), (c, p[c] - ref)
        cc = v - u
        csq = (cc * cc).sum(1)
        q = 0.5 * m * (cc * (csq * f[c])[:, None]).sum(0)
        tau = -m * (cc[:, :, None] * cc[:, None, :] * f[c][:, None, None]).sum(0)
        tau += np.eye(3) * 0.5 * n * t
        nz = f[c] != 0
        ent = -(np.abs(f[c][nz]) * np.log(np.abs(f[c][nz]))).sum()
        ref_ext = np.array([q[0], q[1], q[2], tau[0, 0], tau[1, 1], tau[2, 2], tau[0, 1], tau[0, 2], tau[1, 2], ent])
        scale = np.abs(ref_ext).max()
        assert np.allclose(ext[c], ref_ext, rtol=1e-11, atol=1e-12 * scale), (c, ext[c] - ref_ext)


@pytest.mark.parametrize("name", ["u25", "u18", "u357", "nonuni", "shift"])
def test_grid_integers_match_reference(pkg, name):
    lib = _lib()
    case = golden_cases("remap_grid.npz")[name]
    cfg = pkg.SolverConfig(vx=case["ax"], vy=case["ay"], vz=case["az"], samples_per_cell=0)
    s = oracle_solver(cfg)
    ints = np.zeros(11, np.int32)
    b3 = C.c_double(0.0)
    assert lib.orc_grid_info(s.h, I(ints), C.byref(b3)) == 0
    assert np.array_equal(ints, case["grid_ints"].astype(np.int32)), (ints, case["grid_ints"])
    assert b3.value == float(case["beta3_min"])
