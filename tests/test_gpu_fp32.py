"""FP32-collision path: FP32 arithmetic for relative speeds, cross sections,
post-collision velocities and weights, FP64 remap weights, FP64 moments and
an FP64 conservation correction. Checked against the FP64 path on the same
seed (same partner picks): moments within 1e-5 relative, and mass, momentum
and energy conserved to FP64 round-off every step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP32 = 1


def cons(f, v):
    return np.stack([f.sum(1), f @ v[:, 0], f @ v[:, 1], f @ v[:, 2], f @ (v ** 2).sum(1)], axis=1)


@pytest.mark.parametrize("case", ["homogeneous", "vss_nrepl2", "global_atomics"])
def test_fp32_collision_conserves_to_roundoff(pkg, case):
    cfg, f = pkg.homogeneous_relaxation(n_v=29 if case == "global_atomics" else 17, samples=40000)
    if case == "vss_nrepl2":
        cfg = cfg.with_(omega=0.81, alpha=1.4, num_repl=2)
    cfg = cfg.with_(precision=FP32)
    print(f"[fp32 conservation] {case}", flush=True)
    v = pkg.node_velocities(cfg)
    with pkg.Solver(cfg) as s:
        s.initialize(f)
        for n in range(1, 6):
            a = cons(s.download_f(), v)
            s.compute_collision_integral(n)
            b = cons(s.download_f(), v)
            s.calculate_physical_properties()
            scale = np.abs(a) + np.abs(cons(np.abs(s.download_f()), np.abs(v)))
            assert np.all(np.abs(b - a) <= 1e-13 * scale), (n, (b - a) / scale)


@pytest.mark.parametrize("case", ["homogeneous21", "homogeneous29", "cavity10", "vhs21", "nrepl21"])
def test_fp32_collision_matches_fp64_distribution(pkg, case):
    """One collision of the FP32 path against the FP64 path from the same f and
    seed, node by node: the L1 difference of the post-collision distributions
    stays at FP32 rounding of the collision's change (the picks coincide). 21^3: the whole grid in LDS;
    29^3: an LDS window around the bulk velocity with the rest of the grid
    taking global atomics, which the conservation test alone would not see
    (the correction restores the conserved moments of any scatter); a cavity
    on 10^3 nodes: an even node count, where every other cell's sign bytes
    start 2 bytes off a 4-byte boundary; VHS (g^(2 - 2 omega) with omega 0.81)
    and two replenish locations per sample, each with its own Philox block."""
    if case == "cavity10":
        cfg, f = pkg.lid_driven_cavity(nx=4, ny=4, n_v=10, vmax=4.0, samples=20000)
    else:
        cfg, f = pkg.homogeneous_relaxation(n_v=int(case[-2:]), samples=40000)
        if case.startswith("vhs"):
            cfg = cfg.with_(omega=0.81)
        elif case.startswith("nrepl"):
            cfg = cfg.with_(num_repl=2)
    print(f"[fp32 vs fp64 distribution] {cfg.name}", flush=True)
    out = {}
    for prec in (0, FP32):
        with pkg.Solver(cfg.with_(precision=prec)) as s:
            s.initialize(f)
            s.compute_collision_integral(1)
            out[prec] = s.download_f()
    d = np.abs(out[FP32] - out[0]).sum() / np.abs(out[0]).sum()
    change = np.abs(out[0] - f).sum() / np.abs(out[0]).sum()
    print(f"  L1 difference {d:.2e} of a change of {change:.2e}", flush=True)
    # FP32 rounding of the change (about 1e-6 of it); a misplaced scatter
    # would be of the order of the change itself
    assert change > 1e-6, change
    assert d <= 1e-5 * change, (d, change)


@pytest.mark.parametrize("n_v", [9, 10])
def test_fp32_persistent_cells_match_fp64(pkg, n_v):
    """A cavity with enough cells for one workgroup per cell (576 cells, so
    the FP32 collision runs whole cells on persistent workgroups with the
    moments pass's conservation factor): one collision against the FP64 one
    from the same f and seed, node by node, and mass, momentum and energy of
    every cell conserved to round-off; 10^3 nodes put every other cell's sign
    words 2 bytes off a 4-byte boundary."""
    cfg, f = pkg.lid_driven_cavity(nx=24, ny=24, n_v=n_v, vmax=4.0, samples=2000)
    print(f"[fp32 persistent cells] {cfg.name}", flush=True)
    v = pkg.node_velocities(cfg)
    out = {}
    for prec in (0, FP32):
        with pkg.Solver(cfg.with_(precision=prec)) as s:
            s.initialize(f)
            s.compute_collision_integral(1)
            out[prec] = s.download_f()
    d = np.abs(out[FP32] - out[0]).sum() / np.abs(out[0]).sum()
    change = np.abs(out[0] - f).sum() / np.abs(out[0]).sum()
    print(f"  L1 difference {d:.2e} of a change of {change:.2e}", flush=True)
    assert change > 1e-6, change
    # near equilibrium with 2000 samples per cell the net change is a small
    # difference of depletions and replenishments that are each rounded in
    # FP32 (measured 4e-6 and 1e-5 of it); a misplaced scatter or sign would
    # be of the order of the change itself
    assert d <= 1e-4 * change, (d, change)
    a, b = cons(f, v), cons(out[FP32], v)
    scale = np.abs(a) + cons(np.abs(out[FP32]), np.abs(v))
    assert np.all(np.abs(b - a) <= 1e-13 * scale), np.max(np.abs(b - a) / scale)


def _moment_err(pa, pb):
    """Max relative error of density, velocity (in units of sqrt(T)) and T."""
    th = np.sqrt(np.abs(pa[:, 4]))
    return max(np.max(np.abs(pb[:, 0] - pa[:, 0]) / np.abs(pa[:, 0])),
               np.max(np.abs(pb[:, 1:4] - pa[:, 1:4]).max(axis=1) / th),
               np.max(np.abs(pb[:, 4] - pa[:, 4]) / np.abs(pa[:, 4])))


def _builders(pkg, case):
    if case == "homogeneous":
        return pkg.homogeneous_relaxation(n_v=21, samples=100000)
    return pkg.lid_driven_cavity(nx=12, ny=12, n_v=15, vmax=4.0, samples=5000)


@pytest.mark.parametrize("case", ["homogeneous", "cavity"])
def test_fp32_step_moments_within_1e5_of_fp64(pkg, case):
    """Per-step precision of the FP32-collision path: from the same state and
    seed, one FP32 step and one FP64 step agree in every moment field to far
    below 1e-5 (the partner picks are FP64 in both, so the samples coincide;
    a pick can differ only when a uniform lands within rounding of a CDF
    boundary). The FP64 trajectory is advanced and re-shared every step."""
    cfg, f = _builders(pkg, case)
    print(f"[fp32 vs fp64, per step] {cfg.name}", flush=True)
    a = pkg.Solver(cfg)
    b = pkg.Solver(cfg.with_(precision=FP32))
    a.initialize(f)
    worst = 0.0
    for n in range(1, 9):
        state = a.download_f()
        b.initialize(state)
        a.step(n)
        b.step(n)
        err = _moment_err(a.calculate_physical_properties(), b.calculate_physical_properties())
        print(f"  step {n}: max moment rel err {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst <= 1e-6, worst


def test_fp32_trajectory_tracks_fp64(pkg):
    """Same seed over several steps on the cavity: the FP32 run stays within
    1e-5 of the FP64 run in every moment field (pick flips near CDF
    boundaries make the two Monte Carlo trajectories drift apart slowly)."""
    cfg, f = pkg.lid_driven_cavity(nx=16, ny=16, n_v=15, vmax=4.0, samples=20000)
    print(f"[fp32 vs fp64, trajectory] {cfg.name}", flush=True)
    a = pkg.Solver(cfg)
    b = pkg.Solver(cfg.with_(precision=FP32))
    a.initialize(f)
    b.initialize(f)
    for n in range(1, 6):
        a.step(n)
        b.step(n)
        err = _moment_err(a.calculate_physical_properties(), b.calculate_physical_properties())
        print(f"  step {n}: max moment rel err {err:.2e}", flush=True)
        assert err <= 1e-5, (n, err)


def test_fp32_lds_brackets_pick_the_same_nodes(pkg, monkeypatch):
    """The persistent FP32 collision stages each cell's bucket brackets in
    LDS beside the sign words. Four full steps of a 32x32 cavity on the 25^3
    grid with the tables in LDS against the same steps with the global tables
    (DVM_FAST_LDS_TABLES=0): equal up to the summation order of the LDS
    atomics, while a single different pick (a bracket staged from the wrong
    cell or entry) would move its nodes by the sample's weight, many orders
    above that."""
    cfg, f = pkg.lid_driven_cavity(nx=32, ny=32, n_v=25, vmax=5.0, samples=2000)
    print(f"[fp32 lds brackets] {cfg.name}", flush=True)
    out = {}
    for lds in ("1", "0"):
        monkeypatch.setenv("DVM_FAST_LDS_TABLES", lds)
        with pkg.Solver(cfg.with_(precision=FP32)) as s:
            s.initialize(f)
            s.solve(4)
            out[lds] = s.download_f()
    scale = np.abs(out["0"]).max()
    d = np.abs(out["1"] - out["0"]).max() / scale
    change = np.abs(out["0"] - f).max() / scale
    print(f"  max difference {d:.2e} of max f, change {change:.2e}", flush=True)
    assert change > 1e-5, change
    assert d <= 1e-12, d
