"""The workloads at their full sizes, checked on the GPU:

* 64x64-cell lid-driven cavity, 25^3 nodes on +-5, 20 000 samples per cell,
  hard sphere, Kn 0.1 (the FP64 LDS slice of 125 KB per cell): two FP64 steps
  against the oracle (cell-parallel on the host) at 1e-12 per cell; the FP32
  collision path one step from a developed FP64 state against the FP64 path
  in every moment field (density, velocity, temperature, heat flux, viscous
  stress, entropy) and in the L1 norm of f at 1e-5;
* 1-D normal shock, 200 cells x 31^3 nodes on +-6, VHS omega 0.81, Mach 2,
  50 000 samples per cell, minmod sweep (the windowed FP64 path, each cell
  split over several workgroups): two FP64 steps against the oracle at 1e-12;
* 4096 batched space-homogeneous instances (edges 15..41, 50 samples per
  node): one FP64 and one FP32 step of the whole batch, with sampled
  instances (the largest and smallest grids among them) against standalone
  contexts at 1e-12, and a 15^3 instance against the oracle;
* the serving mix (70 % homogeneous 15^3..41^3, 30 % shocks 50..400 cells,
  21^3..31^3): sampled requests end in the same state and step alone as in
  the mix.
"""
import ctypes as C
import os

import numpy as np
import pytest

from tests._oracle import CellParallelOracle, load_oracle, oracle_solver

pytestmark = pytest.mark.gpu

FP64, FP32, FP32R = 0, 1, 2  # DVM_PREC_FP64, DVM_PREC_FP32, DVM_PREC_FP32_REMAP32


def cellwise(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return float(np.max(np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-300)))


def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _philox_pair(lib, seed, t1, s, cell, ntime):
    """(u_A, u_B) of sample s of integral t1 (1-based) of a cell: words 0 and 1
    of the Philox block (s, t1, cell, ntime) keyed by the seed."""
    out = (C.c_uint32 * 4)()
    ctr = (C.c_uint32 * 4)(int(s), int(t1), int(cell), int(ntime))
    key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    lib.orc_philox4x32_10(ctr, key, out)
    return out[0] * 2.0 ** -32, out[1] * 2.0 ** -32


def explain_differing_cells(pkg, cfg, f, cells, ntime=1):
    """explain_differing_cell for several cells on one pair of contexts."""
    if len(cells) == 0:
        return 0
    lib = load_oracle()
    g = pkg.Solver(cfg)
    g.initialize(f)
    o = oracle_solver(cfg)
    o.initialize(f)
    try:
        return sum(explain_differing_cell(lib, cfg, g, o, c, ntime) for c in cells)
    finally:
        g.close()
        o.close()


def explain_differing_cell(lib, cfg, g, o, cell, ntime):
    """Traces one cell's collision phase in both codes from the same uploaded
    state (both rebuild the equilibrium at their first collision) and checks
    that they differ only in picks whose scaled uniform u * total lies within
    rounding of the CDF value between the two picked nodes (the reference's
    sequential sums and the device's parallel scans round the CDF
    differently; Q5's difference of equilibrium prefix sums carries that
    rounding into the small deviation CDF). Returns the number of such
    samples; every other record (picks, skip flag, both stencils) must be
    identical."""
    N = cfg.num_points
    ig, _, n = g.collide_trace(ntime, int(cell), 10 ** 6)
    io, _, n2 = o.collide_trace(ntime, int(cell), 10 ** 6)
    eqc, neqc = np.zeros(N + 1), np.zeros(N + 1)
    tabs = [np.zeros(1000, np.int32) for _ in range(4)]
    cnt = np.zeros(3, np.int32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    I = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    assert lib.orc_cell_cdf(o.h, int(cell), P(eqc), P(neqc), *(I(t) for t in tabs), I(cnt)) == 0
    assert n == n2 and len(ig) == n
    diff = np.where(np.any(ig[:, :19] != io[:, :19], axis=1))[0]
    for q in diff:
        t1, s = int(io[q, 0]), int(io[q, 1])
        assert ig[q, 0] == t1 and ig[q, 1] == s
        uA, uB = _philox_pair(lib, cfg.seed, t1, s, cell, ntime)
        moved = False
        for col, u, eq_side in ((2, uA, t1 == 1), (3, uB, t1 == 2)):
            a, b = int(ig[q, col]), int(io[q, col])
            if a == b:
                continue
            moved = True
            cum = eqc if eq_side else neqc
            adj = u * cum[-1]
            lo, hi = min(a, b), max(a, b)
            gap = max(abs(adj - cum[lo + 1]), abs(adj - cum[hi])) / cum[-1]
            print(f"    cell {cell} sample {q}: picks {a} / {b}, |u total - CDF| = {gap:.2e} of the total",
                  flush=True)
            assert gap <= 1e-10, (cell, q, gap)
        assert moved, (cell, q, "records differ with equal picks")
    return len(diff)


def run_against_oracle(pkg, cfg, f, steps):
    """Full steps (collide, convect, moments) in both codes. Step 1 (from the
    uploaded state) per cell after the collision phase: every cell within
    1e-12, except cells where a pick flipped at a CDF rounding boundary, which
    explain_differing_cell verifies from traces. Then every step: at least 75 %
    of the cells within 1e-12 after the full step (a 200-cell shock with three
    flipped cells per step and a sweep that spreads each to its neighbours
    leaves about 85 % after two steps; the cavity 98.7 %), the rest within the
    weight of a few samples."""
    gpu = pkg.Solver(cfg)
    gpu.initialize(f)
    orc = CellParallelOracle(cfg, threads=host_threads())
    orc.initialize(f)
    try:
        gpu.compute_collision_integral(1)
        fc = orc.collide(1)
        err = cellwise_all(gpu.download_f(), fc)
        bad = np.where(err > 1e-12)[0]
        print(f"  step 1 collision: {len(bad)} of {len(err)} cells over 1e-12 (max {err.max():.2e})", flush=True)
        assert len(bad) <= max(4, len(err) // 200)
        flipped = explain_differing_cells(pkg, cfg, f, bad)
        assert flipped <= 2 * len(bad)
        gpu.compute_convection(1)
        gpu.calculate_physical_properties()
        orc.finish_step(1)
        for n in range(1, steps + 1):
            if n > 1:
                gpu.step(n)
                orc.step(n)
            pg = gpu.calculate_physical_properties()
            ef = cellwise_all(gpu.download_f(), orc.download_f())
            ep = cellwise_all(pg, orc.props)
            frac = float(np.mean(ef <= 1e-12))
            print(f"  step {n}: cells within 1e-12 {frac:.4f}, max f {ef.max():.2e}, moments {ep.max():.2e}",
                  flush=True)
            assert frac >= 0.75 and ef.max() <= 1e-5 and ep.max() <= 1e-5, (n, frac, ef.max(), ep.max())
            if n == 1:
                assert len(bad) > 0 or ef.max() <= 1e-12
    finally:
        gpu.close()
        orc.close()


def cellwise_all(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-300)


def test_cavity_64x64_full_size_matches_oracle(pkg):
    cfg, f = pkg.lid_driven_cavity()
    assert (cfg.nx, cfg.ny, cfg.nvx, cfg.samples_per_cell) == (64, 64, 25, 20000)
    print(f"[config] {cfg.name}, FP64, 2 steps against the oracle on {host_threads()} host threads", flush=True)
    run_against_oracle(pkg, cfg.with_(precision=FP64), f, 2)


def test_shock_200_cells_full_size_matches_oracle(pkg):
    cfg, f = pkg.normal_shock_1d()
    cfg = cfg.with_(conv_order=pkg.capi.DVM_CONV_MINMOD2, precision=FP64)
    assert (cfg.nx, cfg.nvx, cfg.samples_per_cell) == (200, 31, 50000)
    print(f"[config] {cfg.name}, minmod, FP64, 2 steps against the oracle", flush=True)
    run_against_oracle(pkg, cfg, f, 2)


@pytest.mark.parametrize("prec32", [FP32, FP32R], ids=["fp64_remap", "fp32_remap"])
def test_cavity_64x64_fp32_step_every_moment_within_1e5(pkg, prec32):
    """From a developed FP64 state (200 steps: the lid has set up shear and heat
    flux), one FP32-collision step (remap in FP64, or in FP32) against one FP64
    step with the same seed: density, temperature and entropy relative per
    cell; velocity in units of the thermal speed; heat flux and viscous stress
    relative to the largest magnitude of the field over the cavity (both vanish
    at equilibrium); sum |f32 - f64| / sum |f64|. All <= 1e-5."""
    cfg, f = pkg.lid_driven_cavity()
    print(f"[config] {cfg.name}: precision {prec32} vs FP64, one step from a developed state", flush=True)
    with pkg.Solver(cfg.with_(precision=FP64)) as s:
        s.initialize(f)
        for n in range(1, 201):
            s.step(n)
        state = s.download_f()
    res = {}
    for name, prec in (("64", FP64), ("32", prec32)):
        with pkg.Solver(cfg.with_(precision=prec)) as s:
            s.initialize(state)
            s.step(201)
            res[name] = (s.calculate_physical_properties(), s.output_moments(), s.download_f())
    (p64, e64, f64), (p32, e32, f32) = res["64"], res["32"]
    th = np.sqrt(p64[:, 4])

    def deviatoric(p, e):
        # tau_ij - p delta_ij: the viscous part, which vanishes at equilibrium
        t = e[:, 3:9].copy()
        t[:, 0:3] -= p[:, 6:7]
        return t

    d32, d64 = deviatoric(p32, e32), deviatoric(p64, e64)
    err = {
        "density": np.max(np.abs(p32[:, 0] - p64[:, 0]) / p64[:, 0]),
        "velocity": np.max(np.abs(p32[:, 1:4] - p64[:, 1:4]).max(1) / th),
        "temperature": np.max(np.abs(p32[:, 4] - p64[:, 4]) / p64[:, 4]),
        "heat_flux": np.max(np.abs(e32[:, 0:3] - e64[:, 0:3])) / np.max(np.abs(e64[:, 0:3])),
        "viscous_stress": np.max(np.abs(d32 - d64)) / np.max(np.abs(d64)),
        "entropy": np.max(np.abs(e32[:, 9] - e64[:, 9]) / np.abs(e64[:, 9])),
        "f_l1": np.abs(f32 - f64).sum() / np.abs(f64).sum(),
    }
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert np.max(np.abs(e64[:, 0:3])) > 1e-3  # the developed state has heat flux to compare
    for k, v in err.items():
        assert v <= 1e-5, (k, v)


@pytest.mark.parametrize("prec32", [FP32, FP32R], ids=["fp64_remap", "fp32_remap"])
def test_cavity_64x64_fp32_collision_conserves(pkg, prec32):
    cfg, f = pkg.lid_driven_cavity(precision=prec32)
    v = pkg.node_velocities(cfg)
    print(f"[config] {cfg.name}: precision {prec32} collision conservation, 2 collision phases", flush=True)

    def cons(g, vel):
        return np.stack([g.sum(1), g @ vel[:, 0], g @ vel[:, 1], g @ vel[:, 2], g @ (vel ** 2).sum(1)], axis=1)

    with pkg.Solver(cfg) as s:
        s.initialize(f)
        for n in (1, 2):
            s.step(n)  # moves the state off the uniform start
        for n in (3, 4):
            g0 = s.download_f()
            s.compute_collision_integral(n)
            g1 = s.download_f()
            s.calculate_physical_properties()
            # round-off scale of each moment: sum |f| |phi(v)|
            scale = cons(np.abs(g0), np.abs(v))
            rel = np.max(np.abs(cons(g1, v) - cons(g0, v)) / scale)
            print(f"  collide {n}: max conservation defect {rel:.2e}", flush=True)
            assert rel <= 1e-13, (n, rel)
            s.compute_convection(n)
            s.calculate_physical_properties()


def _batch_check(pkg, precision, picks, steps=1):
    inst = pkg.batched_homogeneous(4096, precision=precision)
    nodes = sum(c.num_points for c, _ in inst)
    edges = np.array([c.nvx for c, _ in inst])
    print(f"[config] batch of {len(inst)} instances ({nodes} nodes), precision {precision}, {steps} step(s)",
          flush=True)
    b = pkg.Batch(inst[0][0], max_slots=len(inst), node_capacity=nodes, cell_capacity=len(inst))
    try:
        slots = [b.admit(c, f) for c, f in inst]
        b.step(steps)
        b.synchronize()
        sel = sorted(set([int(np.argmax(edges)), int(np.argmin(edges))] + list(picks)))
        for i in sel:
            c, f = inst[i]
            got = b.download_f(slots[i])
            with pkg.Solver(c) as s:
                s.initialize(f)
                for n in range(1, steps + 1):
                    s.step(n)
                ref = s.download_f()
            err = cellwise(got, ref)
            print(f"  instance {i} ({c.nvx}^3): {err:.2e}", flush=True)
            assert err <= 1e-12, (i, err)
    finally:
        b.close()
    return inst


def test_batch_4096_instances_fp64_match_standalone(pkg):
    inst = _batch_check(pkg, FP64, picks=(7, 1000, 2047, 4095))
    # and one small instance against the oracle, two steps
    i = int(np.argmin([c.nvx for c, _ in inst]))
    c, f = inst[i]
    gpu = pkg.Solver(c)
    orc = oracle_solver(c)
    gpu.initialize(f)
    orc.initialize(f)
    for n in (1, 2):
        gpu.step(n)
        orc.step(n)
    err = cellwise(gpu.download_f(), orc.download_f())
    print(f"  instance {i} ({c.nvx}^3) against the oracle: {err:.2e}", flush=True)
    assert err <= 1e-12
    gpu.close()
    orc.close()


def test_batch_4096_instances_fp32_match_standalone(pkg):
    _batch_check(pkg, FP32, picks=(3, 1500, 4000))


@pytest.mark.parametrize("precision", [FP32, FP64])
def test_cavity_64x64_as_a_batch_slot_matches_standalone(pkg, precision):
    """The 64x64 cavity admitted into a batch (its cells in the shared CDF,
    collision and moment kernels, its walls and 2-D sweep launched for the
    slot inside the replayed graph) equals the standalone context after two
    steps (1e-12: FP32 runs receive identical FP32 inputs in the first step;
    the atomics order can move the last bit of the second step's inputs)."""
    cfg, f = pkg.lid_driven_cavity(precision=precision)
    print(f"[config] {cfg.name} as a batch slot, precision {precision}, 2 steps", flush=True)
    b = pkg.Batch(cfg, max_slots=2, node_capacity=cfg.nx * cfg.ny * cfg.num_points, cell_capacity=cfg.nx * cfg.ny)
    try:
        s = b.admit(cfg, f)
        b.step(2)
        b.synchronize()
        got = b.download_f(s)
    finally:
        b.close()
    with pkg.Solver(cfg) as ref:
        ref.initialize(f)
        ref.step(1)
        ref.step(2)
        want = ref.download_f()
    err = cellwise_all(got, want)
    print(f"  cells within 1e-12 {np.mean(err <= 1e-12):.4f}, max {err.max():.2e}", flush=True)
    assert np.mean(err <= 1e-12) >= 0.99 and err.max() <= 1e-5


class TickClock:
    def __init__(self, tick):
        self.t = 0.0
        self.tick = tick

    def __call__(self):
        self.t += self.tick
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


def test_serving_mix_alone_equals_in_the_mix(pkg):
    """The configured serving workload (Poisson arrivals; 70 % homogeneous
    mixtures with edges 15..41 and 50 samples per node, 30 % shocks of
    50..400 cells with edges 21..31 and 20 000 samples per cell, run to steady
    state), 36 requests in FP64, on a simulated clock fast enough that many
    solves are resident together. Sampled requests are run again alone with
    the same seed and alone with another seed. The device's floating-point
    atomics add in an order that varies from run to run, so the alone and the
    mix runs of a request differ in the last bits until a pick lands on a CDF
    rounding boundary; from there they are two realisations of one Monte Carlo
    trajectory. Checked: the same stop reason; homogeneous requests stop at the
    same step with density, velocity and temperature equal to 1e-12 (the
    collisions conserve them exactly); shocks stop within three windows (300
    steps) of each other; and f (homogeneous) or the density and temperature
    profiles (shocks) of alone and mix differ by at most twice the difference
    between two seeds (for shocks the larger of two other seeds' differences),
    and by 1e-12 when no pick moved."""
    reqs = pkg.synthetic_requests(count=36, rate=20.0, seed=5, precision=FP64)
    pools = dict(node_capacity=160_000_000, cell_capacity=65536, max_slots=512)

    def serve(rs):
        clock = TickClock(0.05)
        with pkg.Server(rs[0].cfg, check_every=10, keep_f=True, keep_moments=True, clock=clock,
                        sleep=clock.sleep, **pools) as srv:
            return srv.run(rs)

    def reseeded(r, k=1):
        return pkg.Request(r.rid, r.arrival, r.cfg.with_(seed=r.cfg.seed + 99991 * k), r.f0, r.kind, r.tol,
                           r.max_steps)

    print("[config] serving mix, 36 requests", flush=True)
    st = serve(reqs)
    mix = {o.rid: o for o in st.outcomes}
    reasons = sorted(set((o.kind, o.reason) for o in st.outcomes))
    print(f"  max resident {st.max_resident}, reasons {reasons}", flush=True)
    assert st.max_resident >= 4
    assert ("shock", "steady") in reasons
    shocks = [r for r in reqs if r.kind == "shock"]
    homs = sorted((r for r in reqs if r.kind == "homogeneous"), key=lambda r: r.cfg.num_points)
    for r in (homs[0], homs[-1]):
        alone, other, o = serve([r]).outcomes[0], serve([reseeded(r)]).outcomes[0], mix[r.rid]
        d_mix, d_seed = cellwise(alone.f, o.f), cellwise(alone.f, other.f)
        dm = cellwise(alone.props[:, :5], o.props[:, :5])
        print(f"  request {r.rid} (homogeneous {r.cfg.nvx}^3): {alone.steps} steps ({alone.reason}) alone, "
              f"{o.steps} ({o.reason}) in the mix; f {d_mix:.2e} (seed spread {d_seed:.2e}), moments {dm:.2e}",
              flush=True)
        assert alone.steps == o.steps and alone.reason == o.reason
        assert dm <= 1e-12
        assert d_mix <= max(1e-12, 2.0 * d_seed), (r.rid, d_mix, d_seed)

    def profile_diff(a, b):
        dn = np.max(np.abs(a[:, 0] - b[:, 0])) / np.max(b[:, 0])
        dt = np.max(np.abs(a[:, 4] - b[:, 4])) / np.max(b[:, 4])
        return max(dn, dt)

    # A steady shock's position wanders with the Monte Carlo noise, so the
    # profile difference of two realisations is heavy-tailed: the mix is
    # compared with the larger spread of two other seeds.
    for r in (shocks[0], max(shocks, key=lambda r: r.cfg.nx * r.cfg.num_points)):
        alone, o = serve([r]).outcomes[0], mix[r.rid]
        others = [serve([reseeded(r, k)]).outcomes[0] for k in (1, 2)]
        other = max(others, key=lambda x: profile_diff(alone.props, x.props))
        d_mix, d_seed = profile_diff(alone.props, o.props), profile_diff(alone.props, other.props)
        print(f"  request {r.rid} (shock {r.cfg.nx} cells x {r.cfg.nvx}^3): {alone.steps} steps ({alone.reason}) "
              f"alone, {o.steps} ({o.reason}) in the mix, {other.steps} with another seed; profile difference "
              f"{d_mix:.2e}, seed spread {d_seed:.2e}", flush=True)
        assert alone.reason == o.reason and abs(alone.steps - o.steps) <= 300
        assert d_mix <= max(1e-12, 2.0 * d_seed), (r.rid, d_mix, d_seed)
