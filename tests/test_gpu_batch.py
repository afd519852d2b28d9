"""Batched solves (dvm_batch_*): every resident instance follows the trajectory
that a standalone context with the same config produces ("alone and in the
mix"), through admissions, retirements and autotuning.

Bars: FP64, per cell max|a - b| <= 1e-12 max|b| (the two runs differ only in
the order of floating-point atomic additions). FP32 collision arithmetic: the
same bar after one step (identical FP32 inputs); afterwards 1e-6, because a
last-bit FP64 difference can move an FP32 rounding of the next step's inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def cellwise(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    err = np.max(np.abs(a - b), axis=1)
    scale = np.maximum(np.max(np.abs(b), axis=1), 1e-300)
    return float(np.max(err / scale))


def instances(pkg, precision=0):
    """Uneven mix sharing one species: two mixtures in LDS, a 29^3 grid that
    takes global atomics over several work items, a 1-D shock with the
    second-order sweep."""
    rng = np.random.default_rng(5)
    out = [
        pkg.mixture_instance(rng, 11, samples_per_node=4, seed=101, precision=precision),
        pkg.mixture_instance(rng, 15, samples_per_node=3, seed=102, precision=precision),
    ]
    c, f = pkg.homogeneous_relaxation(n_v=29, samples=40000, seed=103)
    out.append((c.with_(precision=precision), f))
    c, f = pkg.normal_shock_1d(nx=8, n_v=9, vmax=6.0, samples=600, omega=0.5, seed=104)
    out.append((c.with_(precision=precision, conv_order=2), f))
    return out


def standalone(pkg, cfg, f, steps):
    with pkg.Solver(cfg) as s:
        s.initialize(f)
        for n in range(1, steps + 1):
            s.step(n)
        return s.download_f()


def make_batch(pkg, insts, max_slots=8, item=16384):
    b = pkg.Batch(insts[0][0], max_slots=max_slots, node_capacity=200_000, cell_capacity=64)
    b.set_tiles(1024, 1, item)
    return b


def test_instances_match_standalone_in_the_mix(pkg):
    insts = instances(pkg)
    print("[batch] 11^3, 15^3, 29^3 (global atomics, 16384-sample items), 1-D shock 8x9^3; 3 steps", flush=True)
    with make_batch(pkg, insts) as b:
        slots = [b.admit(c, f) for c, f in insts]
        infos = [b.slot_info(s) for s in slots]
        assert [i["in_lds"] for i in infos] == [True, True, False, True]
        b.step(3)
        b.synchronize()
        for (c, f), s in zip(insts, slots):
            assert b.slot_info(s)["steps"] == 3
            got = b.download_f(s)
            want = standalone(pkg, c, f, 3)
            assert cellwise(got, want) <= RTOL, (c.name, cellwise(got, want))
            with pkg.Solver(c) as ref:
                ref.initialize(want)
                pm = ref.calculate_physical_properties()
            assert cellwise(b.moments(s), pm) <= RTOL


def test_fp32_cdf_kernel_switch_keeps_equilibrium(pkg):
    """A batch on the fast FP32 path stores no f_eq in its many-cell CDF pass.
    With a 1100-cell 1-D shock resident, the first steps run that pass; once
    the shock retires, the few-cell pass runs, which reads f_eq back and so
    must compute it again on a cache hit (the relaxing instance, at zero bulk
    velocity and with its moments conserved, keeps its cached equilibrium
    every step). The instance still follows its standalone FP32 trajectory."""
    relax = pkg.homogeneous_relaxation(n_v=9, samples=2000, seed=111)
    relax = (relax[0].with_(precision=1), relax[1])
    c, f = pkg.normal_shock_1d(nx=1100, n_v=7, vmax=6.0, samples=100, omega=0.5, seed=112)
    shock = (c.with_(precision=1), f)
    print("[batch] FP32: 1101 cells (many-cell CDF pass), then 1 cell (few-cell pass)", flush=True)
    with pkg.Batch(relax[0], max_slots=4, node_capacity=400_000, cell_capacity=1200) as b:
        b.set_tiles(1024, 1, 16384)
        r = b.admit(*relax)
        sh = b.admit(*shock)
        b.step(2)
        b.synchronize()
        b.retire(sh)
        b.step(2)
        b.synchronize()
        got = b.download_f(r)
    want = standalone(pkg, relax[0], relax[1], 4)
    assert cellwise(got, want) <= 1e-6, cellwise(got, want)


def test_admit_and_retire_between_steps(pkg):
    insts = instances(pkg)
    print("[batch] admit / retire between steps", flush=True)
    with make_batch(pkg, insts) as b:
        a = b.admit(*insts[0])
        g = b.admit(*insts[2])
        sh = b.admit(*insts[3])
        b.step(2)
        b.retire(g)
        late = b.admit(*insts[1])
        b.step(2)
        assert b.slot_info(a)["steps"] == 4 and b.slot_info(late)["steps"] == 2
        assert cellwise(b.download_f(a), standalone(pkg, *insts[0], 4)) <= RTOL
        assert cellwise(b.download_f(sh), standalone(pkg, *insts[3], 4)) <= RTOL
        assert cellwise(b.download_f(late), standalone(pkg, *insts[1], 2)) <= RTOL
        # the retired range is reusable
        again = b.admit(*insts[2])
        b.step(1)
        assert cellwise(b.download_f(again), standalone(pkg, *insts[2], 1)) <= RTOL


def test_tuner_keeps_the_state(pkg):
    insts = instances(pkg)
    print("[batch] autotune on the resident set", flush=True)
    with make_batch(pkg, insts) as b:
        slots = [b.admit(c, f) for c, f in insts]
        b.step(1)
        before = [b.download_f(s) for s in slots]
        tiles, ms = b.tune(1)
        assert tiles["lanes"] in (256, 512, 768, 1024) and tiles["pick_k"] in (1, 2) and ms > 0.0
        for s, x in zip(slots, before):
            assert np.array_equal(b.download_f(s), x)
        b.step(1)
        for (c, f), s in zip(insts, slots):
            assert cellwise(b.download_f(s), standalone(pkg, c, f, 2)) <= RTOL


def test_admission_reports_a_full_pool(pkg):
    c, f = pkg.homogeneous_relaxation(n_v=11, samples=1000)
    print("[batch] admission beyond the node pool", flush=True)
    with pkg.Batch(c, max_slots=2, node_capacity=2000, cell_capacity=4) as b:
        s0 = b.admit(c, f)
        assert b.try_admit(c, f) is None  # 2 x 1331 nodes > 2000
        with pytest.raises(pkg.DvmError) as e:
            b.admit(c, f)
        assert e.value.code == -3  # DVM_E_NOMEM
        b.retire(s0)
        assert b.try_admit(c, f) is not None
        assert b.capacity()["free_slots"] == 1


def test_l1_distance_from_equilibrium(pkg):
    rng = np.random.default_rng(9)
    c, f = pkg.mixture_instance(rng, 13, samples_per_node=20, seed=7)
    print("[batch] L1 to the discrete Maxwellian", flush=True)
    with pkg.Batch(c, max_slots=3, node_capacity=10_000, cell_capacity=4) as b:
        s = b.admit(c, f)
        l1 = b.l1()
        assert np.all(l1[[q for q in range(3) if q != s]] == -1.0)
        p = b.moments(s)[0]
        m = pkg.maxwellian(c, p[0], p[1], p[2], p[3], p[4], renormalise=True)
        want = np.abs(f[0] - m).sum() / p[0]
        assert abs(l1[s] - want) <= 1e-12 * want
        b.step(20)
        assert b.l1()[s] < l1[s]


def test_fp32_batch_matches_fp32_standalone(pkg):
    insts = instances(pkg, precision=1)
    print("[batch] FP32 collision arithmetic, 3 steps", flush=True)
    with make_batch(pkg, insts) as b:
        slots = [b.admit(c, f) for c, f in insts]
        b.step(1)
        for (c, f), s in zip(insts, slots):
            assert cellwise(b.download_f(s), standalone(pkg, c, f, 1)) <= RTOL, c.name
        b.step(2)
        for (c, f), s in zip(insts, slots):
            assert cellwise(b.download_f(s), standalone(pkg, c, f, 3)) <= 1e-6, c.name


def test_full_monte_carlo_instances_match_standalone(pkg):
    """The batch with the full Monte Carlo solver (iColl = 2): an LDS mixture,
    a windowed 29^3 grid and a 1-D shock equal standalone contexts."""
    insts = [(c.with_(coll_solver=pkg.capi.DVM_COLL_FMC), f) for c, f in instances(pkg)]
    insts = [insts[0], insts[2], insts[3]]
    print("[batch] full Monte Carlo: 11^3, 29^3, 1-D shock; 3 steps", flush=True)
    with make_batch(pkg, insts) as b:
        slots = [b.admit(c, f) for c, f in insts]
        b.step(3)
        b.synchronize()
        for (c, f), s in zip(insts, slots):
            got = b.download_f(s)
            want = standalone(pkg, c, f, 3)
            assert cellwise(got, want) <= RTOL, (c.name, cellwise(got, want))


def test_2d_slots_match_standalone_in_the_mix(pkg):
    """2-D cavities as batch slots (their walls and 2-D sweep launched per slot
    inside the step's graph) next to a homogeneous instance and a 1-D shock:
    diffuse walls with a moving lid, partially specular walls, Couette sides;
    every slot equals a standalone context to 1e-12 after three steps,
    admitted and retired around each other."""
    rng = np.random.default_rng(9)
    cav, fc = pkg.lid_driven_cavity(nx=6, ny=5, n_v=9, vmax=4.0, samples=800, seed=301)
    spec, fs = pkg.lid_driven_cavity(nx=4, ny=7, n_v=11, vmax=4.0, samples=600, seed=302)
    spec = spec.with_(acc_coeff=0.6)
    cou, fo = pkg.lid_driven_cavity(nx=5, ny=4, n_v=9, vmax=4.0, samples=500, seed=303)
    cou = cou.with_(bc_left=pkg.capi.DVM_BC_COUETTE, bc_right=pkg.capi.DVM_BC_COUETTE)
    hom = pkg.mixture_instance(rng, 11, samples_per_node=4, seed=304)
    shk = pkg.normal_shock_1d(nx=8, n_v=9, vmax=6.0, samples=600, omega=0.5, seed=305)
    hom = (hom[0].with_(omega=0.5), hom[1])
    insts = [(cav, fc), (spec, fs), (cou, fo), hom, shk]
    print("[batch] 2-D cavities (diffuse + lid, partially specular, Couette) with a homogeneous and a 1-D slot",
          flush=True)
    b = pkg.Batch(cav, max_slots=8, node_capacity=200_000, cell_capacity=128)
    try:
        slots = [b.admit(c, f) for c, f in insts]
        b.step(2)
        b.retire(slots[1])
        again = b.admit(spec, fs)
        b.step(1)
        for (c, f), s in zip(insts, slots):
            if s == slots[1]:
                continue
            err = cellwise(b.download_f(s), standalone(pkg, c, f, 3))
            print(f"  {c.name}: {err:.2e}", flush=True)
            assert err <= RTOL, (c.name, err)
        err = cellwise(b.download_f(again), standalone(pkg, spec, fs, 1))
        assert err <= RTOL, err
    finally:
        b.close()
