"""Edge cases of the time step on the GPU against the oracle (1e-12 per cell,
sample counts equal): a cavity with an empty cell (zero density: no samples,
its moments zero, it fills only by convection) under both collision solvers,
the longest velocity axis the library takes (128 points) and the shortest
(3 points), and a context that is asked for more than the axis limit."""
import numpy as np
import pytest

from tests._oracle import oracle_solver
from tests.test_gpu_parity import cellwise_close, close

pytestmark = pytest.mark.gpu


def run_pair(pkg, cfg, f, steps=3):
    gpu, orc = pkg.Solver(cfg), oracle_solver(cfg)
    gpu.initialize(f)
    orc.initialize(f)
    for n in range(1, steps + 1):
        sg = gpu.compute_collision_integral(n)
        so = orc.compute_collision_integral(n)
        assert np.array_equal(sg, so), (n, sg, so)
        ok, bad, worst = cellwise_close(gpu.download_f(), orc.download_f())
        assert ok, f"step {n} collide: cells {bad[:10]} worst {worst:.3e}"
        gpu.compute_convection(n)
        orc.compute_convection(n)
        ok, bad, worst = cellwise_close(gpu.download_f(), orc.download_f())
        assert ok, f"step {n} convect: cells {bad[:10]} worst {worst:.3e}"
        ok, err, scale = close(gpu.calculate_physical_properties(), orc.calculate_physical_properties())
        assert ok, (n, err, scale)
    counts = sg
    gpu.close()
    orc.close()
    return counts


@pytest.mark.parametrize("solver", [2, 3])
def test_empty_cell(pkg, solver):
    cfg, f = pkg.lid_driven_cavity(nx=4, ny=4, n_v=9, vmax=3.0, samples=400)
    cfg = cfg.with_(coll_solver=solver)
    f = f.copy()
    f[5] = 0.0
    print(f"[edges] empty cell, coll_solver {solver}", flush=True)
    gpu = pkg.Solver(cfg)
    gpu.initialize(f)
    assert gpu.compute_collision_integral(1)[5] == 0
    gpu.close()
    run_pair(pkg, cfg, f)


@pytest.mark.parametrize("n_x", [128, 3])
def test_axis_length_limits(pkg, n_x):
    cfg, _ = pkg.homogeneous_relaxation(n_v=7, samples=4000)
    ax = pkg.uniform_axis(n_x, -5.0, 5.0)
    cfg = cfg.with_(vx=ax)
    f = (pkg.maxwellian(cfg, 0.5, 1.0, 0.0, 0.0, 1.0) + pkg.maxwellian(cfg, 0.5, -1.0, 0.0, 0.0, 2.0)).reshape(1, -1)
    print(f"[edges] x axis of {n_x} points", flush=True)
    assert run_pair(pkg, cfg, f, steps=2)[0] > 0


def test_axis_longer_than_the_limit_is_refused(pkg):
    cfg, _ = pkg.homogeneous_relaxation(n_v=5, samples=10)
    with pytest.raises(pkg.DvmError, match="3..128"):
        pkg.Solver(cfg.with_(vx=pkg.uniform_axis(129, -5.0, 5.0)))
