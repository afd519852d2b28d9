"""Cells whose samples are split over several collide workgroups (the parts of
one cell run on one XCD, in a block order that is a permutation of the
(cell, part) order): every part of every cell must be processed exactly once,
for grids that are and are not multiples of the 8 XCDs. Checked against the
oracle, which has no split, with the same bars as the other step tests."""
import numpy as np
import pytest

from tests._oracle import oracle_solver
from tests.test_gpu_parity import cellwise_close, close

pytestmark = pytest.mark.gpu

CASES = {
    # name: (builder, forced split or None for the automatic choice)
    "shock_12_auto": (lambda pkg: pkg.normal_shock_1d(nx=12, n_v=11, vmax=6.0, samples=20000), None),
    "shock_12_split5": (lambda pkg: pkg.normal_shock_1d(nx=12, n_v=11, vmax=6.0, samples=20000), 5),
    "shock_13_split3": (lambda pkg: pkg.normal_shock_1d(nx=13, n_v=9, vmax=6.0, samples=9000), 3),
    "cavity_5x5_split3": (lambda pkg: pkg.lid_driven_cavity(nx=5, ny=5, n_v=9, vmax=3.0, samples=9000), 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_split_cells_match_oracle(pkg, case, monkeypatch):
    build, split = CASES[case]
    cfg, f = build(pkg)
    if split is not None:
        monkeypatch.setenv("DVM_COLL_SPLIT", str(split))
    print(f"[split] {case}: {cfg.name}, split {split or 'auto'}", flush=True)
    gpu = pkg.Solver(cfg)
    orc = oracle_solver(cfg)
    gpu.initialize(f)
    orc.initialize(f)
    for n in range(1, 3):
        sg = gpu.compute_collision_integral(n)
        so = orc.compute_collision_integral(n)
        assert np.array_equal(sg, so), "sample counts differ"
        ok, bad, worst = cellwise_close(gpu.download_f(), orc.download_f())
        assert ok, f"step {n} collide: cells {bad[:10]} worst {worst:.3e}"
        gpu.compute_convection(n)
        orc.compute_convection(n)
        ok, bad, worst = cellwise_close(gpu.download_f(), orc.download_f())
        assert ok, f"step {n} convect: cells {bad[:10]} worst {worst:.3e}"
        ok, err, scale = close(gpu.calculate_physical_properties(), orc.calculate_physical_properties())
        assert ok, (n, err, scale)


@pytest.mark.parametrize("split", [3, 7])
def test_fp32_split_equals_unsplit(pkg, split, monkeypatch):
    """FP32 collision arithmetic (no FP64 oracle at 1e-12): a split run equals
    the one-workgroup-per-cell run of the same path (same picks, stencils and
    FP32 values; only the order of the FP64 accumulation differs)."""
    cfg, f = pkg.lid_driven_cavity(nx=4, ny=5, n_v=9, vmax=3.0, samples=9000)
    cfg = cfg.with_(precision=1)
    print(f"[split fp32] 4x5 cavity, split {split} against split 1", flush=True)
    monkeypatch.setenv("DVM_COLL_SPLIT", "1")
    one = pkg.Solver(cfg)
    monkeypatch.setenv("DVM_COLL_SPLIT", str(split))
    many = pkg.Solver(cfg)
    one.initialize(f)
    many.initialize(f)
    for n in range(1, 4):
        one.step(n)
        many.step(n)
        ok, bad, worst = cellwise_close(many.download_f(), one.download_f())
        assert ok, f"step {n}: cells {bad[:10]} worst {worst:.3e}"
