"""The cell-parallel oracle harness (tests/_oracle.py: CellParallelOracle),
which the full-size GPU checks use, equals the single-context oracle step
bit for bit: a cell's collision reads only that cell's f, moments and
equilibrium cache, and its samples are keyed by the global cell index."""
import numpy as np
import pytest

from tests._oracle import CellParallelOracle, oracle_solver


@pytest.mark.parametrize("case", ["cavity", "shock"])
def test_parallel_oracle_equals_single_context(pkg, case):
    if case == "cavity":
        cfg, f = pkg.lid_driven_cavity(nx=6, ny=5, n_v=9, vmax=4.0, samples=500)
    else:
        cfg, f = pkg.normal_shock_1d(nx=9, n_v=9, samples=400)
        cfg = cfg.with_(conv_order=2)
    a = oracle_solver(cfg)
    a.initialize(f)
    b = CellParallelOracle(cfg, threads=4)
    b.initialize(f)
    for n in (1, 2, 3):
        a.step(n)
        pb = b.step(n)
        assert np.array_equal(a.download_f(), b.download_f()), n
        assert np.array_equal(a.calculate_physical_properties(), pb), n
    b.close()
