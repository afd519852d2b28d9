"""Wall heat fluxes on the GPU (dvm_set_wall_heat_flux_freq /
dvm_wall_heat_flux) against the oracle through the C-ABI, on every recorded
step, with diffuse and partially specular walls and a moving lid; the
recorded step follows the interval. The flux is a small difference of the
incoming and outgoing half-space sums, so the bar is 1e-12 of the size of
those sums, m/2 sum |v_n| f |v|^2, not of the flux itself."""
import numpy as np
import pytest

from tests._oracle import oracle_solver

pytestmark = pytest.mark.gpu


def flux_scale(pkg, cfg, f):
    """m/2 max over cells of sum |v_n| |f| |v|^2 (both wall normals)."""
    v = pkg.node_velocities(cfg)
    w = (v ** 2).sum(1)
    return 0.5 * cfg.mass * max(np.max(np.abs(f) @ (np.abs(v[:, 0]) * w)), np.max(np.abs(f) @ (np.abs(v[:, 1]) * w)))


@pytest.mark.parametrize("acc", [1.0, 0.6])
def test_wall_heat_flux_matches_oracle(pkg, acc):
    cfg, f = pkg.lid_driven_cavity(nx=5, ny=4, n_v=9, vmax=3.0, samples=500, t_wall=1.3)
    cfg = cfg.with_(acc_coeff=acc)
    print(f"[wall heat] cavity 5x4 9^3 acc={acc}", flush=True)
    gpu, orc = pkg.Solver(cfg), oracle_solver(cfg)
    gpu.initialize(f)
    orc.initialize(f)
    for s in (gpu, orc):
        s.set_wall_heat_flux_freq(2)
    for n in range(1, 6):
        gpu.step(n)
        orc.step(n)
        qg, ng = gpu.wall_heat_flux()
        qo, no = orc.wall_heat_flux()
        assert ng == no == (n // 2) * 2 if n >= 2 else ng == no == -1
        scale = flux_scale(pkg, cfg, orc.download_f())
        for side, along in ((0, cfg.ny), (1, cfg.ny), (2, cfg.nx), (3, cfg.nx)):
            a, b = qg[side, :along], qo[side, :along]
            assert np.max(np.abs(a - b)) <= 1e-12 * scale, (n, side, a, b, scale)
        if n >= 2:
            assert np.all(qo[:, :cfg.ny] != 0.0)
    gpu.close()
    orc.close()


def test_wall_heat_flux_needs_a_2d_grid(pkg):
    cfg, f = pkg.normal_shock_1d(nx=8, n_v=9, samples=100)
    with pkg.Solver(cfg) as s:
        with pytest.raises(pkg.DvmError, match="2-D"):
            s.set_wall_heat_flux_freq(1)
