"""End to end on the device: the reference's cavity deck run by the Fortran
driver through ISO_C_BINDING equals the oracle run of the same configuration
from the same initial distribution (1e-12 per cell), and a Fortran host
reads the same moments the library computes."""
import numpy as np
import pytest

from tests._deck import deck_text, have_driver, parse_check, run_driver, solver_config
from tests._oracle import oracle_solver

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_driver(), reason="dvm_gpu is not built")]


@pytest.mark.parametrize("over", [{}, {"dvm_samples": 400, "dvm_precision": 64, "nx_space": "'8'", "ny_space": "'1'",
                                      "dvm_dim": 1, "dvm_conv_order": 2, "LWallBC": "'2'", "RWallBC": "'1'",
                                      "x_vel_LW": "'0.5d0'"}])
def test_deck_run_matches_the_oracle(pkg, tmp_path, over):
    steps = 3
    print(f"[deck] cavity10_ar overrides={over} steps={steps}", flush=True)
    r = run_driver(str(tmp_path), deck_text(**over), "--steps", str(steps), "--dump-init", "f0.bin",
                   "--dump", "f.bin", "--props", "props.txt")
    c = parse_check(run_driver(str(tmp_path), deck_text(**over), "--check").stdout)
    cfg = solver_config(pkg, c)
    ncell, N = cfg.nx * cfg.ny, cfg.num_points
    f0 = np.fromfile(tmp_path / "f0.bin", dtype=np.float64).reshape(ncell, N)
    f = np.fromfile(tmp_path / "f.bin", dtype=np.float64).reshape(ncell, N)
    props = np.loadtxt(tmp_path / "props.txt").reshape(ncell, -1)
    orc = oracle_solver(cfg)
    orc.initialize(f0)
    for n in range(1, steps + 1):
        orc.step(n)
    want = orc.download_f()
    err = np.max(np.abs(f - want), axis=1) / np.max(np.abs(want), axis=1)
    assert np.all(err <= 1e-12), err.max()
    pw = orc.calculate_physical_properties()
    # norm-wise per field; velocities against the thermal speed sqrt(T)
    scale = np.max(np.abs(pw), axis=0)
    scale[1:4] = np.sqrt(np.max(pw[:, 4]))
    assert np.all(np.max(np.abs(props - pw), axis=0) <= 1e-12 * scale)
    assert "steps: 3" in r.stdout
