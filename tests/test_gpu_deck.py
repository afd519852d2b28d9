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


def test_properties_file_matches_the_oracle(pkg, tmp_path):
    """macro_properties.dat of a 4-step run with dumps every 2 steps: three
    records (t = 0, 2 dt, 4 dt) of 100 cells whose columns equal the oracle's
    moments and output moments at those steps."""
    deck = deck_text(properties_vis_dump_freq=2, write_tot_energy=1, write_entropy=1, write_mach=1,
                     write_stress=2, stress_directions="'11 12'")
    print("[deck] properties file, 4 steps", flush=True)
    run_driver(str(tmp_path), deck, "--steps", "4", "--dump-init", "f0.bin")
    text = (tmp_path / "macro_properties.dat").read_text().splitlines()
    cfg = solver_config(pkg, parse_check(run_driver(str(tmp_path), deck, "--check").stdout))
    records = []
    for ln in text[2:]:
        if ln.startswith("zone"):
            records.append([])
        else:
            records[-1].append([float(x) for x in ln.split()])
    assert [len(r) for r in records] == [100, 100, 100]
    ncell = cfg.nx * cfg.ny
    f0 = np.fromfile(tmp_path / "f0.bin", dtype=np.float64).reshape(ncell, cfg.num_points)
    orc = oracle_solver(cfg)
    orc.initialize(f0)
    done = 0
    for rec, step in zip(records, (0, 2, 4)):
        while done < step:
            done += 1
            orc.step(done)
        p = orc.calculate_physical_properties()
        e = orc.output_moments()
        got = np.array(rec)
        # records run x-major (nx outer, ny inner); cells are x-fastest
        order = [ix + cfg.nx * iy for ix in range(cfg.nx) for iy in range(cfg.ny)]
        p, e = p[order], e[order]
        dens, u, v, T = p[:, 0], p[:, 1], p[:, 2], p[:, 4]
        want = np.stack([np.full(ncell, step * cfg.dt), got[:, 1], got[:, 2], dens, u, v,
                         u / np.sqrt(0.5 * (5.0 / 3.0) * T / 1.0), T, 0.75 * dens * T, e[:, 9], e[:, 0], e[:, 1],
                         0.5 * dens * T, e[:, 3], e[:, 6]], axis=1)
        # physical scales: velocities and Mach against the thermal speed, stresses
        # against the pressure, heat fluxes against pressure x thermal speed
        pmax, cth = np.max(0.5 * dens * T), np.sqrt(np.max(T))
        scale = np.max(np.abs(want), axis=0) + 1e-300
        scale[[4, 5, 6]] = cth
        scale[[10, 11]] = pmax * cth
        scale[[13, 14]] = pmax
        assert np.all(np.max(np.abs(got - want), axis=0) <= 1e-12 * scale), (step, np.max(np.abs(got - want), axis=0) / scale)
