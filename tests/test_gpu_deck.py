"""End to end on the device: the reference's cavity deck run by the Fortran
driver through ISO_C_BINDING equals the oracle run of the same configuration
from the same initial distribution (1e-12 per cell), and a Fortran host
reads the same moments the library computes."""
import numpy as np
import pytest

from tests._deck import deck_text, have_driver, parse_check, run_driver, solver_config, write_grid_file
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


def _compare_run(pkg, tmp_path, deck, steps, grid=None, axes=None):
    run_driver(str(tmp_path), deck, "--steps", str(steps), "--dump-init", "f0.bin", "--dump", "f.bin", grid=grid)
    cfg = solver_config(pkg, parse_check(run_driver(str(tmp_path), deck, "--check", grid=grid).stdout))
    if axes is not None:
        cfg = cfg.with_(vx=axes[0], vy=axes[1], vz=axes[2])
    ncell, N = cfg.nx * cfg.ny, cfg.num_points
    f0 = np.fromfile(tmp_path / "f0.bin", dtype=np.float64).reshape(ncell, N)
    f = np.fromfile(tmp_path / "f.bin", dtype=np.float64).reshape(ncell, N)
    orc = oracle_solver(cfg)
    orc.initialize(f0)
    for n in range(1, steps + 1):
        orc.step(n)
    want = orc.download_f()
    err = np.max(np.abs(f - want), axis=1) / np.max(np.abs(want), axis=1)
    return cfg, err


def test_shock_deck_initc4_matches_the_oracle(pkg, tmp_path):
    """The deck's own shock set-up (InitC = 4: Rankine-Hugoniot jump, quadratic
    profile) as a 1-D run with in/out flow, 3 steps."""
    deck = deck_text(nx_space="'20'", ny_space="'1'", dvm_dim=1, LWallBC="'2'", RWallBC="'2'", InitC="'4'",
                     spatial_middle="'10'", shock_thickness="'6'", dens_RW="'1.0d0'", x_vel_RW="'-1.5d0'",
                     tr_temp_RW="'1.0d0'", xv_min="'-6.0d0'", xv_max="'6.0d0'", num_points_x="'15'",
                     num_points_y="'9'", num_points_z="'9'", dvm_samples=500)
    print("[deck] InitC 4 stationary shock, 20 cells, 3 steps", flush=True)
    cfg, err = _compare_run(pkg, tmp_path, deck, 3)
    assert cfg.dim == 1 and cfg.wall[0][0] > 1.5
    assert np.all(err <= 1e-12), err.max()


def test_velocity_grid_file_run_matches_the_oracle(pkg, tmp_path):
    """A cavity on non-uniform velocity axes read from the velocity-grid file
    (the remap's bracket walk and bisection, the boundary stencils and the
    quadrature weights on uneven spacing), 3 steps."""
    axes = (np.array([-4.0, -2.8, -1.9, -1.2, -0.6, 0.0, 0.55, 1.15, 1.85, 2.7, 4.0]),
            np.array([-3.6, -2.0, -1.0, -0.35, 0.35, 1.0, 2.0, 3.6]),
            np.array([-3.5, -2.25, -1.25, -0.5, 0.0, 0.5, 1.25, 2.25, 3.5]))
    write_grid_file(tmp_path / "grid.dat", axes)
    deck = deck_text(x_vel_grid_type="'0'", y_vel_grid_type="'0'", z_vel_grid_type="'0'", nx_space="'5'",
                     ny_space="'4'", dvm_samples=600)
    print("[deck] cavity 5x4 on a non-uniform 11x8x9 grid file, 3 steps", flush=True)
    cfg, err = _compare_run(pkg, tmp_path, deck, 3, grid=str(tmp_path / "grid.dat"), axes=axes)
    assert cfg.nvx == 11 and cfg.nvy == 8 and cfg.nvz == 9
    assert np.all(err <= 1e-12), err.max()


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


def test_restart_continues_the_run(pkg, tmp_path):
    """write_rst = 1 writes the restart file at the last step (ntime, then one
    record per cell equal to the resident f); a run that reads it and steps
    on to 5 equals a straight 5-step run (1e-12 per cell: the order of the
    floating-point atomics may differ between runs)."""
    from tests.test_deck_driver import read_fortran_records
    small = dict(nx_space="'4'", ny_space="'3'", dvm_samples=300)
    print("[deck] restart: 3 steps + 2 steps vs 5 steps", flush=True)
    run_driver(str(tmp_path), deck_text(write_rst=1, write_restart_filename="a.rst", **small), "--steps", "3",
               "--dump", "f3.bin")
    recs = read_fortran_records(tmp_path / "a.rst")
    assert np.frombuffer(recs[0], dtype=np.int32)[0] == 3 and len(recs) == 1 + 12
    f3 = np.fromfile(tmp_path / "f3.bin").reshape(12, -1)
    assert np.array_equal(np.stack([np.frombuffer(r, dtype=np.float64) for r in recs[1:]]), f3)
    r = run_driver(str(tmp_path), deck_text(read_rst=1, read_restart_filename="a.rst", **small), "--steps", "5",
                   "--dump", "f5r.bin")
    assert "steps: 2" in r.stdout
    run_driver(str(tmp_path), deck_text(**small), "--steps", "5", "--dump", "f5.bin")
    a = np.fromfile(tmp_path / "f5r.bin").reshape(12, -1)
    b = np.fromfile(tmp_path / "f5.bin").reshape(12, -1)
    err = np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)
    assert np.all(err <= 1e-12), err.max()


@pytest.mark.parametrize("dim", [3, 2])
def test_distribution_function_records(pkg, tmp_path, dim):
    """write_df records of a 2-step run at every step: zones at t = 0, dt,
    2 dt; f equals the resident distribution (16 printed digits) and the
    deviation is f - beta^3 f_M(n, u, T) with the reference's pi at the
    cell's moments, taken at z = 0 on the plane output
    (visualization.f90:1085-1135)."""
    over = dict(enable_vel_df_vis=1, df_output_dimension=dim, num_output_locations=1, x_output_locations="'4'",
                y_output_locations="'7'", df_vis_dump_freq=1, nx_space="'6'", ny_space="'8'", dvm_samples=300)
    print(f"[deck] df output dim {dim}", flush=True)
    run_driver(str(tmp_path), deck_text(**over), "--steps", "2", "--dump-init", "f0.bin", "--dump", "f2.bin")
    text = (tmp_path / "dist_func_1_4_7.dat").read_text()          # --check below rewrites the header
    cfg = solver_config(pkg, parse_check(run_driver(str(tmp_path), deck_text(**over), "--check").stdout))
    zones, rows = [], []
    for ln in text.splitlines()[2:]:
        if ln.startswith("zone"):
            zones.append(ln)
            rows.append([])
        else:
            rows[-1].append([float(x) for x in ln.split()])
    assert len(zones) == 3
    cell, N, nv = 3 + cfg.nx * 6, cfg.num_points, cfg.nvx
    v = pkg.node_velocities(cfg)
    beta = [np.concatenate([[a[1] - a[0]], (a[2:] - a[1:-1]) / 2 + (a[1:-1] - a[:-2]) / 2, [a[-1] - a[-2]]])
            for a in (cfg.vx, cfg.vy, cfg.vz)]
    b3 = (beta[2][:, None, None] * beta[1][None, :, None] * beta[0][None, None, :]).reshape(-1)
    kz = max(k for k in range(cfg.nvz) if cfg.vz[k] <= 0.0)
    sel = np.arange(N) if dim == 3 else kz * nv * cfg.nvy + np.arange(nv * cfg.nvy)
    orc = oracle_solver(cfg)
    for step, name in ((0, "f0.bin"), (2, "f2.bin")):
        rec = np.array(rows[step])
        assert zones[step] == ("zone i=  18     , j=  18     , k=  18" if dim == 3 else "zone i= 18, j= 18")
        assert rec.shape == (len(sel), dim + 4)
        assert np.all(rec[:, dim + 1] == step * cfg.dt)
        assert np.allclose(rec[:, dim], b3[sel], rtol=1e-15, atol=0)
        fall = np.fromfile(tmp_path / name).reshape(-1, N)
        f = fall[cell]
        assert np.allclose(rec[:, dim + 2], f[sel], rtol=1e-15, atol=1e-300)
        orc.upload_f(fall)
        p = orc.calculate_physical_properties()[cell]
        dens, u, w_, T = p[0], p[1:4], None, p[5]
        coeff = dens * np.sqrt(cfg.mass ** 3) / np.sqrt(pkg.config.PI_REF ** 3 * T ** 3)
        at = v[sel].copy()
        if dim == 2:
            at[:, 2] = 0.0
        feq = coeff * np.exp(-((at - u) ** 2).sum(1) * cfg.mass / T) * b3[sel]
        dev = rec[:, dim + 3]
        assert np.max(np.abs(dev - (f[sel] - feq))) <= 1e-12 * np.max(np.abs(f[sel]))
    orc.close()


def test_wall_heat_flux_files(pkg, tmp_path):
    """wall_heat_flux_dump_freq = 2 over 4 steps: each of the four wall files
    holds its header, then per dump step a zone line and one record per cell
    along the wall (time, cell-centre location, heat flux) equal to the
    oracle's fluxes from the same initial f (1e-12 of the half-space sums the
    flux is the difference of, m/2 sum |v_n| f |v|^2)."""
    over = dict(wall_heat_flux_dump_freq=2, nx_space="'5'", ny_space="'4'", dvm_samples=300)
    print("[deck] wall heat-flux files, 4 steps", flush=True)
    run_driver(str(tmp_path), deck_text(**over), "--steps", "4", "--dump-init", "f0.bin")
    files = {s: (tmp_path / f"{s}_wall.dat").read_text().splitlines() for s in ("left", "right", "bottom", "top")}
    cfg = solver_config(pkg, parse_check(run_driver(str(tmp_path), deck_text(**over), "--check").stdout))
    f0 = np.fromfile(tmp_path / "f0.bin").reshape(cfg.nx * cfg.ny, -1)
    orc = oracle_solver(cfg)
    orc.initialize(f0)
    orc.set_wall_heat_flux_freq(2)
    want = {}
    for n in range(1, 5):
        orc.step(n)
        if n % 2 == 0:
            want[n] = orc.wall_heat_flux()[0]
    orc.close()
    for side, (name, lines) in enumerate(files.items()):
        along = cfg.ny if side < 2 else cfg.nx
        assert lines[:2] == [f'title = "{name} wall fluxes"', 'variables = "Time", "Location", "Heat Flux"']
        body = lines[2:]
        assert len(body) == 2 * (1 + along)
        for z, n in enumerate((2, 4)):
            block = body[z * (1 + along):(z + 1) * (1 + along)]
            assert block[0] == f"zone i={along:4d}"
            rec = np.array([[float(x) for x in ln.split()] for ln in block[1:]])
            h = cfg.dy if side < 2 else cfg.dx
            assert np.allclose(rec[:, 0], n * cfg.dt, rtol=1e-13)
            assert np.allclose(rec[:, 1], (np.arange(along) + 0.5) * h, rtol=1e-13)
            q = want[n][side, :along]
            v = pkg.node_velocities(cfg)
            scale = 0.5 * cfg.mass * np.max(np.abs(f0) @ (np.abs(v[:, 0 if side < 2 else 1]) * (v ** 2).sum(1)))
            # records carry 15 significant digits
            assert np.max(np.abs(rec[:, 2] - q)) <= 1e-12 * scale + 1e-14 * np.max(np.abs(q))


def test_species_properties_file(pkg, tmp_path):
    """macro_properties_1.dat of a 2-step run with dumps every step: zones at
    t = 0, dt, 2 dt (written because nx > 1) of 12 cells with locations from
    0; directional temperatures (2/3) m sum (c_d - u_d)^2 f / n, the analytic
    relaxation columns (t = 0: T_rot = 0 and T_tr; later T_eq = T_tr and
    (4/3) E_tr) and the speed moments m^(k/2) sum |c - u|^k f / (n (k+1)!! T^(3/2+k)
    2^(-k/2)) equal the oracle's moments of the same f (1e-12)."""
    over = dict(nx_space="'4'", ny_space="'3'", dvm_samples=300, properties_vis_dump_freq=1, write_temp_x=1,
                write_temp_y=1, write_temp_z=1, write_rot_temp_relax=1, write_kin_temp_relax=1, write_moment=2,
                moment_numbers="'2 3'")
    print("[deck] per-species properties file, 2 steps", flush=True)
    run_driver(str(tmp_path), deck_text(**over), "--steps", "2", "--dump-init", "f0.bin", "--dump", "f2.bin")
    text = (tmp_path / "macro_properties_1.dat").read_text().splitlines()
    cfg = solver_config(pkg, parse_check(run_driver(str(tmp_path), deck_text(**over), "--check").stdout))
    records = []
    for ln in text[2:]:
        if ln.startswith("zone"):
            assert ln == "zone i=   12"
            records.append([])
        else:
            records[-1].append([float(x) for x in ln.split()])
    assert [len(r) for r in records] == [12, 12, 12]
    v = pkg.node_velocities(cfg)
    orc = oracle_solver(cfg)
    order = [ix + cfg.nx * iy for ix in range(cfg.nx) for iy in range(cfg.ny)]
    for step, name in ((0, "f0.bin"), (2, "f2.bin")):
        f = np.fromfile(tmp_path / name).reshape(-1, cfg.num_points)
        orc.upload_f(f)
        p = orc.calculate_physical_properties()
        got = np.array(records[step])
        assert np.allclose(got[:, 1], np.repeat(np.arange(cfg.nx) * cfg.dx, cfg.ny), rtol=1e-14, atol=1e-300)
        assert np.allclose(got[:, 2], np.tile(np.arange(cfg.ny) * cfg.dy, cfg.nx), rtol=1e-14, atol=1e-300)
        for row, c in zip(got, order):
            n, u, T = p[c, 0], p[c, 1:4], p[c, 4]
            cc = v - u
            tdir = (2.0 / 3.0) * cfg.mass * (f[c][:, None] * cc ** 2).sum(0) / n
            mom = []
            for k, nf in ((2, 3.0 * 1.0), (3, 4.0 * 2.0)):
                mb = nf * T ** (1.5 + k) * 2.0 ** (-0.5 * k)
                mom.append(cfg.mass ** (0.5 * k) * (np.sqrt((cc ** 2).sum(1)) ** k * f[c]).sum() / (n * mb))
            relax = [0.0, T] if step == 0 else [T, (4.0 / 3.0) * 0.75 * n * T]
            # columns: t x y | dens u v | T_tr | Tx Ty Tz | hf x y | rot relax, kin relax | p | moments
            want = np.concatenate([[n, u[0], u[1], T], tdir, [relax[0], relax[1], 0.5 * n * T], mom])
            sel = np.r_[3:7, 7:10, 12:15, 15:17]
            have = row[sel]
            assert np.allclose(have, want, rtol=1e-12, atol=1e-14 * np.max(np.abs(want))), (step, c, have, want)
    orc.close()
