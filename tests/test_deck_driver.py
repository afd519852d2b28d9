"""The Fortran deck driver on the CPU (--check: no device work): a trunk deck
(tests/golden/decks/cavity10_ar.inp, the reference's own software/trunk/bin/
velocity_refined.inp) maps to the library's configuration, its scaling
reproduces the reference's reference_values.txt for that deck, and the
initial distribution is the renormalised Maxwellian of the deck's state."""
import os

import numpy as np
import pytest

from tests._deck import DECKS, deck_text, have_driver, parse_check, run_driver, solver_config

pytestmark = pytest.mark.skipif(not have_driver(), reason="dvm_gpu is not built (amdflang missing)")


def test_scaling_reproduces_the_reference_values_file(tmp_path):
    run_driver(str(tmp_path), deck_text(), "--check")
    got = (tmp_path / "reference_values.txt").read_text()
    want = open(os.path.join(DECKS, "cavity10_ar.reference_values.txt")).read()
    assert got == want


def test_cavity_deck_maps_to_the_library_config(tmp_path):
    c = parse_check(run_driver(str(tmp_path), deck_text(), "--check").stdout)
    assert (c["dim"], c["nx"], c["ny"], c["nvx"], c["nvy"], c["nvz"]) == (2, 10, 10, 18, 18, 18)
    assert c["vx0"] == -3.25 and c["vx1"] == 3.25 and c["dt"] == 0.02 and c["kn"] == 1.091
    # the reference species scales to unit mass and diameter; VHS keeps omega, sets alpha = 1
    assert c["mass"] == 1.0 and c["diam"] == 1.0 and c["omega"] == 0.81 and c["alpha"] == 1.0
    assert c["csection_coeff"] == 2 and c["coll_solver"] == 3 and c["num_repl"] == 1
    assert c["sample_mode"] == 0 and c["coln_rms"] == 0.001 and c["cdf_tolerance"] == 0.01
    assert c["enable_convection"] == 1 and c["bc_left"] == 1 and c["bc_right"] == 1 and c["acc_coeff"] == 1.0
    assert c["wall3"] == [1.0, 0.02965, 0.0, 0.0, 1.0] and c["wall0"] == [1.0, 0.0, 0.0, 0.0, 1.0]
    assert c["ntstep"] == 10000 and c["species"] == "Ar"


@pytest.mark.parametrize("method, omega, alpha", [(1, 0.5, 1.0), (2, 1.0, 1.0), (3, 0.81, 1.0), (4, 0.81, 1.40)])
def test_cross_section_model_sets_omega_and_alpha(tmp_path, method, omega, alpha):
    c = parse_check(run_driver(str(tmp_path), deck_text(csection_method=method), "--check").stdout)
    assert c["omega"] == omega and c["alpha"] == alpha


def test_other_species_scale_by_the_reference_species(tmp_path):
    c = parse_check(run_driver(str(tmp_path), deck_text(species_codes="'He'"), "--check").stdout)
    assert c["mass"] == pytest.approx(np.float32(6.65e-27) / np.float32(66.3e-27), rel=1e-15)
    assert c["diam"] == pytest.approx(np.float32(2.33e-10) / np.float32(4.17e-10), rel=1e-15)
    assert c["omega"] == 0.66


def test_extension_keys(tmp_path):
    c = parse_check(run_driver(str(tmp_path), deck_text(dvm_samples=500, dvm_precision=32, nx_space="'12'",
                                                         ny_space="'1'", dvm_dim=1, dvm_conv_order=2,
                                                         LWallBC="'2'", RWallBC="'3'"), "--check").stdout)
    assert c["sample_mode"] == 1 and c["samples_per_cell"] == 500 and c["precision"] == 1
    assert c["dim"] == 1 and c["nx"] == 12 and c["conv_order"] == 2 and (c["bc_left"], c["bc_right"]) == (2, 3)


@pytest.mark.parametrize("key, value, text", [("iColl", 1, "iColl"), ("csection_method", 7, "csection_method"),
                                              ("species_codes", "'H2'", "monatomic"), ("InitDF", "'5'", "InitDF"),
                                              ("x_vel_grid_type", "'0'", "velocity-grid file")])
def test_unsupported_decks_stop_with_a_message(tmp_path, key, value, text):
    r = run_driver(str(tmp_path), deck_text(**{key: value}), "--check", check=False)
    assert r.returncode != 0 and text in r.stdout


@pytest.mark.parametrize("initc", [5, 3])
def test_initial_distribution_is_the_deck_state(pkg, tmp_path, initc):
    over = dict(InitC=f"'{initc}'", dens_LW="'0.8d0'", x_vel_LW="'0.3d0'", tr_temp_LW="'1.4d0'",
                dens_RW="'1.2d0'", tr_temp_RW="'0.9d0'", spatial_middle="'4'", nx_space="'6'", ny_space="'2'")
    r = run_driver(str(tmp_path), deck_text(**over), "--check", "--dump-init", "f0.bin")
    cfg = solver_config(pkg, parse_check(r.stdout))
    f0 = np.fromfile(tmp_path / "f0.bin", dtype=np.float64).reshape(cfg.nx * cfg.ny, cfg.num_points)
    for cell in range(cfg.nx * cfg.ny):
        ix = cell % cfg.nx
        if initc == 5:
            state = (1.0, 0.0, 0.0, 0.0, 1.0)            # dens_init ... tr_temp_init of the deck
        elif ix + 1 > 4:
            state = (1.2, 0.0, 0.0, 0.0, 0.9)
        else:
            state = (0.8, 0.3, 0.0, 0.0, 1.4)
        want = pkg.maxwellian(cfg, *state, renormalise=True)
        assert np.allclose(f0[cell], want, rtol=1e-13, atol=0), cell


def test_properties_file_header_follows_the_reference_layout(tmp_path):
    run_driver(str(tmp_path), deck_text(), "--check")
    lines = (tmp_path / "macro_properties.dat").read_text().splitlines()
    # (a3,a14) on ", " and a blank-padded name: " , " then the name's leftmost 14 characters;
    # the closing (1x) only moves the position, so nothing follows the last name
    cols = "".join(" , " + n.ljust(14)[:14] for n in ['"Density"', '"X-Velocity"', '"Y-Velocity"',
                                                       '"Trans. Temp."', '"Heat Flux X"', '"Heat Flux Y"',
                                                       '"Pressure"'])
    assert lines == ['title = "Macroscopic Properties"',
                     'variables = "Time", "x-Location", "y-Location"' + cols]


def test_full_monte_carlo_deck(tmp_path):
    c = parse_check(run_driver(str(tmp_path), deck_text(iColl=2), "--check").stdout)
    assert c["coll_solver"] == 2


def fortran_records(*payloads):
    """Unformatted sequential file: each record framed by its 4-byte length."""
    import struct
    out = b""
    for p in payloads:
        out += struct.pack("<i", len(p)) + p + struct.pack("<i", len(p))
    return out


def read_fortran_records(path):
    import struct
    data, recs, i = open(path, "rb").read(), [], 0
    while i < len(data):
        (n,) = struct.unpack_from("<i", data, i)
        recs.append(data[i + 4:i + 4 + n])
        assert struct.unpack_from("<i", data, i + 4 + n)[0] == n
        i += n + 8
    return recs


def test_restart_file_is_read_in_record_order(pkg, tmp_path):
    """read_rst = 1: ntime from the first record, then one record of f per
    cell (restart.f90:147-250 order); the run would start at ntime + 1."""
    over = dict(nx_space="'3'", ny_space="'2'", read_rst=1, read_restart_filename="r.rst")
    c = parse_check(run_driver(str(tmp_path), deck_text(**{k: v for k, v in over.items() if "rst" not in k}),
                               "--check").stdout)
    N, ncell = c["nvx"] * c["nvy"] * c["nvz"], 6
    f = np.random.default_rng(2).random((ncell, N))
    (tmp_path / "r.rst").write_bytes(fortran_records(np.int32(7).tobytes(), *(row.tobytes() for row in f)))
    r = run_driver(str(tmp_path), deck_text(**over), "--check", "--dump-init", "f0.bin")
    assert parse_check(r.stdout)["ntime_init"] == 7
    assert np.array_equal(np.fromfile(tmp_path / "f0.bin").reshape(ncell, N), f)
    # a file shorter than the deck's grid stops with a message
    (tmp_path / "r.rst").write_bytes(fortran_records(np.int32(7).tobytes(), f[0].tobytes()))
    r = run_driver(str(tmp_path), deck_text(**over), "--check", check=False)
    assert r.returncode != 0 and "shorter" in r.stdout


def test_distribution_function_headers(tmp_path):
    """enable_vel_df_vis: one file per output location, named
    <df_vis_filename>_<species>_<x>_<y>.dat, with the header of
    visualization.f90:497-560 for the chosen dimension."""
    over = dict(enable_vel_df_vis=1, df_output_dimension=2, num_output_locations=2, x_output_locations="'3 5'",
                y_output_locations="'2 4'", df_vis_filename="dfout")
    run_driver(str(tmp_path), deck_text(**over), "--check")
    for name in ("dfout_1_3_2.dat", "dfout_1_5_4.dat"):
        assert (tmp_path / name).read_text().splitlines() == [
            'title = "Distribution Function"',
            'variables = "X-Velocity", "Y-Velocity", "Beta3", "Time", "Vel DistFunc", "Deviation"']
    over.update(df_output_dimension=3)
    run_driver(str(tmp_path), deck_text(**over), "--check")
    assert (tmp_path / "dfout_1_3_2.dat").read_text().splitlines()[1] == (
        'variables = "X-Velocity", "Y-Velocity", "Z-Velocity", "Beta3", "Time", "Vel DistFunc", "Deviation"')


@pytest.mark.parametrize("over, text", [
    (dict(enable_vel_df_vis=1, df_output_dimension=1), "df_output_dimension 1"),      # even y grid: no y = 0 node
    (dict(enable_rot_df_vis=1), "internal energy levels")])
def test_distribution_function_output_errors(tmp_path, over, text):
    r = run_driver(str(tmp_path), deck_text(**over), "--check", check=False)
    assert r.returncode != 0 and text in r.stdout


def test_distribution_function_location_outside_the_grid_is_skipped(tmp_path):
    r = run_driver(str(tmp_path), deck_text(num_output_locations=2, x_output_locations="'11 2'",
                                            y_output_locations="'3 3'"), "--check")
    assert "(11,3) is outside the physical grid" in r.stdout
    assert not (tmp_path / "dist_func_1_11_3.dat").exists() and (tmp_path / "dist_func_1_2_3.dat").exists()


def test_wall_files_have_headers_on_2d_grids(tmp_path):
    run_driver(str(tmp_path), deck_text(), "--check")
    for s in ("left", "right", "bottom", "top"):
        assert (tmp_path / f"{s}_wall.dat").read_text().splitlines() == [
            f'title = "{s} wall fluxes"', 'variables = "Time", "Location", "Heat Flux"']


def test_no_wall_files_on_1d_grids(tmp_path):
    run_driver(str(tmp_path), deck_text(nx_space="'6'", ny_space="'1'", dvm_dim=1), "--check")
    assert not (tmp_path / "left_wall.dat").exists()


def test_species_properties_file_header(tmp_path):
    """<properties_vis_filename>_1.dat: the total file's header with the names
    trimmed (visualization.f90:660-700)."""
    run_driver(str(tmp_path), deck_text(write_moment=2, moment_numbers="'3 4'"), "--check")
    lines = (tmp_path / "macro_properties_1.dat").read_text().splitlines()
    names = ['"Density"', '"X-Velocity"', '"Y-Velocity"', '"Trans. Temp."', '"Heat Flux X"', '"Heat Flux Y"',
             '"Pressure"']
    assert lines == ['title = "Macroscopic Properties"',
                     'variables = "Time", "x-Location", "y-Location"' + "".join(", " + n for n in names)
                     + ', "Moment  3", "Moment  4"']


def _f0(pkg, tmp_path, r, grid=None):
    cfg = solver_config(pkg, parse_check(r.stdout))
    if grid is not None:
        cfg = cfg.with_(vx=grid[0], vy=grid[1], vz=grid[2])
    return cfg, np.fromfile(tmp_path / "f0.bin", dtype=np.float64).reshape(cfg.nx * cfg.ny, cfg.num_points)


def _moments(pkg, cfg, f):
    v = pkg.node_velocities(cfg)
    n = f.sum()
    u = f @ v / n
    return n, u, 2.0 * cfg.mass * (f * ((v - u) ** 2).sum(1)).sum() / (3.0 * n)


SHOCK = dict(nx_space="'24'", ny_space="'1'", dvm_dim=1, LWallBC="'2'", RWallBC="'2'", InitC="'4'",
             spatial_middle="'12'", shock_thickness="'6'", dens_RW="'1.0d0'", x_vel_RW="'-1.6d0'",
             tr_temp_RW="'1.0d0'", dens_LW="'1.0d0'", x_vel_LW="'0.0d0'", tr_temp_LW="'1.0d0'",
             xv_min="'-6.0d0'", xv_max="'6.0d0'", num_points_x="'17'", num_points_y="'9'", num_points_z="'9'",
             yv_min="'-4.0d0'", yv_max="'4.0d0'", zv_min="'-4.0d0'", zv_max="'4.0d0'")


def test_initc4_stationary_shock_jump_and_profile(pkg, tmp_path):
    """InitC = 4 with in/out-flow sides: a stationary shock (shock_conditions.
    f90:54-171) whose downstream state goes to the left wall. Pinned by the
    Rankine-Hugoniot ratios (checked against the reference's shock-calculator
    fixture) at Mach |u_up| / sqrt(gamma T / 2m), gamma = 5/3, and by mass-flux
    conservation across the shock; the columns between follow
    initial_shock_profile (initial_conditions.f90:652-671) around
    spatial_middle, each cell the renormalised Maxwellian of its state."""
    r = run_driver(str(tmp_path), deck_text(**SHOCK), "--check", "--dump-init", "f0.bin")
    c = parse_check(r.stdout)
    mach = 1.6 / pkg.sound_speed(5.0 / 3.0, 1.0, 1.0)
    dens, vel, temp, _ = pkg.rankine_hugoniot(mach, 5.0 / 3.0)
    lw = c["wall0"]
    assert abs(lw[0] - dens) < 1e-13 and abs(lw[4] - temp) < 1e-13
    assert abs(lw[0] * lw[1] - (-1.6)) < 1e-12          # n u is the same on both sides
    assert abs(lw[1] - (-1.6) * vel) < 1e-12
    cfg, f0 = _f0(pkg, tmp_path, r)
    half, mid = 3, 12
    tmid = int(np.ceil(mid + 0.5 * half))
    for ix in range(1, 25):
        n_, u_, t_ = _moments(pkg, cfg, f0[ix - 1])
        if ix < mid - half:
            want_n, want_t = lw[0], lw[4]
        elif ix > mid + half and ix > tmid + half:
            want_n, want_t = 1.0, 1.0
        else:
            continue
        assert abs(n_ - want_n) < 1e-12 and abs(t_ - want_t) < 2e-3 * want_t, (ix, n_, t_)
    # the density profile: LW + (RW - LW)/2 (x - mid + half)^2 / half^2 left of the
    # middle, RW + (LW - RW)/2 (mid - x + half)^2 / half^2 right of it
    for ix in range(mid - half, mid + half + 1):
        x = (ix - mid + half) if ix <= mid else (mid - ix + half)
        want = (lw[0] + 0.5 * (1.0 - lw[0]) * x * x / half ** 2) if ix <= mid else \
            (1.0 + 0.5 * (lw[0] - 1.0) * x * x / half ** 2)
        assert abs(f0[ix - 1].sum() - want) < 1e-12, ix


def test_initc4_reflected_shock_leaves_the_gas_at_rest(tmp_path):
    """With a reflecting wall the shock moves (the reference solves a quadratic
    for its speed); behind a shock reflected from a wall the gas is at rest."""
    over = dict(SHOCK, LWallBC="'1'")
    c = parse_check(run_driver(str(tmp_path), deck_text(**over), "--check").stdout)
    assert abs(c["wall0"][1]) < 1e-12 and c["wall0"][0] > 1.0 and c["wall0"][4] > 1.0


@pytest.mark.parametrize("initdf", [2, 3, 4])
def test_initdf_shapes(pkg, tmp_path, initdf):
    """InitDF 2: BKW at t = 0 (dist_func.f90:374-401): every BKW profile has
    the Maxwellian's temperature; 3: all of the density on the node of
    x/y/z_delta_loc; 4: the Maxwellian integrated over each node's cell
    (dist_func.f90:346-372), which approaches the nodal Maxwellian."""
    over = dict(InitC="'5'", InitDF=f"'{initdf}'", nx_space="'2'", ny_space="'2'", dens_init="'1.3d0'",
                x_vel_init="'0.25d0'", tr_temp_init="'1.2d0'", num_points_x="'29'", num_points_y="'29'",
                num_points_z="'29'", xv_min="'-6.0d0'", xv_max="'6.0d0'", yv_min="'-6.0d0'", yv_max="'6.0d0'",
                zv_min="'-6.0d0'", zv_max="'6.0d0'", x_delta_loc="'0.0d0'", y_delta_loc="'1.28571428571428570d0'",
                z_delta_loc="'0.0d0'")
    r = run_driver(str(tmp_path), deck_text(**over), "--check", "--dump-init", "f0.bin")
    cfg, f0 = _f0(pkg, tmp_path, r)
    n, u, t = _moments(pkg, cfg, f0[0])
    assert abs(n - 1.3) < 1e-13
    if initdf == 2:
        assert abs(t - 1.2) < 1e-6 and abs(u[0] - 0.25) < 1e-9
        m = pkg.maxwellian(cfg, 1.3, 0.25, 0.0, 0.0, 1.2)
        assert np.abs(f0[0] - m).max() > 1e-2 * m.max()  # not the Maxwellian
    elif initdf == 3:
        v = pkg.node_velocities(cfg)
        nz = np.nonzero(f0[0])[0]
        assert len(nz) == 1 and abs(f0[0][nz[0]] - 1.3) < 1e-15
        assert np.allclose(v[nz[0]], [0.0, 9.0 / 7.0, 0.0], atol=1e-12)
    else:
        m = pkg.maxwellian(cfg, 1.3, 0.25, 0.0, 0.0, 1.2)
        assert np.abs(f0[0] - m).max() < 0.05 * m.max()
        assert abs(t - 1.2) < 0.05


def test_velocity_grid_file(pkg, tmp_path):
    """x/y/z_vel_grid_type = 0 takes the axes from the velocity-grid file of the
    command line (read_vel_grid: the three point counts, then the values)."""
    gx = np.array([-4.0, -2.5, -1.5, -0.75, 0.0, 0.75, 1.5, 2.5, 4.0])
    gy = np.array([-3.0, -1.0, 0.0, 1.0, 3.0])
    gz = np.array([-3.5, -2.0, -0.5, 0.5, 2.0, 3.5, 4.0])
    with open(tmp_path / "grid.dat", "w") as fh:
        fh.write(f"{len(gx)}\n{len(gy)}\n{len(gz)}\n")
        for a in (gx, gy, gz):
            fh.writelines(f"{x:.17g}\n" for x in a)
    over = dict(x_vel_grid_type="'0'", y_vel_grid_type="'0'", z_vel_grid_type="'0'", InitC="'5'",
                x_vel_init="'0.3d0'", nx_space="'2'", ny_space="'1'")
    path = tmp_path / "run.inp"
    path.write_text(deck_text(**over))
    import subprocess
    from tests._deck import DRIVER
    r = subprocess.run([DRIVER, str(path), str(tmp_path / "grid.dat"), "--check", "--dump-init", "f0.bin"],
                       cwd=tmp_path, capture_output=True, text=True, check=True)
    c = parse_check(r.stdout)
    assert (c["nvx"], c["nvy"], c["nvz"]) == (9, 5, 7) and c["vx0"] == -4.0 and c["vz1"] == 4.0
    cfg, f0 = _f0(pkg, tmp_path, r, grid=(gx, gy, gz))
    want = pkg.maxwellian(cfg, 1.0, 0.3, 0.0, 0.0, 1.0, pi=3.1459265358979323)
    assert np.allclose(f0[0], want, rtol=1e-13, atol=0)
