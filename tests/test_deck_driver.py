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
                                              ("species_codes", "'H2'", "monatomic"), ("InitC", "'4'", "InitC")])
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
