"""The drop-in DVMSolver / ScreenOutput modules under the reference's own
main program: oracle/_ref/dvmboltz_main is software/trunk/src/main.f90,
unchanged, compiled against this repository's Fortran modules (oracle/Makefile
target ref-main, built where the reference sources are present). It reads the
deck, writes reference_values.txt and configures the run exactly as the
repository's own driver does; on the GPU it runs the deck to the same moments."""
import os
import subprocess

import numpy as np
import pytest

from tests._deck import DRIVER, ROOT, deck_text, have_driver

MAIN = os.path.join(ROOT, "oracle", "_ref", "dvmboltz_main")

pytestmark = pytest.mark.skipif(not (os.path.exists(MAIN) and have_driver()),
                                reason="the reference's main program was not built against the drop-in modules")


def _run(exe, workdir, deck, *args):
    (workdir / "run.inp").write_text(deck)
    return subprocess.run([exe, "run.inp", *args], cwd=workdir, capture_output=True, text=True, check=True)


def test_reference_main_configures_like_the_driver(tmp_path):
    a, b = tmp_path / "main", tmp_path / "driver"
    a.mkdir()
    b.mkdir()
    ra = _run(MAIN, a, deck_text(), "--check")
    rb = _run(DRIVER, b, deck_text(), "--check")
    assert ra.stdout == rb.stdout and "species Ar" in ra.stdout
    assert (a / "reference_values.txt").read_text() == (b / "reference_values.txt").read_text()


@pytest.mark.gpu
def test_reference_main_runs_the_deck_on_the_gpu(tmp_path):
    deck = deck_text(nx_space="'4'", ny_space="'4'", num_points_x="'9'", num_points_y="'9'", num_points_z="'9'",
                     dvm_samples=300, verbosity="'2'", display_freq="'1'", enable_properties_vis="'0'")
    a, b = tmp_path / "main", tmp_path / "driver"
    a.mkdir()
    b.mkdir()
    print("[drop-in] the reference's main.f90 over DVMSolver, 4x4 cavity deck, 3 steps", flush=True)
    ra = _run(MAIN, a, deck, "--steps", "3", "--props", "p.txt")
    _run(DRIVER, b, deck, "--steps", "3", "--props", "p.txt")
    assert "Beginning Solve" in ra.stdout and "Finishing Solve" in ra.stdout
    pa, pb = np.loadtxt(a / "p.txt"), np.loadtxt(b / "p.txt")
    assert pa.shape == (16, 8)
    assert np.allclose(pa, pb, rtol=1e-12, atol=1e-15)
