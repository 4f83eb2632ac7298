"""Periodic plasmas on the device against the reference's own run output,
from the same seed (both sides load the species with std::mt19937_64):

- 3D3V with a tabulated B_y(x) (tests/golden/run3d,
  tests/golden/make_run3d_golden.py): 16^3 cells, 16,384 electrons and
  16,384 ions (mass 100), Boris push with the tabulated field, FFT Poisson,
  60 steps, against Simulation3D::run();
- 2D3V with a uniform B_z (tests/golden/run2d, make_run2d_golden.py): 64^2
  cells, the reference's 2D3V Boris push, 60 steps, against
  Simulation2D::run() (depth 0.5: the field energy and charge L1 take the
  node area times the out-of-plane depth, Diagnostics.cpp:127-135).

The host test program runs each case through pic::gpu::Simulation3D/2D::run(),
which write scalars.csv with the reference's columns; compared here:

- headers, step, time, live and absorbed counts identical;
- FP64 particles (the untiled push): kinetic, field and total energy and
  charge L1 within 1e-11 relative (measured: 4e-15, 2e-13, 3e-15, 4e-14);
- FP32 particles (the LDS-tiled push of the benchmark): kinetic and total
  energy within 1e-7, charge L1 and the field energy within 1e-6 relative
  (measured: 3e-9, 3e-9, 5e-8, 5e-8) -- FP32 positions and velocities
  against the reference's FP64 ones over 60 steps."""
import csv
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(REPO, "aurora-pic_amd", "lib", "aurora_host_tests")
CASE = os.path.join(REPO, "tests", "golden", "run3d", "plasma3d_btable")
EXACT = ("step", "time", "live_particles", "absorbed_left", "absorbed_right", "absorbed_bottom",
         "absorbed_top", "absorbed_back", "absorbed_front", "live_particles_electrons",
         "live_particles_ions")
CASE2 = os.path.join(REPO, "tests", "golden", "run2d", "plasma2d_bz")
TOL = {64: dict(kinetic_energy=1e-11, field_energy=1e-11, total_energy=1e-11, charge_l1=1e-11),
       32: dict(kinetic_energy=1e-7, field_energy=1e-6, total_energy=1e-7, charge_l1=1e-6)}


def _read(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], np.array([[float(v) for v in r] for r in rows[1:]])


@pytest.mark.parametrize("precision", [64, 32])
def test_run3d_matches_reference_scalars(precision, tmp_path):
    out = str(tmp_path / f"p{precision}")
    r = subprocess.run([TESTS, "--run3d", os.path.join(CASE, "case.txt"), out, str(precision)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    hm, mine = _read(os.path.join(out, "scalars.csv"))
    hr, ref = _read(os.path.join(CASE, "scalars.csv"))
    assert hm == hr
    assert mine.shape == ref.shape
    for j, col in enumerate(hr):
        if col in EXACT:
            assert np.array_equal(mine[:, j], ref[:, j]), (col, mine[:, j], ref[:, j])
        else:
            rel = np.max(np.abs(mine[:, j] - ref[:, j]) / np.abs(ref[:, j]))
            assert rel <= TOL[precision][col], (precision, col, rel, mine[:, j], ref[:, j])


@pytest.mark.parametrize("precision", [64, 32])
def test_run2d_matches_reference_scalars(precision, tmp_path):
    out = str(tmp_path / f"p{precision}")
    r = subprocess.run([TESTS, "--run2d", os.path.join(CASE2, "case.txt"), out, str(precision)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    hm, mine = _read(os.path.join(out, "scalars.csv"))
    hr, ref = _read(os.path.join(CASE2, "scalars.csv"))
    assert hm == hr
    assert mine.shape == ref.shape
    for j, col in enumerate(hr):
        if col in EXACT:
            assert np.array_equal(mine[:, j], ref[:, j]), (col, mine[:, j], ref[:, j])
        else:
            rel = np.max(np.abs(mine[:, j] - ref[:, j]) / np.abs(ref[:, j]))
            assert rel <= TOL[precision][col], (precision, col, rel, mine[:, j], ref[:, j])
