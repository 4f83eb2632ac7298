"""1D runs on the device against the reference's own outputs.

tests/golden/run1d/<case>/ holds a case file and what the reference's
Simulation::run() wrote for it (tests/golden/make_run1d_golden.py):
scalars.csv (Diagnostics.cpp:28-72), boundary_losses.csv and
power_transfer.csv (Simulation.cpp:207-297) and fields_<step>.csv.  The host
test program runs the same case through pic::gpu::Simulation::run() (FP64
particles) and the files are compared here:

- identical headers (the reference's column contracts);
- step and time columns identical (time accumulates t += dt on both sides);
- wall counts, live counts and absorbed charge identical;
- energies, electric work and fields within 1e-6 relative (the device's
  fixed-point deposit and FFT-free Thomas solve round differently from the
  reference's FP64 sums, and a plasma amplifies rounding over 400 steps).

rf_1d1v_aps starts both sides from the same .aps state file; rf_1d3v_seeded
loads both from the config seed with std::mt19937_64."""
import csv
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(REPO, "aurora-pic_amd", "lib", "aurora_host_tests")
GOLDEN = os.path.join(REPO, "tests", "golden", "run1d")
CASES = ["rf_1d1v_aps", "rf_1d3v_seeded"]
EXACT = ("step", "time", "counter_origin_step", "live_particles", "phi_left", "phi_right")
RTOL = 1e-6


def _read(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def _exact_column(name):
    return (name in EXACT or name.startswith("live_particles_") or "_count_" in name
            or "_charge_" in name)


def _compare(mine, ref, name):
    hm, rm = _read(mine)
    hr, rr = _read(ref)
    assert hm == hr, f"{name}: header differs"
    assert len(rm) == len(rr), f"{name}: {len(rm)} rows, reference {len(rr)}"
    for j, col in enumerate(hr):
        a = np.array([float(r[j]) for r in rm])
        b = np.array([float(r[j]) for r in rr])
        if _exact_column(col):
            # phi at the electrodes is set, not computed: equal to the ulp
            tol = 1e-15 * max(np.max(np.abs(b)), 1e-300) if col.startswith("phi_") else 0.0
            assert np.all(np.abs(a - b) <= tol), (name, col, a, b)
        else:
            scale = max(np.max(np.abs(b)), 1e-300)
            assert np.max(np.abs(a - b)) <= RTOL * scale, (name, col, a, b)


@pytest.mark.parametrize("case", CASES)
def test_run1d_matches_reference_outputs(case, tmp_path):
    src = os.path.join(GOLDEN, case)
    out = str(tmp_path / case)
    r = subprocess.run([TESTS, "--run1d", os.path.join(src, "case.txt"), out],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    for f in ("scalars.csv", "boundary_losses.csv", "power_transfer.csv", "fields_0.csv",
              "fields_400.csv"):
        _compare(os.path.join(out, f), os.path.join(src, f), f"{case}/{f}")
