"""The reference's own collisional 1D runs (tests/golden/run1d/rf_1d3v_mcc,
rf_1d3v_mcc_ions, rf_1d1v_mcc) pin the invariants the device runs are held
to: the product count (live particles = initial + ionizations - absorbed,
per species) and the energy ledger, which the reference closes to rounding:
kinetic-energy change = electric work - wall energy + the tracked energy
change of every channel (an ionization's includes its secondary's and ion's
kinetic energy, Simulation.cpp:880-902).  The per-seed statistics files
hold 12 seeds of each case."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import make_run1d_mcc_golden as golden_case  # noqa: E402


@pytest.mark.parametrize("case", golden_case.CASES)
def test_reference_products_and_ledger_close(case):
    resid, work, products_ok, nion = golden_case.ledger_residual(golden_case.case_dir(case))
    assert products_ok
    assert np.max(np.abs(resid)) <= 1e-12 * work
    if case != "rf_1d1v_mcc":
        assert nion[-1] > 1500


@pytest.mark.parametrize("case", golden_case.CASES)
def test_reference_seed_statistics_are_consistent(case):
    g = np.load(os.path.join(golden_case.case_dir(case), "seeds.npz"), allow_pickle=False)
    nch = {"rf_1d3v_mcc": 3, "rf_1d3v_mcc_ions": 5, "rf_1d1v_mcc": 2}[case]
    assert g["windows"].shape == (12, 10, 2 + nch)
    assert np.allclose(g["eedf"].sum(axis=1), 1.0)
    # every channel fires; counts spread across seeds like a Poisson variable
    # or wider (the discharge's own fluctuation: 1-6x Poisson in these runs,
    # the widest for the most frequent channel), never far narrower
    total = g["windows"][:, :, 2:].sum(axis=1)
    assert np.all(total.mean(axis=0) > 50)
    ratio = total.std(axis=0, ddof=1) / np.sqrt(total.mean(axis=0))
    assert np.all(ratio > 0.25) and np.all(ratio < 10)
