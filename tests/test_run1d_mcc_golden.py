"""The reference's own collisional 1D3V run (tests/golden/run1d/rf_1d3v_mcc)
pins the invariants the device run is held to: its product count (live
particles = initial + ionizations - absorbed, per species) and its energy
ledger, which closes to rounding (kinetic-energy change = electric work -
wall energy + elastic and excitation energy changes - ionizations x
threshold x weight: the primary and its secondary share the energy above
the threshold).  The per-seed statistics file holds 12 seeds of the same
case."""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import make_run1d_mcc_golden as golden_case  # noqa: E402

D = golden_case.OUT


def test_reference_products_and_ledger_close():
    hc, c = golden_case.read_csv(os.path.join(D, "collisions.csv"))
    hs, s = golden_case.read_csv(os.path.join(D, "scalars.csv"))
    hb, b = golden_case.read_csv(os.path.join(D, "boundary_losses.csv"))
    hp, p = golden_case.read_csv(os.path.join(D, "power_transfer.csv"))
    nion = c[:, hc.index("cumulative_collisions_e_he.ionization")]
    assert nion[-1] > 1500
    for sp in ("electrons", "ions"):
        absorbed = b[:, hb.index(f"absorbed_left_count_{sp}")] + \
            b[:, hb.index(f"absorbed_right_count_{sp}")]
        assert np.array_equal(s[:, hs.index(f"live_particles_{sp}")],
                              golden_case.N + nion - absorbed)
    w = 2.56e14 * golden_case.L / golden_case.N
    ke = s[:, hs.index("kinetic_energy")]
    work = p[:, [j for j, h in enumerate(hp) if h.startswith("electric_work")]].sum(axis=1)
    wall = b[:, [j for j, h in enumerate(hb) if "kinetic_energy" in h]].sum(axis=1)
    el = c[:, hc.index("cumulative_tracked_kinetic_energy_change_e_he.elastic_J_m-2")]
    ex = c[:, hc.index("cumulative_tracked_kinetic_energy_change_e_he.excitation_J_m-2")]
    resid = (ke - ke[0]) - (work - wall + el + ex - 24.6 * golden_case.wl.EV * w * nion)
    assert np.max(np.abs(resid)) <= 1e-12 * np.max(np.abs(work))


def test_reference_seed_statistics_are_consistent():
    g = np.load(os.path.join(D, "seeds.npz"), allow_pickle=False)
    assert g["windows"].shape == (12, 10, 5)
    # the EEDF fractions of every run sum to one; counts per window are
    # Poisson-like across seeds (spread within a factor 3 of sqrt(mean))
    assert np.allclose(g["eedf"].sum(axis=1), 1.0)
    ion = g["windows"][:, :, 4].sum(axis=1)
    assert ion.std(ddof=1) < 3 * np.sqrt(ion.mean())
