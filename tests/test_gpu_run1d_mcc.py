"""Running 1D discharges with MCC against the reference's own runs
(tests/golden/run1d/<case>, written by tests/golden/make_run1d_mcc_golden.py
from the reference's Simulation::run() over 12 seeds):

  rf_1d3v_mcc        electron impact with ionization products
  rf_1d3v_mcc_ions   two named models: electrons, then ions on 300 K neutrals
  rf_1d1v_mcc        1D1V, the scalar collide

The device (pic::gpu::Simulation::run() through the host test program,
FP32 particles, Philox streams) runs each case over 6 seeds; both sides
load their initial particles from the seed the reference's way, but their
collision random numbers differ, so the bar is statistical: every compared
quantity's device mean lies within 5 standard errors of the reference's
mean, the standard error combining both sides' seed-to-seed spread (for
counts at least the Poisson variance).  Compared: per-window collision
counts by channel and their energy changes (collisions.csv), kinetic and
field energy and live counts at every output (scalars.csv), the wall
ledgers and electric work at the end, the mean electron energy and the
electron energy distribution (fraction per log-energy bin, each bin above
1 % and a chi-square over all of them).

Exact on every device run: the column contracts of the four CSV files, the
product count (live particles = initial + ionizations - absorbed, per
species) and the energy ledger, which the reference closes to rounding:
kinetic-energy change = electric work - wall energy + every channel's
tracked energy change (within 1e-6 of the work: FP32 velocities)."""
import csv
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import make_run1d_mcc_golden as golden_case  # noqa: E402

pytestmark = pytest.mark.gpu

TESTS = os.path.join(REPO, "aurora-pic_amd", "lib", "aurora_host_tests")
DEVICE_SEEDS = 6
_RUNS = {}


def _device_runs(case, tmp_path_factory):
    if case not in _RUNS:
        tmp = tmp_path_factory.mktemp(case)
        outs = []
        for k in range(DEVICE_SEEDS):
            seed = golden_case.SEED + k
            out = str(tmp / f"seed{seed}")
            print(f"device run {case} seed {seed}", flush=True)
            r = subprocess.run([TESTS, "--run1d",
                                os.path.join(golden_case.case_dir(case), "case.txt"), out],
                               capture_output=True, text=True, timeout=300,
                               env=dict(os.environ, AURORA_CASE_SEED=str(seed)))
            assert r.returncode == 0, r.stdout + r.stderr
            outs.append(out)
        _RUNS[case] = outs
    return _RUNS[case]


def _header(path):
    with open(path) as f:
        return next(csv.reader(f))


def _within(name, dev, ref, counts=False, sigmas=5.0):
    dev, ref = np.asarray(dev, float), np.asarray(ref, float)
    md, mr = dev.mean(axis=0), ref.mean(axis=0)
    vd, vr = dev.var(axis=0, ddof=1), ref.var(axis=0, ddof=1)
    if counts:
        vd, vr = np.maximum(vd, md), np.maximum(vr, mr)
    se = np.sqrt(vd / len(dev) + vr / len(ref))
    z = np.abs(md - mr) / np.where(se > 0, se, np.inf)
    assert np.all((z <= sigmas) | (np.abs(md - mr) <= 1e-12 * np.abs(mr))), \
        (name, np.max(z), md, mr, se)
    return z


@pytest.mark.parametrize("case", golden_case.CASES)
def test_column_contracts(case, tmp_path_factory):
    runs = _device_runs(case, tmp_path_factory)
    for f in ("scalars.csv", "collisions.csv", "boundary_losses.csv", "power_transfer.csv"):
        assert _header(os.path.join(runs[0], f)) == \
            _header(os.path.join(golden_case.case_dir(case), f)), f


@pytest.mark.parametrize("case", golden_case.CASES)
def test_statistics_match_the_reference(case, tmp_path_factory):
    runs = _device_runs(case, tmp_path_factory)
    ref = np.load(os.path.join(golden_case.case_dir(case), "seeds.npz"), allow_pickle=False)
    dev = [golden_case.summarise(d) for d in runs]
    stack = {k: np.stack([r[k] for r in dev]) for k in dev[0]}
    assert stack["windows"].shape[1:] == ref["windows"].shape[1:]
    zs = {}
    zs["windows"] = _within("collision counts", stack["windows"], ref["windows"], counts=True)
    zs["energy"] = _within("window energy change", stack["window_energy"], ref["window_energy"])
    zs["scalars"] = _within("scalars", stack["scalars"], ref["scalars"])
    zs["walls"] = _within("wall ledgers", stack["walls"], ref["walls"])
    zs["work"] = _within("electric work", stack["work"], ref["work"])
    zs["mean_energy"] = _within("mean electron energy", stack["mean_energy_ev"],
                                ref["mean_energy_ev"])
    f_ref, f_dev = ref["eedf"], stack["eedf"]
    keep = f_ref.mean(axis=0) > 0.01
    z = _within("EEDF bins", f_dev[:, keep], f_ref[:, keep])
    chi2 = float(np.mean(z ** 2))
    print(case, {k: round(float(np.max(v)), 3) for k, v in zs.items()},
          "EEDF bins", np.round(z, 2).tolist(), "chi2/dof", round(chi2, 3), flush=True)
    assert chi2 < 4.0
    assert stack["windows"][:, :, 2:].sum(axis=1).min() > 50     # every channel fired


@pytest.mark.parametrize("case", golden_case.CASES)
def test_products_and_energy_ledger_close(case, tmp_path_factory):
    for d in _device_runs(case, tmp_path_factory):
        resid, work, products_ok, _ = golden_case.ledger_residual(d)
        assert products_ok
        assert np.max(np.abs(resid)) <= 1e-6 * work, (resid, work)


def test_full_size_discharge_invariants():
    """The 1D3V discharge at its full size (513 nodes, 5e5 electrons + 5e5
    ions, FP32, synthetic gas at 1e21 m^-3) for 200 steps through the C ABI:
    products equal ionizations (live = initial + ionizations - absorbed, per
    species), charge is conserved with the wall losses, and the energy
    ledger closes within 1e-6 of the electric work."""
    from aurora_pic import hip, workloads as wl
    cfg = wl.config_discharge()
    steps = 200
    with hip.Context(0) as c:
        c.mesh(cfg["n"], cfg["L"], hip.DIRICHLET, cfg["eps"])
        c.electrodes([0.0, 0.0, 0.0, 0.0], [0.0, cfg["phi_right_amplitude"], cfg["frequency"], 0.0])
        sids = []
        for k, s in enumerate(cfg["species"]):
            sid = c.add_species(s["charge"], s["mass"], s["weight"], 3 * s["particles"])
            c.init_maxwellian(sid, s["particles"], wl.maxwellian_sigma(s["temperature_ev"],
                                                                       s["mass"]), 100 + k)
            sids.append(sid)
        c.mcc_load(sids[0], cfg["gas"], product_species=(sids[0], sids[1]), seed=cfg["seed"])
        c.initialize(cfg["dt"])
        ke0, _, _, live0 = c.energy(cfg["dt"])
        c.step(cfg["dt"], steps)
        ke1, _, _, live1 = c.energy(cfg["dt"])
        counts, ech = c.mcc_stats()
        nion = int(counts[4])
        assert nion > 100
        absorbed, wall, work = {}, 0.0, 0.0
        for sid in sids:
            a, e = c.losses(sid)
            absorbed[sid] = int(a[:2].sum())
            wall += float(e.sum())
            work += c.power_transfer(sid)
            _, live = c.count(sid)
            assert live == cfg["species"][sid]["particles"] + nion - absorbed[sid], sid
        # charge: every ionization adds +q and -q; the walls remove what they absorb
        w, q = cfg["species"][0]["weight"], wl.EV
        charge = q * w * ((live1 - live0) - 2 * nion + absorbed[0] + absorbed[1])
        assert charge == 0.0
        resid = (ke1.sum() - ke0.sum()) - (work - wall + ech.sum())
        print(f"ionizations {nion}, work {work:.4e}, wall {wall:.4e}, residual {resid:.3e}",
              flush=True)
        assert abs(resid) <= 1e-6 * abs(work)
