"""1D loss and power ledgers on the device (Simulation.cpp:1008-1138).

- The reference's wall-loss case (test_core.cpp:3290-3340): one neutral
  particle per species leaves through each wall in one step; counts, the
  represented impact energy 0.5 m w v^2 = 300, and zero electric work.
- The reference's power case (test_core.cpp:3470-3516): 32 cold electrons
  in a Dirichlet gap; after one step the electric work equals the change in
  kinetic energy to 1e-12, and is not zero.
- Ledgers, MCC energy tallies and wall counters are summed in a fixed order:
  two identical runs of an RF discharge with MCC and ionization agree to the
  bit, and a restart from a checkpoint continues the ledgers exactly.
- aurora_hip_initialize starts the ledgers over."""
import math

import numpy as np
import pytest

from aurora_pic import workloads

pytestmark = pytest.mark.gpu


def test_reference_wall_loss_case():
    from aurora_pic import hip
    c = hip.Context(0)
    c.mesh([3], [1.0], 1, 1.0)                      # Dirichlet, absorbing walls
    left = c.add_species(0.0, 2.0, 3.0, 4, precision=64)
    right = c.add_species(0.0, 2.0, 3.0, 4, precision=64)
    c.upload(left, np.array([0.5]), np.array([[-10.0, 0.0, 0.0]]))
    c.upload(right, np.array([0.5]), np.array([[10.0, 0.0, 0.0]]))
    c.initialize(0.2)
    c.step(0.2, 1)
    (cl, el), (cr, er) = c.losses(left), c.losses(right)
    assert list(cl[:2]) == [1, 0] and list(cr[:2]) == [0, 1]
    assert abs(el[0] - 300.0) <= 1e-12 and el[1] == 0.0
    assert abs(er[1] - 300.0) <= 1e-12 and er[0] == 0.0
    assert abs(c.power_transfer(left)) <= 1e-14 and abs(c.power_transfer(right)) <= 1e-14
    assert c.energy(0.2)[3] == 0
    c.close()


def test_reference_power_case_work_equals_energy_change():
    from aurora_pic import hip
    rng = np.random.default_rng(7)
    dt = 1e-4
    c = hip.Context(0)
    c.mesh([17], [1.0], 1, 1.0)
    sid = c.add_species(-1.0, 1.0, 1.0, 64, precision=64)
    c.upload(sid, rng.uniform(0.1, 0.9, 32), np.zeros((32, 3)))
    c.initialize(dt)
    ke0 = c.energy(dt)[0][0]
    c.step(dt, 1)
    ke1 = c.energy(dt)[0][0]
    work = c.power_transfer(sid)
    assert abs(work) > 1e-12
    assert abs(work - (ke1 - ke0)) <= 1e-12
    # and over many steps for survivors: still the kinetic-energy change
    c.step(dt, 50)
    ke2 = c.energy(dt)[0][0]
    assert c.losses(sid)[0][:2].sum() == 0
    assert abs(c.power_transfer(sid) - (ke2 - ke0)) <= 1e-9 * max(abs(ke2 - ke0), 1e-300)
    # initialize starts the ledger over
    c.initialize(dt)
    assert c.power_transfer(sid) == 0.0
    c.close()


def test_fp32_species_work_matches_fp64_bound():
    """FP32 storage: the work still tracks the kinetic-energy change of the
    survivors, within the FP32 rounding of the stored velocities."""
    from aurora_pic import hip
    rng = np.random.default_rng(8)
    dt = 1e-4
    c = hip.Context(0)
    c.mesh([33], [1.0], 1, 1.0)
    sid = c.add_species(-1.0, 1.0, 1.0, 4096)
    c.upload(sid, rng.uniform(0.2, 0.8, 2000), rng.normal(0.0, 0.05, (2000, 3)))
    c.initialize(dt)
    ke0 = c.energy(dt)[0][0]
    c.step(dt, 20)
    ke1 = c.energy(dt)[0][0]
    assert c.losses(sid)[0][:2].sum() == 0
    assert abs(c.power_transfer(sid) - (ke1 - ke0)) <= 1e-5 * ke0


def _discharge(seed=77):
    from aurora_pic import hip
    rng = np.random.default_rng(5)
    n, L, N = 65, 0.067, 3000
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    c = hip.Context(0)
    c.mesh([n], [L], 1, workloads.EPS0)
    c.electrodes([0, 0, 0, 0], [0.0, 200.0, 13.56e6, 0.0])
    sids = []
    w = 2.56e14 * L / N
    for q, m, te in ((-workloads.EV, workloads.ME, 8.0), (workloads.EV, 6.67e-27, 0.026)):
        sid = c.add_species(q, m, w, 4 * N)
        c.upload(sid, rng.uniform(0.1 * L, 0.9 * L, size=N),
                 rng.normal(0, math.sqrt(te * workloads.EV / m), size=(N, 3)))
        sids.append(sid)
    c.mcc_load(sids[0], gas, product_species=(sids[0], sids[1]), seed=seed)
    return c, sids


def _tallies(c, sids):
    out = [c.mcc_stats()]
    for s in sids:
        out.append(c.losses(s))
        out.append(c.power_transfer(s))
    return out


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, tuple):
            for u, v in zip(x, y):
                assert np.array_equal(u, v)
        else:
            assert x == y


def test_ledgers_are_run_to_run_identical_and_restart_exact(tmp_path):
    dt = 1.0 / (400 * 13.56e6)
    runs = []
    for _ in range(2):
        c, sids = _discharge()
        c.initialize(dt)
        c.step(dt, 400)
        runs.append(_tallies(c, sids))
        c.close()
    _same(runs[0], runs[1])
    stats, (ce, ee), we, (ci, ei), wi = runs[0]
    assert ce[:2].sum() > 0                         # electrons reach the walls
    assert np.any(stats[1] != 0.0)                  # channel energy tallies
    # the RF field does work on the electrons
    assert we != 0.0 and wi != 0.0
    path = str(tmp_path / "ledger.ckpt")
    c, sids = _discharge()
    c.initialize(dt)
    c.step(dt, 250)
    c.checkpoint_save(path)
    c.close()
    d, dsids = _discharge()
    d.checkpoint_load(path)
    d.step(dt, 150)
    _same(runs[0], _tallies(d, dsids))
    d.close()
