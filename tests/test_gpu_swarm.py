"""GPU batched swarms: closed form, parity with the oracle (same Philox
streams) and the reference (statistical), and identical results alone and
in a mix (continuous batching)."""
import math

import numpy as np
import pytest

import oracle_py
from aurora_pic import workloads
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx():
    from aurora_pic import hip
    c = hip.Context(0)
    yield c
    c.close()


def _request(td, particles, steps, burn, blocks, dt, seed, e0=1.0, emax=900.0):
    from aurora_pic import hip
    return hip.SwarmRequest(reduced_field_td=td, timestep=dt, particles=particles, steps=steps,
                            burn_in=burn, blocks=blocks, initial_mean_energy_ev=e0,
                            max_energy_ev=emax, seed=seed)


def _run_all(ctx, requests, chunk):
    handles = [ctx.swarm_submit(r) for r in requests]
    results = {}
    while len(results) < len(handles):
        ctx.swarm_advance(chunk)
        for h in handles:
            if h in results:
                continue
            done, r = ctx.swarm_poll(h)
            if done:
                results[h] = r
    return [results[h] for h in handles]


def test_swarm_zero_cross_section_closed_form(ctx):
    e = np.array([0.0, 1.0e-16])
    gas = dict(channels=[(workloads.ELASTIC, 0.0, e, np.zeros(2))], neutral_density=1e20,
               max_frequency=1e7, neutral_mass=workloads.HELIUM_MASS)
    ctx.swarm_gas(gas)
    r = _run_all(ctx, [_request(1.0, 300, 40, 10, 5, 1e-9, 3, e0=0.0, emax=600.0)], 64)[0]
    a_dt = workloads.EV * 1.0e-21 * 1e20 / workloads.ME * 1e-9
    s1 = np.arange(11, 41, dtype=float)
    assert r.status == 0
    # FP32 velocity accumulation over 40 half-kick pairs
    assert abs(r.drift - a_dt * s1.mean()) <= 1e-5 * a_dt * s1.mean()
    e_exp = 0.5 * workloads.ME * a_dt ** 2 * np.mean(s1 ** 2) / workloads.EV
    assert abs(r.mean_energy_ev - e_exp) <= 1e-5 * e_exp
    assert r.collisions[0] == 0


def _initial_velocities(oracle, n, e0, seed):
    """The device's initial Maxwellian (Box-Muller on the Philox stream at
    step slot 0xFFFFFFFF), restated in FP64."""
    sigma = math.sqrt(2.0 * e0 * workloads.EV / (3.0 * workloads.ME))
    v = np.zeros((n, 3))
    for p in range(n):
        s = oracle.stream(seed, 0, p, 0xFFFFFFFF)
        u1, u2, u3, u4 = [oracle.stream_u01(s) for _ in range(4)]
        r1, r2 = math.sqrt(-2 * math.log(u1)), math.sqrt(-2 * math.log(u3))
        v[p] = [sigma * r1 * math.cos(2 * math.pi * u2), sigma * r1 * math.sin(2 * math.pi * u2),
                sigma * r2 * math.cos(2 * math.pi * u4)]
    return v


def test_swarm_matches_oracle_same_stream(ctx, oracle):
    gas = workloads.synthetic_gas(rows=200)
    model = oracle_py.model_from_gas(workloads.synthetic_gas(rows=200))
    ctx.swarm_gas(gas)
    dt = 0.099 / gas["max_frequency"]
    n, steps, burn, blocks, seed = 512, 1200, 400, 8, 4242
    r = _run_all(ctx, [_request(300.0, n, steps, burn, blocks, dt, seed)], 300)[0]
    vel0 = _initial_velocities(oracle, n, 1.0, seed)
    cfg = dict(neutral_density=gas["neutral_density"], max_frequency=gas["max_frequency"],
               timestep=dt, reduced_field_td=300.0, steps=steps, burn_in=burn, particles=n,
               blocks=blocks, initial_mean_energy_ev=1.0, max_energy_ev=900.0)
    o = oracle.swarm(model, cfg, vel0, seed, 0)
    assert r.status == 0 and o.status == 0
    for mine, se, theirs, tse in [(r.drift, r.drift_se, o.drift, o.drift_se),
                                  (r.mean_energy_ev, r.mean_energy_se, o.mean_energy_ev,
                                   o.mean_energy_se)]:
        assert abs(mine - theirs) <= 4 * math.hypot(se, tse) + 0.01 * abs(theirs)
    assert abs(r.ionization_rate - o.ion_rate) <= 5 * math.hypot(r.ionization_rate_se,
                                                                 o.ion_rate_se) + 1.0
    # the same displacement histories: diffusion within 5 % (a few FP32/FP64
    # branch flips change single trajectories)
    for mine, theirs in [(r.longitudinal_diffusion, o.longitudinal_diffusion),
                         (r.transverse_diffusion, o.transverse_diffusion)]:
        assert theirs > 0 and abs(mine - theirs) <= 0.05 * theirs, (mine, theirs)


def test_swarm_matches_reference_golden(ctx):
    g = golden("swarm_300td.npz")
    gas = workloads.synthetic_gas(rows=200)
    ctx.swarm_gas(gas)
    req = _request(float(g["reduced_field_td"]), int(g["particles"]), int(g["steps"]),
                   int(g["burn_in"]), int(g["blocks"]), float(g["timestep"]), 99,
                   e0=float(g["initial_mean_energy_ev"]), emax=float(g["max_energy_ev"]))
    r = _run_all(ctx, [req], 1000)[0]
    ref = g["out"]
    assert r.status == 0
    for mine, se, theirs, tse in [(r.drift, r.drift_se, ref[0], ref[1]),
                                  (r.mean_energy_ev, r.mean_energy_se, ref[2], ref[3])]:
        assert abs(mine - theirs) <= 5 * math.hypot(se, tse) + 0.02 * abs(theirs)
    assert abs(r.ionization_rate - ref[4]) <= 5 * math.hypot(r.ionization_rate_se, ref[5]) + 1.0


def test_swarm_alone_equals_mixed():
    """Continuous batching: a request's result is bit-identical alone and
    mixed with others of different sizes, advanced in different launch
    chunks (the statistics are integer sums)."""
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=200)
    dt = 0.099 / gas["max_frequency"]
    target = _request(100.0, 1000, 600, 200, 10, dt, 11)
    others = [_request(30.0, 3000, 900, 300, 10, dt, 12), _request(700.0, 257, 350, 50, 5, dt, 13)]
    with hip.Context(0) as a:
        a.swarm_gas(gas)
        alone = _run_all(a, [target], 600)[0]
    with hip.Context(0) as b:
        b.swarm_gas(gas)
        mixed = _run_all(b, [others[0], target, others[1]], 77)[1]
    for f in ["drift", "drift_se", "mean_energy_ev", "mean_energy_se", "ionization_rate",
              "candidates", "nulls", "max_energy_ev"]:
        assert getattr(alone, f) == getattr(mixed, f), f
    assert list(alone.collisions) == list(mixed.collisions)


def test_swarm_admission_budget(ctx):
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=200)
    ctx.swarm_gas(gas)
    dt = 0.099 / gas["max_frequency"]
    r = _request(10.0, 10000, 100, 10, 5, dt, 1)
    fp = ctx.swarm_footprint(r)
    h1 = ctx.swarm_submit(r, hbm_budget=fp + 10)
    with pytest.raises(hip.HipError, match="HBM budget"):
        ctx.swarm_submit(r, hbm_budget=fp + 10)
    while True:
        ctx.swarm_advance(100)
        done, res = ctx.swarm_poll(h1)
        if done:
            break
    ctx.swarm_submit(r, hbm_budget=fp + 10)    # memory freed: admitted now
    with pytest.raises(hip.HipError, match="max_frequency \\* timestep"):
        ctx.swarm_submit(_request(10.0, 100, 100, 10, 5, 1.0, 1))


def test_reference_swarm_example_through_readers(ctx):
    """The reference's examples/synthetic_swarm.swarm read by the gas and
    swarm readers and run on the device (300 K neutrals, Henyey-Greenstein
    elastic channel): drift, mean energy, elastic collisions and candidates
    per field within 4 sigma of the reference's own runs, sigma being the
    reference's spread over 16 seeds."""
    import os
    from aurora_pic import gasdata
    from test_gasdata import SWARM_DIR, check_against_reference
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    ctx.swarm_gas(gasdata.swarm_gas(ds, cfg))
    gasdata.apply_channel_options(ctx, ds, model=1)
    results = _run_all(ctx, gasdata.swarm_requests(cfg), 500)
    assert all(r.status == 0 for r in results)
    check_against_reference([(r.drift, r.mean_energy_ev, r.collisions[0], r.candidates,
                              r.longitudinal_diffusion, r.transverse_diffusion)
                             for r in results])


def test_reference_swarm_example_csv(ctx, tmp_path):
    """The device run of the reference's example written in the reference's
    CSV schema: every column that describes the run (dataset, config, seeds,
    collision model signature, field, fixed-population columns) equals the
    reference's own CSV; the statistics agree within 4 sigma of the
    reference's seed spread (checked by the previous test's helper)."""
    import csv
    import os
    from aurora_pic import gasdata, swarmio
    from test_gasdata import SWARM_DIR, _reference_csv, _strip_paths, check_against_reference
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    ctx.swarm_gas(gasdata.swarm_gas(ds, cfg))
    gasdata.apply_channel_options(ctx, ds, model=1)
    results = _run_all(ctx, gasdata.swarm_requests(cfg), 700)
    out = str(tmp_path / "swarm.csv")
    swarmio.write_swarm_benchmark_csv(out, cfg, ds, results)
    head, ref_rows = _reference_csv()
    with open(out) as f:
        mine = list(csv.reader(f))
    assert mine[0] == head
    statistical = {"mean_velocity_x_m_s", "mean_velocity_x_standard_error_m_s",
                   "electron_drift_velocity_m_s", "reduced_mobility_1_v_m_s", "mean_energy_ev",
                   "mean_energy_standard_error_ev", "longitudinal_diffusion_m2_s",
                   "transverse_diffusion_m2_s", "maximum_observed_energy_ev",
                   "collision_candidates", "null_collisions"}
    statistical |= {h for h in head if h.endswith(("_collisions", "_rate_per_electron_s",
                                                   "_rate_standard_error_s"))}
    for got, want in zip(mine[1:], ref_rows):
        g, w = _strip_paths(got, head), _strip_paths(want, head)
        for name, a, b in zip(head, g, w):
            if name not in statistical:
                assert a == b, (name, a, b)
    check_against_reference([(r.drift, r.mean_energy_ev, r.collisions[0], r.candidates,
                              r.longitudinal_diffusion, r.transverse_diffusion)
                             for r in results])
